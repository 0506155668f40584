/*
 * fastice_hip.h -- C ABI of libfasticehip.so, the MI355X (gfx950) implementation
 * of FastIce.jl's pseudo-transient (PT) Stokes hot path.
 *
 * Reference interface replaced (FastIce.jl v0.3.3; paths relative to the
 * FastIce.jl source tree).  The reference has no FFI: its boundary is Julia
 * multiple dispatch on the Models / LevelSets API, which a Julia package would
 * re-expose by `ccall`ing the functions below.
 *
 *   fi_model_create      <- IsothermalFullStokesModel(; arch, grid, boundary_conditions,
 *                           gravity, rheology, solver_params, outer_width)
 *                           src/Models/FullStokes/Isothermal/Isothermal.jl:60-89 and
 *                           IsothermalImmersedBoundaryFullStokesModel(...)
 *                           src/Models/ImmersedBoundaryFullStokes/Isothermal/Isothermal.jl:63-98
 *   fi_model_destroy     <- garbage collection of the model's device Fields
 *   fi_field_upload      <- set!(field, array) on model.stress/velocity/viscosity fields
 *                           (scripts_future_API/tm_stokes_chmy.jl:86-91)
 *   fi_field_download    <- interior(field) / Array(field) reads
 *                           (scripts_future_API/tm_stokes_chmy.jl:110-113)
 *   fi_field_fill        <- set!(field, value) (tm_stokes_chmy.jl:86-88)
 *   fi_apply_bc          <- bc!(arch, grid, batch) for stress / velocity / eta groups
 *                           (tm_stokes_chmy.jl:93-94; Isothermal.jl:102-104)
 *   fi_advance_iterations<- advance_iteration!(model, t, dt), n times
 *                           (FullStokes Isothermal.jl:91-112; IB Isothermal.jl:100-122)
 *   fi_compute_residuals <- compute_residuals!(model) + norm(r, Inf) per residual
 *                           (Isothermal.jl:114-125; tm_stokes_chmy.jl:99-106)
 *   fi_solve             <- the driver loop of tm_stokes_chmy.jl:96-108 with a tolerance exit
 *   fi_compute_tau       <- compute_tau! helper kernel
 *                           (src/Models/FullStokes/Isothermal/kernels_3d.jl:44-64, kernels_2d.jl:33-44;
 *                           IB: src/Models/ImmersedBoundaryFullStokes/Isothermal/kernels_3d.jl:78-116)
 *   fi_levelset_from_dem <- compute_levelset_from_dem!(arch, launch, psi, dem, grid2D, grid, R)
 *                           (src/LevelSets/compute_levelsets.jl:16-25)
 *   fi_invert_levelset   <- invert_levelset!(arch, launch, psi, grid) (src/LevelSets/compute_levelsets.jl:27-35)
 *   fi_omega_from_psi    <- omega_from_psi!(arch, launch, mask, psi, grid)
 *                           (src/LevelSets/volume_fractions_kernels.jl:101-110)
 *
 * Conventions
 *   - Every call returns FI_OK (0) or a negative fi_status; fi_last_error()
 *     returns a thread-local message for the last failure.
 *   - The library owns all device memory; host pointers are borrowed for the
 *     duration of a call.  One handle must not be used from two threads at once;
 *     distinct handles are independent.  Each handle owns one HIP stream.
 *   - Calls that return host data (downloads, norms, fi_solve) synchronise the
 *     handle's stream; the others only enqueue work.
 *   - Field arrays on the host are x-fastest (Julia column-major order).  Layout
 *     FI_LAYOUT_PARENT is the field with its 1-cell halo in a uniform
 *     (nx+3)*(ny+3)*(nz+3) box (2D: (nx+3)*(ny+3)) for every staggered
 *     location; element (i,j,k) of the 1-based Chmy index is at
 *     i + (nx+3)*j + (nx+3)*(ny+3)*k.  FI_LAYOUT_INTERIOR is the interior only,
 *     (mx,my,mz) with m = n for Center and n+1 for Vertex dimensions.
 */
#ifndef FASTICE_HIP_H
#define FASTICE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FI_ABI_VERSION 1

typedef enum fi_status {
    FI_OK = 0,
    FI_ERR_ARG = -1,
    FI_ERR_OOM = -2,
    FI_ERR_HIP = -3,
    FI_ERR_NONFINITE = -4, /* tm_stokes_chmy.jl:104-106 "simulation failed" */
    FI_ERR_UNSUPPORTED = -5,
    FI_ERR_STATE = -6
} fi_status;

/* BoundaryCondition{Traction|Velocity|Slip}: boundary_conditions.jl:1-13 */
enum { FI_BC_TRACTION = 0, FI_BC_VELOCITY = 1, FI_BC_SLIP = 2 };

/* rheologies: src/Physics.jl:7-47 */
enum { FI_RHEO_LINEAR_CONST = 0, FI_RHEO_LINEAR_FIELD = 1, FI_RHEO_POWERLAW = 2 };
enum { FI_RELAX_NONE = 0, FI_RELAX_LIN = 1, FI_RELAX_LOG = 2, FI_RELAX_LOG10 = 3 };

enum { FI_PREC_FP64 = 0, FI_PREC_FP32_REFINED = 1 };

enum { FI_LAYOUT_PARENT = 0, FI_LAYOUT_INTERIOR = 1 };

/* fi_apply_bc groups */
enum { FI_BCGROUP_STRESS = 0, FI_BCGROUP_VELOCITY = 1, FI_BCGROUP_ETA = 2, FI_BCGROUP_RESIDUAL = 3 };

/* field identifiers (2D models use the x/y members only) */
enum {
    FI_F_P = 0,
    FI_F_TXX, FI_F_TYY, FI_F_TZZ, FI_F_TXY, FI_F_TXZ, FI_F_TYZ,
    FI_F_VX, FI_F_VY, FI_F_VZ,
    FI_F_ETA, FI_F_ETA_NEXT,
    FI_F_RP, FI_F_RVX, FI_F_RVY, FI_F_RVZ,
    FI_F_GX, FI_F_GY, FI_F_GZ,  /* gravity fields when gravity_is_field */
    FI_F_RHEO,                  /* LinearViscousRheology eta field / PowerLawRheology A field */
    FI_F_PSI_NA, FI_F_PSI_NS,   /* level sets at vertices */
    FI_F_OM_NA0,                /* + component: ccc vvv vcc cvc ccv vvc vcv cvv (2D: cc vv vc cv) */
    FI_F_OM_NS0 = FI_F_OM_NA0 + 8,
    FI_F_COUNT = FI_F_OM_NS0 + 8
};

/* mask phases */
enum { FI_PHASE_NA = 0, FI_PHASE_NS = 1 };

typedef struct fi_rheology {
    int32_t kind;       /* FI_RHEO_* */
    int32_t has_reg;    /* wrap in ViscoplasticRegularisation(inner, eta_reg) */
    int32_t relax;      /* wrap in Relaxation(eta, inner, rate; scale) */
    int32_t A_is_field; /* POWERLAW: A read from FI_F_RHEO */
    double eta;         /* LINEAR_CONST viscosity */
    double A;           /* POWERLAW prefactor (constant) */
    double n;           /* POWERLAW exponent */
    double eta_reg;     /* regularisation viscosity */
    double rate;        /* relaxation rate */
} fi_rheology;

typedef struct fi_model_desc {
    int32_t ndims;              /* 2 or 3 */
    int32_t n[3];               /* cells per dimension */
    double origin[3];
    double extent[3];
    int32_t bc_kind[3][2];      /* [dim][low/high] FI_BC_* */
    double bc_value[3][2][3];   /* components of BoundaryCondition{Kind}(c1, c2, c3) */
    fi_rheology rheology;
    int32_t gravity_is_field;   /* 0: constant vector (ValueField/ZeroField), 1: FI_F_G* fields */
    double gravity[3];
    double dtau_P;              /* solver_params.dtau.Pr */
    double dtau_tau[6];         /* xx yy zz xy xz yz (2D: xx yy xy) */
    double dtau_V[3];
    int32_t immersed_boundary;  /* 1: IsothermalImmersedBoundaryFullStokesModel */
    int32_t precision;          /* FI_PREC_* */
} fi_model_desc;

typedef struct fi_model_s* fi_model_t;

int fi_abi_version(void);
const char* fi_last_error(void);

int fi_model_create(const fi_model_desc* desc, fi_model_t* out);
int fi_model_destroy(fi_model_t m);

int fi_field_upload(fi_model_t m, int field_id, const double* host, size_t count, int layout);
int fi_field_download(fi_model_t m, int field_id, double* host, size_t count, int layout);
int fi_field_fill(fi_model_t m, int field_id, double value);
/* fi_field_download with FI_LAYOUT_INTERIOR that allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  The interior
 * is packed on the device, a run of whole x rows at a time, into a static
 * staging area of 16 MiB (2^21 doubles) that each GPU holds until the process
 * ends, and copied out chunk by chunk; a field of any size is transferred,
 * and one whose x rows are longer than the staging area is rejected with
 * FI_ERR_ARG.  Every call on a GPU shares that GPU's staging area: calls that
 * overlap on one GPU (from several threads) race.  Synchronises the model's
 * stream; same count rule as fi_field_download. */
int fi_field_download_noalloc(fi_model_t m, int field_id, double* host, size_t count);
/* number of doubles of a field in the given layout (0 on a bad id) */
size_t fi_field_size(fi_model_t m, int field_id, int layout);

int fi_apply_bc(fi_model_t m, int group);
int fi_advance_iterations(fi_model_t m, int n_iter);
int fi_compute_residuals(fi_model_t m, double norms_inf[4]);
int fi_compute_tau(fi_model_t m);
int fi_solve(fi_model_t m, int max_iter, int ncheck, double tol, int* iters_out, double norms_inf[4]);
int fi_synchronize(fi_model_t m);
/* Measurement hook (no counterpart in the reference, whose timing is host
 * wall-clock around the loop, scripts_future_API/tm_stokes_mpi_wip.jl:151-183):
 * runs n_iter iterations with HIP events around every launch on the model's
 * stream, synchronising after each iteration.  out_ms = {iteration, fused
 * interior sweep, pointwise (shell) sweeps, boundary-condition kernels}
 * averages in ms. */
int fi_time_iterations(fi_model_t m, int n_iter, double out_ms[4]);
/* The interior box [lo, hi] (1-based Chmy indices) swept by the fused kernel;
 * returns 1 when the fused sweep is active, 0 otherwise. */
int fi_fused_region(fi_model_t m, int lo[3], int hi[3]);
/* Measurement hook (no counterpart in the reference): out[0] = cells of the
 * fused interior, out[1] = those the immersed-boundary sweep processes, i.e.
 * on planes of their column that hold a valid location (the sweep reads and
 * writes nothing on the other planes but the viscosity it updates; out[1] =
 * out[0] without an immersed boundary, both 0 without a fused interior). */
int fi_fused_active_cells(fi_model_t m, long long out[2]);
/* Validity bits of an immersed-boundary model, one byte per parent index
 * ((nx+3)(ny+3)(nz+3) bytes, x fastest; 0 outside the work range [1, n+1]).
 * A set bit marks a location the masked kernels update: bits 0-3 the stress
 * locations ccc vvc vcv cvv, valid when isvalid(na) && !isempty(ns), bits 4-6
 * the velocity locations vcc cvc ccv, valid when isvalid(ns) && !isempty(na)
 * (src/Models/ImmersedBoundaryFullStokes/Isothermal/null_spaces.jl:1-43 as
 * kernels_3d.jl:6,26,32,38,50,56,62 of that directory apply them; 2D uses
 * bits 0, 1, 4, 5).  Inspection hook with no counterpart in the reference,
 * where the predicates are evaluated inside every kernel. */
int fi_ib_validity(fi_model_t m, uint8_t* host, size_t count);

/* Options (no counterpart in the reference).  Kernel path, results do not
 * depend on them: FI_OPT_FUSED 1/0 selects the fused interior sweep or the
 * pointwise sweep everywhere; FI_OPT_FUSED_SHELL 1/0 runs the stress update of
 * the boundary shell inside the fused launch (default 1) or in a launch of its
 * own; FI_OPT_ZCHUNK sets the z planes per workgroup of
 * the fused sweep; FI_OPT_FUSED_VARIANT(_FP32) selects a compiled variant of
 * the FP64 (FP32) fused sweep; FI_OPT_TILE_ORDER maps workgroups to tiles:
 * 0 x-fastest, 1 one contiguous run per XCD, 2 runs of 8 tiles per XCD
 * (default 2).  Refined solver: FI_OPT_REFINE_FULL_PHASE 1/0 starts a
 * nonlinear solve with full FP32 PT sweeps (default 1) that stop after
 * FI_OPT_REFINE_PHASE1_DIGITS digits of residual drop (1..7, default 3) or
 * at the FP32 floor; FI_OPT_REFINE_ETA_RATE is the initial log-space
 * relaxation of the outer viscosity update in per mille (default 1000),
 * halved whenever an outer step reduces the residual by less than 10 %;
 * FI_OPT_REFINE_ANDERSON is the depth (0..3, default 2) of the Anderson
 * mixing of the nonlinear outer updates (0: plain defect correction).
 * FI_OPT_REFINE_MODE selects the outer loop: 1 (default) lagged PT, where
 * every outer step follows exactly ncheck FP32 sweeps, the stress correction
 * starts from the lag of the PT stress behind 2 eta edot'(V) and the
 * viscosity is relaxed as the PT iteration would have over those sweeps;
 * the chunk halves (down to 8 sweeps) whenever the outer residual has not
 * halved over 8 n_max sweeps; max_outer x ncheck caps the FP32 sweeps in all
 * (max_inner caps the full-FP32 nonlinear start; inner_rtol, the
 * Anderson depth and the rate halving are not used); 0 correction solves:
 * each outer step solves the correction problem to inner_rtol with the
 * stress lag dropped, the viscosity following a relaxed Picard iteration.
 * FI_OPT_REFINE_COLS 1/0: on a 3D immersed-boundary model with traction on
 * every face, the lagged-PT outer steps after the first visit only the
 * planes of each column that hold (or neighbour) a valid location (default
 * 1), or every cell (0); the results are the same bit for bit. */
enum {
    FI_OPT_FUSED = 1,
    FI_OPT_ZCHUNK = 2,
    FI_OPT_REFINE_FULL_PHASE = 3,
    FI_OPT_REFINE_ETA_RATE = 4,
    FI_OPT_REFINE_PHASE1_DIGITS = 5,
    FI_OPT_FUSED_VARIANT = 6,
    FI_OPT_FUSED_VARIANT_FP32 = 7,
    FI_OPT_TILE_ORDER = 8,
    FI_OPT_REFINE_ANDERSON = 9,
    FI_OPT_REFINE_MODE = 10,
    FI_OPT_FUSED_SHELL = 11,
    FI_OPT_REFINE_COLS = 12
};
int fi_set_option(fi_model_t m, int key, int value);
/* Compiled variants of the fused sweep (FI_OPT_FUSED_VARIANT* take an index
 * into this list): triples (tile lanes in x: 64 for 62 columns at any
 * alignment or 66 for 64 line-aligned columns, tile rows, minimum waves per
 * SIMD the register allocation allows, 1 = unconstrained); returns the
 * number of variants and writes at most max_variants triples.  Results do
 * not depend on the variant. */
int fi_fused_variants(int fp32, int* tx_ty_minw, int max_variants);

int fi_levelset_from_dem(fi_model_t m, int phase, const double* dem_parent, int dem_nx, int dem_ny,
                         const double dem_origin[2], const double dem_extent[2], const double R[9]);
/* 2D level set of a closed polygon (npts points, x y pairs): signed distance
 * of every vertex, negative inside (even-odd rule), then Neumann on psi.
 * Replaces the 2D distance sd_poly (src/LevelSets/signed_distances.jl:3-19),
 * the 2D counterpart of compute_levelset_from_dem! (compute_levelsets.jl:16-25);
 * the reference defines it but no 2D driver calls it. */
int fi_levelset_from_polygon(fi_model_t m, int phase, const double* xy, int npts);
int fi_invert_levelset(fi_model_t m, int phase);
int fi_omega_from_psi(fi_model_t m, int phase);

/* FP32 sweeps inside an FP64 defect-correction loop (no counterpart in the
 * reference; same solution as fi_solve).  Outer steps run on the GPU in FP64:
 * eta <- rheology(tau_II) (nonlinear rheologies), tau <- 2 eta edot'(V),
 * residuals and Inf-norms; stop when all norms < tol or after max_outer steps.
 * Between them the correction solves the PT iteration in FP32 with frozen
 * eta, homogeneous boundary values and the FP64 defect as forcing, checked
 * every ncheck iterations, until its residual falls below inner_rtol times the
 * outer one (at most max_inner iterations), and is added to the FP64 state.
 * Nonlinear rheologies first run full FP32 PT sweeps (eta updated every
 * sweep) until the FP32 residual drops by FI_OPT_REFINE_PHASE1_DIGITS digits
 * or stalls, and relax the outer viscosity update in log space
 * (FI_OPT_REFINE_ETA_RATE).
 * counts_out = {outer steps, total FP32 iterations, FP32 iterations of the
 * nonlinear phase}. */
int fi_refined_solve(fi_model_t m, int max_outer, int max_inner, int ncheck, double inner_rtol, double tol,
                     int counts_out[3], double norms_inf[4]);
/* n_iter FP32 correction iterations on the current correction state (timing) */
int fi_refined_advance_inner(fi_model_t m, int n_iter);
/* as fi_time_iterations, for the FP32 correction iteration */
int fi_refined_time_inner(fi_model_t m, int n_iter, double out_ms[4]);

/* free / total device memory in bytes (hipMemGetInfo) */
int fi_mem_info(size_t* free_bytes, size_t* total_bytes);
/* device bytes fi_model_create would allocate for desc (0 on a bad desc);
 * admission control of the serving loop */
size_t fi_model_bytes(const fi_model_desc* desc);

/* Batched solves (no counterpart in the reference, whose models are solved
 * one at a time by the driver loop tm_stokes_chmy.jl:96-108).  A batch holds
 * FP64 models; every PT iteration of all running models is a few packed
 * launches (fused interior tiles of every instance, their pointwise shells,
 * boundary conditions per dimension) replayed as a HIP graph.  Each model
 * keeps fi_solve's semantics exactly: its own iteration count, residual
 * check every ncheck of its iterations, stop at the first check with all
 * norms < tol or at max_iter; results are bitwise those of fi_solve on the
 * model alone.  While attached, the model's work runs on the batch's stream.
 * The batch does not own its models: fi_model_destroy detaches one. */
typedef struct fi_batch_s* fi_batch_t;
enum { FI_BATCH_RUNNING = 0, FI_BATCH_CONVERGED = 1, FI_BATCH_MAX_ITER = 2, FI_BATCH_FAILED = 3 };
int fi_batch_create(fi_batch_t* out);
int fi_batch_destroy(fi_batch_t b);
/* attach a model and start its solve (state FI_BATCH_RUNNING) */
int fi_batch_add(fi_batch_t b, fi_model_t m, int max_iter, int ncheck, double tol);
/* attach a model and start its FP32-refined solve (the arguments of
 * fi_refined_solve).  A batch holds either FP64 solves (fi_batch_add) or
 * refined solves: the FP32 sweeps of every running refined model are packed
 * into the same launches up to the nearest model's residual check, then each
 * model whose chunk ended runs its check and FP64 outer step on the batch's
 * stream.  Results are bitwise those of fi_refined_solve on the model alone;
 * fi_batch_status reports FP32 sweeps as iterations. */
int fi_batch_add_refined(fi_batch_t b, fi_model_t m, int max_outer, int max_inner, int ncheck, double inner_rtol,
                         double tol);
/* counts_out = {outer steps, FP32 sweeps, FP32 sweeps of the nonlinear phase} */
int fi_batch_refined_counts(fi_batch_t b, fi_model_t m, int counts_out[3]);
int fi_batch_remove(fi_batch_t b, fi_model_t m);
/* advance every running model to the next residual check of the nearest one,
 * check the models due, and return how many finished in *n_finished */
int fi_batch_step(fi_batch_t b, int* n_finished);
/* step until no model is running */
int fi_batch_solve(fi_batch_t b);
/* state (FI_BATCH_*), iterations done and last residual norms of a model;
 * a failed model (non-finite residual) reports FI_BATCH_FAILED */
int fi_batch_status(fi_batch_t b, fi_model_t m, int* state, int* iters, double norms_inf[4]);
/* FI_BOPT_VARIANT: fused sweep variant (index into fi_fused_variants(0), or
 * fi_fused_variants(1) for a refined batch),
 * FI_BOPT_ZCHUNK: z planes per tile, FI_BOPT_AUTOTUNE 1/0: time every FI_BOPT_GRAPH_ITERS: iterations per graph
 * replay (0: no graph), FI_BOPT_TILE_ORDER: as FI_OPT_TILE_ORDER.  Results do
 * not depend on any of them. */
enum { FI_BOPT_VARIANT = 1, FI_BOPT_ZCHUNK = 2, FI_BOPT_AUTOTUNE = 3, FI_BOPT_GRAPH_ITERS = 4, FI_BOPT_TILE_ORDER = 5 };
int fi_batch_set_option(fi_batch_t b, int key, int value);
/* out = {attached, running, fused work items, (32 * tile lanes + tile rows)
 * * 16 + min waves, 4 * z planes + tile order, graph iterations, autotune
 * runs, table rebuilds} */
int fi_batch_info(fi_batch_t b, int out[8]);

/* ---------------------------------------------------------------------------
 * PT heat diffusion: HeatDiffusionModel (src/Models/Diffusion/Heat/Heat.jl:16-40),
 * advance_iteration! (:42-58) with update_q! / update_T! (kernels.jl:3-29)
 * and the Value / Flux boundary conditions (boundary_conditions.jl:1-69).
 * The reference model does not run as shipped (it reads a missing
 * model.fields and calls KernelLaunch with the wrong arguments); this is the
 * model its kernels and boundary conditions define.  Fields: T and T_o at
 * cell centres, q = (qx, qy[, qz]) at the faces of its own dimension; same
 * layouts as the Stokes fields.  One PT iteration:
 *   q   -= (q + lambda_rho_cp * grad T) * dtau_q,  then q = v on Flux faces
 *   T   -= ((T - T_o) / dt + div q) * dtau_T,       then T ghost = 2 v - T on Value sides
 * --------------------------------------------------------------------------- */
typedef struct fi_heat_s* fi_heat_t;
enum { FI_HEAT_BC_NONE = 0, FI_HEAT_BC_VALUE = 1, FI_HEAT_BC_FLUX = 2 };
enum { FI_HF_T = 0, FI_HF_T_OLD = 1, FI_HF_QX = 2, FI_HF_QY = 3, FI_HF_QZ = 4, FI_HF_COUNT = 5 };
typedef struct {
    int ndims;               /* 2 or 3 */
    int n[3];                /* cells */
    double origin[3];
    double extent[3];
    int bc_kind[3][2];       /* FI_HEAT_BC_* per dimension and side (west/east, south/north, bottom/top) */
    double bc_value[3][2];
    double lambda_rho_cp;    /* physics.λ_ρCp */
    double dtau_q, dtau_T;   /* iter_params.Δτ.q, Δτ.T */
} fi_heat_desc;
/* HeatDiffusionModel(; arch, grid, boundary_conditions, physics, iter_params) */
int fi_heat_create(const fi_heat_desc* desc, fi_heat_t* out);
int fi_heat_destroy(fi_heat_t h);
/* set!(field, values) / interior(field): layouts as fi_field_upload */
int fi_heat_field_upload(fi_heat_t h, int id, const double* host, size_t count, int layout);
int fi_heat_field_download(fi_heat_t h, int id, double* host, size_t count, int layout);
/* the Value boundary conditions on T (after set!) */
int fi_heat_apply_bc(fi_heat_t h);
/* advance_iteration!(model, t, Δt), n_iter times */
int fi_heat_advance_iterations(fi_heat_t h, double dt, int n_iter);
/* Inf-norms of (T - T_o)/dt + div q (cells) and q + lambda grad T (faces not
 * fixed by a Flux condition) */
int fi_heat_residual_norms(fi_heat_t h, double dt, double norms_inf[2]);
/* advance_timestep! (exported by Heat.jl, not defined there): PT iterations
 * with a residual check every ncheck until both norms are below tol or
 * max_iter, then T_o <- T.  FI_ERR_NONFINITE on a non-finite residual. */
int fi_heat_step(fi_heat_t h, double dt, int max_iter, int ncheck, double tol, int* iters_out, double norms_inf[2]);
/* measurement hook: n_iter iterations back to back with HIP events around
 * each kernel; out_ms = {iteration, fused kernel} averages */
int fi_heat_time_iterations(fi_heat_t h, double dt, int n_iter, double out_ms[2]);

#ifdef __cplusplus
}
#endif
#endif
