/*
 * fastice_oracle.h -- CPU FP64 restatement of FastIce.jl's PT-Stokes path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library is compared against; it is never linked into, loaded by, or
 * called from the product path (libfasticehip.so and the Python host
 * package).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Storage convention (shared with the HIP library's "PARENT" layout):
 * every field, whatever its staggered location, is a dense x-fastest
 * array of (nx+3)*(ny+3)*(nz+3) doubles (2D: (nx+3)*(ny+3)).  The storage
 * index of the 1-based Chmy index (i,j,k) is i + sy*j + sz*k with
 * sy = nx+3, sz = (nx+3)*(ny+3); index 0 is the low halo, 1..m the
 * interior (m = n for Center, n+1 for Vertex), m+1 the high halo.
 *
 * Parity status of each function (tests/test_oracle_pins.py; the reference
 * ships no golden vectors for this path and its Julia/Chmy code cannot run
 * here, so every pin is a closed form, an invariant or brute force):
 *   or_init_grid            grid KATs of test/test_chmy_grids.jl:113-122
 *   or_update_sigma/_V,     exact divergence/gradient on linear fields,
 *   or_advance, or_solve    fixed-point invariance, linearity in the forcing,
 *                           O(h^2) convergence to the 2D and 3D manufactured
 *                           solutions (127^2 unit-box case included)
 *   or_compute_tau          closed form on a linear velocity field
 *   or_bc_*                 free-slip / no-slip / traction invariants after
 *                           every iteration
 *   or_compute_residuals,   residual of the manufactured solutions; NaN
 *   or_residual_norms       propagation of the Inf-norm
 *   or_rheology             scalar closed forms of Physics.jl:7-47
 *   or_psi2omega3/2,        exact fractions of planar level sets, complement
 *   or_omega_from_psi       rule, O(h^2) sphere volume
 *   or_levelset_from_dem,   flat and inclined DEMs exactly, brute-force
 *   or_invert_levelset      distance over every DEM triangle
 *   or_sd_poly,             distances to axis-aligned squares in closed
 *   or_levelset_from_polygon form, brute-force point-segment distances and a
 *                           winding-number inside test on random polygons
 *   or_heat_advance,        steady linear profile between two values, heat
 *   or_heat_bc,             conservation under zero flux, backward-Euler
 *   or_heat_residual_norms  decay of a discrete cosine mode
 *   immersed-boundary       (tests/test_oracle_ib_pins.py) or_ib_flags and the
 *   branches (ib != 0) of   masked update_sigma / update_V / compute_tau /
 *   the kernels above,      compute_residuals, 3D and 2D, equal bit for bit an
 *   or_ib_flags             independent numpy transcription that derives every
 *                           neighbour from null_spaces.jl's generic
 *                           flip/left/right rule, on random masks around the
 *                           1e-6 threshold; all-one masks reduce them bit for
 *                           bit to the plain kernels; a hydrostatic column on a
 *                           cut bed is an exact rest state
 * Parity unpinned (choices this library specifies where the reference
 * delegates to unvendored Chmy.jl and no check can reach the reference's own
 * bits): the ghost values at edges and corners that the x-then-y-then-z
 * order of the boundary conditions produces, the definition of Chmy's masked
 * derivative as (w_r f_r - w_l f_l) / dx (the transcription shares it; the
 * indexing built on it is pinned), zero halos of the volume-fraction masks,
 * the Inf-norm over interiors only, and the bitwise trajectory of the
 * KernelAbstractions CPU path (difference times inverse spacing, operation
 * order).
 */
#ifndef FASTICE_ORACLE_H
#define FASTICE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BC_TRACTION = 0, OR_BC_VELOCITY = 1, OR_BC_SLIP = 2 };
enum { OR_RHEO_LINEAR_CONST = 0, OR_RHEO_LINEAR_FIELD = 1, OR_RHEO_POWERLAW = 2 };
enum { OR_RELAX_NONE = 0, OR_RELAX_LIN = 1, OR_RELAX_LOG = 2, OR_RELAX_LOG10 = 3 };

/* Mask component order (3D FieldMask): ccc vvv vcc cvc ccv vvc vcv cvv.
 * 2D: cc vv vc cv. */
enum { OM_CCC = 0, OM_VVV, OM_VCC, OM_CVC, OM_CCV, OM_VVC, OM_VCV, OM_CVV };
enum { OM2_CC = 0, OM2_VV, OM2_VC, OM2_CV };

typedef struct {
    int ndims;
    int n[3];
    double origin[3];
    double extent[3];
    double dx[3];   /* spacing  = extent/n  */
    double idx[3];  /* inverse spacing = 1/dx */
    long sy, sz;
    int bc_kind[3][2];
    double bc_val[3][2][3];
    double dPr;
    double dtau[6];  /* xx yy zz xy xz yz  (2D: xx yy xy) */
    double dV[3];
    /* rheology: inner law, optional viscoplastic cap, optional relaxation */
    int rh_kind;
    double rh_eta;       /* LINEAR_CONST viscosity */
    double rh_A;         /* POWERLAW constant A (used when rh_field is NULL) */
    double rh_n;         /* POWERLAW exponent */
    int rh_has_reg;
    double rh_inv_eta_reg;
    int rh_relax;
    double rh_rate;
    int grav_field;      /* 0: constant vector grav, 1: fields gx,gy,gz */
    double grav[3];
} or_model;

typedef struct {
    double *P, *txx, *tyy, *tzz, *txy, *txz, *tyz;
    double *Vx, *Vy, *Vz;
    double *eta, *eta_next;
    double *rP, *rVx, *rVy, *rVz;
    double *gx, *gy, *gz;
    double *rh_field;         /* LINEAR_FIELD eta, or POWERLAW A field; NULL = constant */
    double *na[8], *ns[8];    /* immersed-boundary masks (NULL when unused) */
} or_fields;

/* grid set-up: fills dx, idx, sy, sz from ndims, n, origin, extent */
void or_init_grid(or_model* m);

/* one PT iteration (update_sigma, stress BCs, update_V, velocity + eta BCs,
 * eta swap), repeated niter times; ib != 0 selects the masked kernels */
void or_advance(const or_model* m, or_fields* f, int niter, int ib);
void or_update_sigma(const or_model* m, or_fields* f, int ib);
void or_update_V(const or_model* m, or_fields* f, int ib);
void or_bc_stress(const or_model* m, or_fields* f);
void or_bc_velocity(const or_model* m, or_fields* f);
void or_bc_neumann_center(const or_model* m, double* a);
void or_bc_neumann_vertex(const or_model* m, double* a);
void or_compute_residuals(const or_model* m, or_fields* f, int ib);
void or_compute_tau(const or_model* m, or_fields* f, int ib);
/* validity bits of every location in the work range (parent layout, one
 * byte per index; bits 0-3 stress ccc vvc vcv cvv, 4-6 velocity vcc cvc ccv) */
void or_ib_flags(const or_model* m, const or_fields* f, unsigned char* out);
/* Inf-norms over the interior of rP, rVx, rVy(, rVz); NaN propagates */
void or_residual_norms(const or_model* m, const or_fields* f, double out[4]);
/* fixed-cadence solve: every ncheck iterations compute residuals and stop
 * when all Inf-norms < tol; returns iterations done, -1 on non-finite */
int or_solve(const or_model* m, or_fields* f, int max_iter, int ncheck, double tol,
             int ib, double norms[4]);

/* rheology closure at one cell (exposed for closed-form tests) */
double or_rheology(const or_model* m, const or_fields* f, long c, double tII);

/* level sets and volume fractions */
void or_omega_from_psi(const or_model* m, const double* psi, double* om[8]);
double or_psi2omega3(const double psi8[8], const double d[3]);
double or_psi2omega2(const double psi4[4], const double d[2]);
void or_levelset_from_dem(const or_model* m, double* psi, const double* dem,
                          const int dem_n[2], const double dem_origin[2],
                          const double dem_extent[2], const double R[9]);
void or_invert_levelset(const or_model* m, double* psi);
double or_sd_poly(double px, double py, const double* xy, int npts);
void or_levelset_from_polygon(const or_model* m, double* psi, const double* xy, int npts);

/* PT heat diffusion (src/Models/Diffusion/Heat: Heat.jl, kernels.jl), run
 * as that model's advance_iteration! intends.  T, To: Center; q[d]: Vertex
 * in dimension d.
 * Same dense parent layout and strides as or_model. */
enum { OR_HEAT_NONE = 0, OR_HEAT_VALUE = 1, OR_HEAT_FLUX = 2 };
typedef struct {
    int ndims;
    int n[3];
    double dx[3];
    long sy, sz;
    int bc_kind[3][2];
    double bc_val[3][2];
    double lam;        /* physics.λ_ρCp */
    double dtq, dtT;   /* iter_params.Δτ.q, Δτ.T */
} or_heat;
typedef struct {
    double *T, *To, *q[3];
} or_heat_fields;
/* update_q! + flux BCs, update_T! + value BCs (Heat.jl:42-58), niter times */
void or_heat_advance(const or_heat* h, or_heat_fields* f, double dt, int niter);
/* the value BCs on T alone (the bc! a caller applies after set!) */
void or_heat_bc(const or_heat* h, or_heat_fields* f);
/* Inf-norms of (T - To)/dt + div q over cells and of q + lam grad T over
 * the faces a flux BC does not fix; NaN propagates */
void or_heat_residual_norms(const or_heat* h, const or_heat_fields* f, double dt, double out[2]);

#ifdef __cplusplus
}
#endif
#endif
