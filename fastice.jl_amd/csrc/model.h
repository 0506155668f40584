// model.h -- the model handle behind fi_model_t and the internal helpers
// shared by the C ABI (api.hip) and the FP32 defect-correction solver
// (refine.hip).
#pragma once

#include "fi_internal.h"

#define HIPCHK(expr)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fi::fail(FI_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace fi {
struct Refine;
int fail(int code, const char* fmt, ...);
} // namespace fi

using fi::BoxList;
using fi::FusedGeom;
using fi::Geo;

struct fi_model_s {
    fi_model_desc desc;
    Geo g;
    int device = 0; // the HIP device current when the model was created
    hipStream_t stream = nullptr;
    double* base[FI_F_COUNT] = {};
    double* f[FI_F_COUNT] = {};
    // second buffer of the stress (P..TYZ) and velocity (VX..VZ) state, in the
    // allocation of the first (base[id] + pair_stride + off)
    double* sf[10] = {};
    int sig_cur = 0; // 1: current stress lives in sf
    int v_cur = 0;   // 1: current velocity lives in sf
    int eta_cur = 0; // 0: logical eta lives in buffer FI_F_ETA, 1: in FI_F_ETA_NEXT
    bool shadow_dirty = true;
    uint8_t* flags_base = nullptr;
    uint8_t* flags = nullptr;
    bool flags_dirty = true;
    // 3D immersed boundary: per-column plane ranges of the flags (launch_col_ranges)
    int2* crange = nullptr;
    int2* erange = nullptr;
    // faces whose layer holds no valid location (face_layers_empty), with the ranges
    unsigned empty_faces = 0;
    int2* tmarch = nullptr; // the z range each tile of the last FP64 fused sweep marched
    // zs[b] / zv[b]: physical buffer b (0: f, 1: sf) of the stress / velocity
    // is known to hold +0 at every invalid location of the fused footprint
    // (a sweep's output is; an upload, a state change outside the sweeps or
    // a mask change makes it unknown)
    bool zs[2] = {false, false};
    bool zv[2] = {false, false};
    // the same for the two FP32 state buffers of the refined solver
    bool zs32[2] = {false, false};
    bool zv32[2] = {false, false};
    // gs[b]: the stress boundary conditions last applied to physical stress
    // buffer b left every ghost whose point and chain start are invalid
    // locations at the value its rules give a +0 start, and nothing has
    // written there since, so the next application may skip those ghosts
    // (set after an application that follows a sweep; anything that clears
    // zs clears it); gs32: the same for the FP32 stress buffers
    bool gs[2] = {false, false};
    bool gs32[2] = {false, false};
    int mask_gen = 0; // bumped whenever a mask component changes
    // 3D immersed boundary: the residual fields hold 0 at every position
    // outside their column's plane range (set by a residual computation over
    // the whole grid; cleared by a mask change or an upload into them)
    bool resid_clean = false;
    int use_fused = 1;
    int refine_full_phase = 1;      // nonlinear refined solves start with a full FP32 PT phase
    double refine_eta_rate = 1.0;   // log-space relaxation of the outer viscosity update (initial)
    int refine_phase1_digits = 3;   // the FP32 nonlinear phase stops after this many digits of residual drop
    int refine_anderson = 2;        // Anderson depth of the nonlinear outer loop (0: plain defect correction)
    int refine_mode = 1;            // outer loop: 1 lagged PT, 0 correction solves
    int fused_shell = 1;            // the shell's stress update rides in the fused launch
    int refine_cols = 1;            // lagged-PT outer steps visit the column ranges only (where exact)
    int kch = 0;                             // z planes per fused workgroup (0: sized to the grid)
    int ftx = 64, fty = 8, fminw = 1;        // fused sweep variant of the FP64 iteration (lanes, rows, min waves)
    int ftx32 = 64, fty32 = 8, fminw32 = 1;  // and of the FP32 sweeps of the refined solver
    int forder = 2;                          // workgroup-to-tile order of the fused sweeps
    FusedGeom fg;
    BoxList shell;
    bool has_interior = false;
    fi::Refine* rf = nullptr; // FP32 defect-correction workspace (refine.hip)
    unsigned long long* d_norm = nullptr;
    unsigned long long* h_norm = nullptr;
    double* d_dem = nullptr;
    size_t dem_elems = 0;
    double* d_stage = nullptr; // contiguous interior of one field, for host transfers
    size_t stage_elems = 0;
    fi_batch_s* batch = nullptr; // the batch this model is attached to (batch.hip)

    // the state changed outside the sweeps: both buffers must be made equal
    // again before the next FP64 sweep, and neither is known to be clean
    void state_written()
    {
        shadow_dirty = true;
        zs[0] = zs[1] = zv[0] = zv[1] = false;
        gs[0] = gs[1] = gs32[0] = gs32[1] = false;
    }
    static bool is_sig(int id) { return id >= FI_F_P && id <= FI_F_TYZ; }
    static bool is_vel(int id) { return id >= FI_F_VX && id <= FI_F_VZ; }
    double* eta() { return f[eta_cur ? FI_F_ETA_NEXT : FI_F_ETA]; }
    double* eta_next() { return f[eta_cur ? FI_F_ETA : FI_F_ETA_NEXT]; }
    // physical buffer holding the current value of a logical field id
    double* buf(int id)
    {
        if (id == FI_F_ETA) return eta();
        if (id == FI_F_ETA_NEXT) return eta_next();
        if (is_sig(id)) return sig_cur ? sf[id] : f[id];
        if (is_vel(id)) return v_cur ? sf[id] : f[id];
        return f[id];
    }
    // buffer the next iteration writes
    double* alt(int id)
    {
        if (is_sig(id)) return sig_cur ? f[id] : sf[id];
        if (is_vel(id)) return v_cur ? f[id] : sf[id];
        return f[id];
    }
};

namespace fi {

StokesArgs<double> make_args(fi_model_s* m, bool in_place);
BcBatch stress_batch(fi_model_s* m, int d);
BcBatch velocity_batch(fi_model_s* m, int d, bool with_eta);
// skip_ranges: the column plane ranges of a stress group whose settled ghosts may be skipped (gs)
// skip (3D immersed boundary, stress group): the settled ghosts of the
// written buffer are skipped (launch_bc_dims with the model's column ranges
// and empty faces)
void run_bc(fi_model_s* m, int group, bool skip = false);
// before kernels run: the immersed-boundary validity bits and, with
// sync_shadow, the ping-pong copies the FP64 iteration needs
int prepare(fi_model_s* m, bool sync_shadow = true);
void iterate(fi_model_s* m, hipEvent_t* ev);
int advance(fi_model_s* m, int niter);
// residuals and their Inf-norms of the current state; sync_shadow = false
// leaves the ping-pong buffers of the FP64 iteration stale (refined solves,
// which do not iterate in FP64, defer that copy to the next prepare)
int residual_norms(fi_model_s* m, double out[4], bool sync_shadow = true);
int residual_norms_enqueue(fi_model_s* m, bool sync_shadow = true);
void residual_norms_read(fi_model_s* m, double out[4]);
void batch_touch(fi_model_s* m);  // the model's buffer parity changed outside its batch
void batch_detach(fi_model_s* m); // remove the model from its batch (model destruction)
double bits_to_double(unsigned long long b);
// padded layout of a grid (rows of 32 elements, index 1 on a 128-byte line)
Geo geo_for(int nd, const int n[3], const double extent[3]);
// host <-> device copy of one field at location L in FI_LAYOUT_PARENT or
// FI_LAYOUT_INTERIOR (the interior through a contiguous staging buffer,
// grown on demand); synchronises the stream
int field_transfer(const Geo& g, double* dev, int L, double* host, size_t count, int layout, bool up, hipStream_t stream,
                   double** stage, size_t* stage_elems);
BcBatch neumann_batch(double* p, int L, int d);
void free_refine(fi_model_s* m);

inline void add_op(BcBatch& b, void* p, int L, int kind, double v, int f32 = 0)
{
    BcOp& o = b.op[b.nop++];
    o.ptr = p;
    o.L = L;
    o.kind = kind;
    o.f32 = f32;
    o.v = v;
}

// Both sides of one dimension share a launch; an op that belongs to one side
// only carries the side in its kind: 0/1 = both sides, 2/3 = low, 4/5 = high.
inline void add_side_op(BcBatch& b, void* p, int L, int kind, double v, int side, int f32 = 0)
{
    add_op(b, p, L, kind + 2 * (side + 1), v, f32);
}

// stress batch (boundary_conditions.jl:26-100) for the fields get(id); homog
// sets every prescribed value to zero (the batch of a correction)
template <class G>
BcBatch stress_batch_g(const fi_model_desc& ds, int d, G get, bool homog, int f32)
{
    BcBatch b;
    b.nop = 0;
    b.d = d;
    for (int s = 0; s < 2; ++s) {
        const int kind = ds.bc_kind[d][s];
        double c[3];
        for (int q = 0; q < 3; ++q) c[q] = homog ? 0.0 : ds.bc_value[d][s][q];
        if (ds.ndims == 3) {
            if (kind == FI_BC_TRACTION) {
                add_side_op(b, get(FI_F_P), L_CCC, 0, c[d], s, f32);
                add_side_op(b, get(d == 0 ? FI_F_TXX : d == 1 ? FI_F_TYY : FI_F_TZZ), L_CCC, 0, c[d], s, f32);
            }
            if (kind == FI_BC_TRACTION || kind == FI_BC_SLIP) {
                if (d == 0) {
                    add_side_op(b, get(FI_F_TXY), L_VVC, 0, c[1], s, f32);
                    add_side_op(b, get(FI_F_TXZ), L_VCV, 0, c[2], s, f32);
                } else if (d == 1) {
                    add_side_op(b, get(FI_F_TXY), L_VVC, 0, c[0], s, f32);
                    add_side_op(b, get(FI_F_TYZ), L_CVV, 0, c[2], s, f32);
                } else {
                    add_side_op(b, get(FI_F_TXZ), L_VCV, 0, c[0], s, f32);
                    add_side_op(b, get(FI_F_TYZ), L_CVV, 0, c[1], s, f32);
                }
            }
        } else {
            if (kind == FI_BC_TRACTION) {
                add_side_op(b, get(FI_F_P), L_CCC, 0, c[d], s, f32);
                add_side_op(b, get(d == 0 ? FI_F_TXX : FI_F_TYY), L_CCC, 0, c[d], s, f32);
            }
            if (kind == FI_BC_TRACTION || kind == FI_BC_SLIP) add_side_op(b, get(FI_F_TXY), L_VVC, 0, c[1 - d], s, f32);
        }
    }
    return b;
}

// velocity batch (boundary_conditions.jl:104-115); eta_next Neumann is added by the caller
template <class G>
BcBatch velocity_batch_g(const fi_model_desc& ds, int d, G get, bool homog, int f32)
{
    BcBatch b;
    b.nop = 0;
    b.d = d;
    const int idV[3] = {FI_F_VX, FI_F_VY, FI_F_VZ};
    const int LV[3] = {L_VCC, L_CVC, L_CCV};
    for (int s = 0; s < 2; ++s) {
        const int kind = ds.bc_kind[d][s];
        double c[3];
        for (int q = 0; q < 3; ++q) c[q] = homog ? 0.0 : ds.bc_value[d][s][q];
        if (kind == FI_BC_VELOCITY) {
            for (int q = 0; q < ds.ndims; ++q) add_side_op(b, get(idV[q]), LV[q], 0, c[q], s, f32);
        } else if (kind == FI_BC_SLIP) {
            add_side_op(b, get(idV[d]), LV[d], 0, c[d], s, f32);
        }
    }
    return b;
}

// residual batch (boundary_conditions.jl:120-128)
template <class G>
BcBatch residual_batch_g(const fi_model_desc& ds, int d, G get, int f32)
{
    BcBatch b;
    b.nop = 0;
    b.d = d;
    const int idR[3] = {FI_F_RVX, FI_F_RVY, FI_F_RVZ};
    const int LV[3] = {L_VCC, L_CVC, L_CCV};
    for (int s = 0; s < 2; ++s)
        if (ds.bc_kind[d][s] != FI_BC_TRACTION) add_side_op(b, get(idR[d]), LV[d], 0, 0.0, s, f32);
    return b;
}

} // namespace fi
