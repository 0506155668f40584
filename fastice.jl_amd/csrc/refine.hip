// refine.hip -- FP32 pseudo-transient sweeps inside an FP64 defect-correction
// (iterative-refinement) loop.
//
// The FP64 model state x = (P, tau, V, eta) is the master.  Two outer loops
// share the FP32 sweeps: lagged PT (default, job_outer_lag below), in which
// the FP32 sweeps are the plain PT iteration in correction form with the
// viscosity lagged by one chunk, and correction solves, described here.
// Each outer step of the correction solves, all in FP64 on the GPU:
//   1. eta <- rheology(tau_II(tau), eta)          (nonlinear rheologies only)
//   2. tau <- 2 eta (edot(V) - div V / 3)          (compute_tau!: zero constitutive defect)
//   3. r_P, r_V <- compute_residuals!(x); Inf-norms; stop when below tol
// then the correction d solves the same PT iteration in FP32 with frozen eta,
// homogeneous boundary values and the FP64 defect as forcing:
//   P_d   -= (div V_d + r_P) eta dtau_P
//   tau_d -= (tau_d - 2 eta edot'(V_d)) dtau_tau
//   V_d   += (-grad P_d + div tau_d + r_V) / eta dtau_V
// (r_V enters as gravity -r_V), until its own residual has dropped by
// inner_rtol; finally x += d in FP64.  Because every operator is linear in the
// state for frozen eta, x + d satisfies the FP64 equations up to the inner
// tolerance and FP32 rounding, and the outer loop removes what remains.
// With a nonlinear rheology the outer loop is a Picard iteration on eta; its
// updates are mixed by Anderson acceleration: with the corrections f_j and
// the applied updates u_j of the last m steps, the update is
//   u_k = f_k - sum_j g_j (u_j + f_{j+1} - f_j),
//   g = argmin || f_k - sum_j g_j (f_{j+1} - f_j) ||   (weighted L2)
// (Walker & Ni 2011, type II with unit mixing), restarted when a step fails
// to reduce the outer residual.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "refine.h"

namespace fi {

struct Refine {
    float* d[10][2] = {};   // correction state P..VZ, two buffers
    float* eta = nullptr;   // frozen viscosity (correction phase) / viscosity (full phase)
    float* eta2 = nullptr;  // second viscosity buffer of the full nonlinear FP32 phase
    float* rh = nullptr;    // FP32 rheology field (LINEAR_FIELD eta or POWERLAW A)
    float* gf[3] = {};      // FP32 gravity fields of the full phase
    int eta_cur = 0;        // full phase: 1 when the current viscosity is eta2
    bool full = false;      // sweeps solve the full state (true) or the correction (false)
    float* fP = nullptr;    // continuity forcing r_P
    float* g[3] = {};       // momentum forcing -r_V
    uint16_t* wna[8] = {};  // 2D masks of the FP32 sweeps (IB), half precision
    uint16_t* wns[8] = {};
    uint16_t* wpk = nullptr; // 3D: the packed mask words (MW_*), 16 bytes per element
    float* r[4] = {};       // inner residuals
    std::vector<void*> allocs;
    int sig_cur = 0, v_cur = 0;
    FusedGeom fg;
    BoxList shell;
    bool has_interior = false;
    int mask_gen = -1; // model mask generation the FP32 masks were cast from
    unsigned long long* d_norm = nullptr;
    unsigned long long* h_norm = nullptr;
    // Anderson history: previous correction f_{k-1} and, per ring slot, the
    // applied update u_j (dX) and the correction difference f_{j+1} - f_j (dF)
    static constexpr int AA_SLOTS = 4;
    float* aa_f[10] = {};
    float* aa_x[AA_SLOTS][10] = {};
    float* aa_df[AA_SLOTS][10] = {};
    double* d_part = nullptr; // per-block partial dot products
    double* h_part = nullptr;

    float* cur(int id) { return d[id][(id <= FI_F_TYZ ? sig_cur : v_cur)]; }
    float* alt(int id) { return d[id][1 - (id <= FI_F_TYZ ? sig_cur : v_cur)]; }
};

// ---------------------------------------------------------------------------
// elementwise kernels over whole allocations
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cast(float* dst, const double* src, double scale, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        dst[t] = (float)(scale * src[t]);
}

// volume fraction -> half precision bits of 2^15 w (round to nearest even),
// the fraction clamped to [0, 1] (h16_decode)
__device__ __forceinline__ uint16_t quantise_h16(double v)
{
    const float w = v > 0.0 ? (v < 1.0 ? (float)v : 1.0f) : 0.0f;
    return __builtin_bit_cast(uint16_t, (_Float16)(w * (float)(1u << H16_SCALE_EXP)));
}

__global__ __launch_bounds__(256) void k_quantise_mask(uint16_t* dst, const double* src, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        dst[t] = quantise_h16(src[t]);
    }
}

// 3D: the seven masks the FP32 sweeps read, in half precision as
// k_quantise_mask writes them, and the validity byte, one 16-byte word per
// element (MW_*)
struct PackSrc {
    const double* w[7];
};
__global__ __launch_bounds__(256) void k_pack_masks(uint4* dst, PackSrc src, const uint8_t* flags, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        uint32_t h[7];
        for (int q = 0; q < 7; ++q) {
            h[q] = quantise_h16(src.w[q][t]);
        }
        dst[t] = make_uint4(h[0] | h[1] << 16, h[2] | h[3] << 16, h[4] | h[5] << 16, h[6] | (uint32_t)flags[t] << 16);
    }
}

__global__ __launch_bounds__(256) void k_fill32(float* p, size_t n, float v)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}

__global__ __launch_bounds__(256) void k_widen(double* dst, const float* src, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        dst[t] = (double)src[t];
}

__global__ __launch_bounds__(256) void k_add_correction(double* x, const float* d, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        x[t] = x[t] + (double)d[t];
}

// ---------------------------------------------------------------------------
// column-restricted kernels of the lagged-PT outer step (3D immersed
// boundary, traction on every face): a thread owns position (i, j, k) of the
// padded box, i in [0, nx+2], j in [0, ny+2], and works only if k lies in its
// column's range (crange: planes holding a valid location; erange: planes
// whose viscosity a valid location reads).  Elsewhere every location is
// invalid: the FP32 corrections, the stress lag and the residuals are 0 there
// and stay 0, so the outer step need not visit it.
// ---------------------------------------------------------------------------
struct ColGeom {
    int nx3, ny3, nz3; // n + 3 per dimension
    long sy, sz;
};

struct LagApply {
    double* x[10];      // FP64 state (current buffers): P, the six stresses, V
    const float* d[10]; // the correction of the chunk (current FP32 buffers)
    float* z[4];        // buffer 0 of the FP32 pressure and velocity, zeroed for the next chunk
};

// x += d for the ten state fields of a 3D model, and buffer 0 of the FP32
// pressure and velocity back to 0, at the positions inside their column's
// range: one thread per position, the workgroups of a launch sweeping whole
// planes (blockIdx.z = plane), and every load issued before the stores.
// Neither the other FP32 buffers nor buffer 0 of the stress need clearing:
// compute_tau writes the stress lag into buffer 0 at every position of the
// ranges (the stress conditions its ghosts), and the next chunk's first
// sweep writes every valid location of buffer 1, which holds +0 at every
// invalid one (no sweep stores anything else there), the conditions after
// it its ghosts.
__global__ __launch_bounds__(256) void k_lag_apply(LagApply a, const int2* crange, ColGeom g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx3 || j >= g.ny3) return;
    const int2 rg = crange[(long)i + g.sy * j];
    if (k < rg.x || k > rg.y) return;
    const long c = (long)i + g.sy * j + g.sz * k;
    double xv[10];
    float dv[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) {
        xv[f] = a.x[f][c];
        dv[f] = a.d[f][c];
    }
#pragma unroll
    for (int f = 0; f < 10; ++f) a.x[f][c] = xv[f] + (double)dv[f];
#pragma unroll
    for (int q = 0; q < 4; ++q) a.z[q][c] = 0.0f;
}

// The forcing and frozen viscosity of the next chunk, cast to FP32 where a
// valid location reads them: the viscosity on the column's erange, the
// forcing (r_P and -r_V) on its crange (a subset).  One thread per column
// and chunk of CAST_KCH planes (blockIdx.z), each plane's five loads issued
// together.
constexpr int CAST_KCH = 16;

struct CastOuter {
    float* eta;
    const double* eta64;
    float* f[4];        // fP, g[0..2]
    const double* r[4]; // rP, rVx, rVy, rVz
    double scale[4];    // 1, -1, -1, -1
};

__global__ __launch_bounds__(256) void k_cast_outer(CastOuter a, const int2* crange, const int2* erange, ColGeom g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx3 || j >= g.ny3) return;
    const long col = (long)i + g.sy * j;
    const int2 re = erange[col], rc = crange[col];
    const int k0 = max((int)blockIdx.z * CAST_KCH, re.x);
    const int k1 = min(min((int)blockIdx.z * CAST_KCH + CAST_KCH - 1, g.nz3 - 1), re.y);
    for (int k = k0; k <= k1; ++k) {
        const long c = col + g.sz * k;
        const bool in = k >= rc.x && k <= rc.y;
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = in ? a.r[q][c] : 0.0;
        a.eta[c] = (float)(1.0 * a.eta64[c]);
        if (in) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a.f[q][c] = (float)(a.scale[q] * v[q]);
        }
    }
}

// Anderson: per-block partial sums of dF_a . dF_b (a <= b < c) and dF_a . f
// for one field, in a fixed block order (deterministic after the host sum)
constexpr int AA_DOT_BLOCKS = 512;
constexpr int AA_MAXC = 3;
constexpr int AA_NDOT = AA_MAXC * (AA_MAXC + 1) / 2 + AA_MAXC + 1;

struct AaDotArgs {
    const float* df[AA_MAXC];
    const float* f;
    int c; // columns in use
};

__global__ __launch_bounds__(256) void k_aa_dots(AaDotArgs a, size_t n, double* part)
{
    double acc[AA_NDOT];
    for (int q = 0; q < AA_NDOT; ++q) acc[q] = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        double v[AA_MAXC];
        for (int q = 0; q < AA_MAXC; ++q) v[q] = q < a.c ? (double)a.df[q][t] : 0.0;
        const double fv = (double)a.f[t];
        int o = 0;
        for (int p = 0; p < AA_MAXC; ++p)
            for (int q = p; q < AA_MAXC; ++q) acc[o++] += v[p] * v[q];
        for (int p = 0; p < AA_MAXC; ++p) acc[o++] += v[p] * fv;
        acc[o] += fv * fv;
    }
    __shared__ double red[256];
    for (int q = 0; q < AA_NDOT; ++q) {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * AA_NDOT + q] = red[0];
        __syncthreads();
    }
}

// dF_new = f - f_prev; f_prev = f (before the mixing)
__global__ __launch_bounds__(256) void k_aa_diff(float* df, float* fprev, const float* f, size_t n, int have_prev)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const float fv = f[t];
        if (have_prev) df[t] = fv - fprev[t];
        fprev[t] = fv;
    }
}

// u = f - sum_j g_j (dX_j + dF_j); x += u; the update is kept as the next dX
struct AaMixArgs {
    const float* dx[AA_MAXC];
    const float* df[AA_MAXC];
    float g[AA_MAXC];
    int c;
};

__global__ __launch_bounds__(256) void k_aa_mix(AaMixArgs a, const float* f, float* u_out, double* x, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        float u = f[t];
        for (int q = 0; q < AA_MAXC; ++q)
            if (q < a.c) u -= a.g[q] * (a.dx[q][t] + a.df[q][t]);
        u_out[t] = u;
        x[t] = x[t] + (double)u;
    }
}

static dim3 ew_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)(b ? b : 1));
}

// Inf-norm of FP32 fields over their interiors (|x| bit patterns as uint32;
// NaN sorts above +Inf)
struct Norm32Args {
    const float* f[4];
    int hi[4][3];
};

__global__ __launch_bounds__(256) void k_norms32(Norm32Args a, long sy, long sz, int nd, unsigned long long* out)
{
    const int q = blockIdx.y;
    const float* f = a.f[q];
    const int hx = a.hi[q][0], hy = a.hi[q][1], hz = a.hi[q][2];
    const long total = (long)hx * hy * hz;
    unsigned m = 0u;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long i = t % hx + 1, r = t / hx;
        long j = r % hy + 1, k = (nd == 3) ? r / hy + 1 : 0;
        unsigned b = __float_as_uint(f[i + sy * j + sz * k]) & 0x7fffffffu;
        m = b > m ? b : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned v = __shfl_xor(m, o);
        m = v > m ? v : m;
    }
    __shared__ unsigned red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = red[w] > r ? red[w] : r;
        atomicMax(out + q, (unsigned long long)r);
    }
}

__global__ void k_zero4(unsigned long long* p)
{
    if (threadIdx.x < 4) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
void free_refine(fi_model_s* m)
{
    Refine* r = m->rf;
    if (!r) return;
    for (void* p : r->allocs) (void)hipFree(p);
    if (r->h_norm) (void)hipHostFree(r->h_norm);
    if (r->h_part) (void)hipHostFree(r->h_part);
    delete r;
    m->rf = nullptr;
}

static int alloc32(fi_model_s* m, float** out)
{
    Refine* r = m->rf;
    float* base = nullptr;
    if (hipMalloc(&base, m->g.elems * sizeof(float)) != hipSuccess)
        return fail(FI_ERR_OOM, "FP32 workspace allocation failed");
    r->allocs.push_back(base);
    HIPCHK(hipMemsetAsync(base, 0, m->g.elems * sizeof(float), m->stream));
    *out = base + m->g.off;
    return FI_OK;
}

// The two buffers of a correction field are one allocation, the second
// pair_stride elements after the first (the same for every field): the fused
// FP32 immersed-boundary sweep addresses both through one base pointer.

static int alloc32_pair(fi_model_s* m, float** b0, float** b1)
{
    Refine* r = m->rf;
    float* base = nullptr;
    const size_t s = pair_stride(m->g);
    if (hipMalloc(&base, 2 * s * sizeof(float)) != hipSuccess)
        return fail(FI_ERR_OOM, "FP32 workspace allocation failed");
    r->allocs.push_back(base);
    HIPCHK(hipMemsetAsync(base, 0, 2 * s * sizeof(float), m->stream));
    *b0 = base + m->g.off;
    *b1 = base + s + m->g.off;
    return FI_OK;
}

static int alloc_mask16(fi_model_s* m, uint16_t** out)
{
    Refine* r = m->rf;
    uint16_t* base = nullptr;
    if (hipMalloc(&base, m->g.elems * sizeof(uint16_t)) != hipSuccess) return fail(FI_ERR_OOM, "FP32 mask allocation failed");
    r->allocs.push_back(base);
    HIPCHK(hipMemsetAsync(base, 0, m->g.elems * sizeof(uint16_t), m->stream));
    *out = base + m->g.off;
    return FI_OK;
}

static int ensure_refine(fi_model_s* m)
{
    if (m->rf) return FI_OK;
    m->rf = new Refine();
    Refine* r = m->rf;
    const int nd = m->g.nd;
    int rc = FI_OK;
    for (int id = FI_F_P; id <= FI_F_VZ && !rc; ++id)
        if (m->f[id])
            rc = alloc32_pair(m, &r->d[id][0], &r->d[id][1]);
    if (!rc) rc = alloc32(m, &r->eta);
    if (!rc) rc = alloc32(m, &r->eta2);
    if (!rc && m->f[FI_F_RHEO]) rc = alloc32(m, &r->rh);
    for (int q = 0; q < nd && !rc; ++q)
        if (m->f[FI_F_GX + q]) rc = alloc32(m, &r->gf[q]);
    if (!rc) rc = alloc32(m, &r->fP);
    for (int q = 0; q < nd && !rc; ++q) rc = alloc32(m, &r->g[q]);
    for (int q = 0; q < (nd == 3 ? 4 : 3) && !rc; ++q) rc = alloc32(m, &r->r[q]);
    if (!rc && m->desc.immersed_boundary) {
        // the masks the sweeps read: na at stress locations, ns at velocity locations
        const int na3[4] = {OM_CCC, OM_VVC, OM_VCV, OM_CVV}, ns3[3] = {OM_VCC, OM_CVC, OM_CCV};
        const int na2[2] = {OM2_CC, OM2_VV}, ns2[2] = {OM2_VC, OM2_CV};
        (void)na3;
        (void)ns3;
        if (nd == 3) {
            uint16_t* base = nullptr;
            if (hipMalloc(&base, m->g.elems * 16) != hipSuccess) {
                rc = fail(FI_ERR_OOM, "FP32 mask allocation failed");
            } else {
                r->allocs.push_back(base);
                if (hipMemsetAsync(base, 0, m->g.elems * 16, m->stream) != hipSuccess) rc = fail(FI_ERR_HIP, "hipMemsetAsync failed");
                r->wpk = base + 8 * m->g.off;
            }
        } else {
            for (int q = 0; q < 2 && !rc; ++q) rc = alloc_mask16(m, &r->wna[na2[q]]);
            for (int q = 0; q < 2 && !rc; ++q) rc = alloc_mask16(m, &r->wns[ns2[q]]);
        }
    }
    if (!rc && (hipMalloc(&r->d_norm, 4 * sizeof(unsigned long long)) != hipSuccess ||
                hipHostMalloc(&r->h_norm, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess))
        rc = fail(FI_ERR_OOM, "norm buffers");
    if (!rc && r->d_norm) r->allocs.push_back(r->d_norm);
    if (rc) {
        free_refine(m);
        return rc;
    }
    // the fused sweep addresses FP32 fields (both buffers of a pair from the
    // first) and the 16-byte mask words with 32-bit byte offsets
    if (nd == 3 && 16 * pair_stride(m->g) < (1ull << 32)) r->has_interior = fused_split(m->g, 16, m->kch, r->fg, r->shell);
    // the allocations are cleared on the model's stream
    m->zs32[0] = m->zs32[1] = m->zv32[0] = m->zv32[1] = true;
    m->gs32[0] = m->gs32[1] = false;
    return FI_OK;
}

static StokesArgs<float> make_args32(fi_model_s* m, bool in_place)
{
    Refine* r = m->rf;
    const fi_model_desc& d = m->desc;
    StokesArgs<float> a;
    std::memset(&a, 0, sizeof(a));
    // every stress field reads buffer sig_cur and every velocity field buffer
    // v_cur of its pair (alloc32_pair): the fused immersed-boundary sweep
    // relies on one input and one output offset per group
    auto in = [&](int id) { return r->cur(id); };
    auto out = [&](int id) { return in_place ? r->cur(id) : r->alt(id); };
    a.P = in(FI_F_P); a.txx = in(FI_F_TXX); a.tyy = in(FI_F_TYY); a.tzz = in(FI_F_TZZ);
    a.txy = in(FI_F_TXY); a.txz = in(FI_F_TXZ); a.tyz = in(FI_F_TYZ);
    a.Vx = in(FI_F_VX); a.Vy = in(FI_F_VY); a.Vz = in(FI_F_VZ);
    a.Pn = out(FI_F_P); a.txxn = out(FI_F_TXX); a.tyyn = out(FI_F_TYY); a.tzzn = out(FI_F_TZZ);
    a.txyn = out(FI_F_TXY); a.txzn = out(FI_F_TXZ); a.tyzn = out(FI_F_TYZ);
    a.Vxn = out(FI_F_VX); a.Vyn = out(FI_F_VY); a.Vzn = out(FI_F_VZ);
    if (r->full) {
        // the full nonlinear PT iteration in FP32 (real forcing, boundary values and rheology)
        a.eta = r->eta_cur ? r->eta2 : r->eta;
        a.eta_next = r->eta_cur ? r->eta : r->eta2;
        a.frozen_eta = 0;
        a.fP = nullptr;
        a.gx = r->gf[0]; a.gy = r->gf[1]; a.gz = r->gf[2];
        a.grav_field = d.gravity_is_field;
        for (int q = 0; q < 3; ++q) a.grav[q] = (float)d.gravity[q];
    } else {
        a.eta = r->eta;
        a.eta_next = nullptr;
        a.frozen_eta = 1;
        a.fP = r->fP;
        a.gx = r->g[0]; a.gy = r->g[1]; a.gz = r->g[2];
        a.grav_field = 1;
    }
    a.rP = r->r[0]; a.rVx = r->r[1]; a.rVy = r->r[2]; a.rVz = r->r[3];
    for (int q = 0; q < 8; ++q) {
        a.na[q] = H16Mask{r->wna[q], 1};
        a.ns[q] = H16Mask{r->wns[q], 1};
    }
    if (r->wpk) {
        a.na[OM_CCC] = H16Mask{r->wpk + MW_NA_CCC, 8};
        a.na[OM_VVC] = H16Mask{r->wpk + MW_NA_VVC, 8};
        a.na[OM_VCV] = H16Mask{r->wpk + MW_NA_VCV, 8};
        a.na[OM_CVV] = H16Mask{r->wpk + MW_NA_CVV, 8};
        a.ns[OM_VCC] = H16Mask{r->wpk + MW_NS_VCC, 8};
        a.ns[OM_CVC] = H16Mask{r->wpk + MW_NS_CVC, 8};
        a.ns[OM_CCV] = H16Mask{r->wpk + MW_NS_CCV, 8};
        a.mw = reinterpret_cast<const uint4*>(r->wpk);
    }
    a.flags = m->flags;
    a.crange = m->crange;
    a.erange = m->erange;
    a.tmarch = m->tmarch; // the full FP32 phase updates eta: k_eta_rest covers the unmarched planes
    a.nx = m->g.n[0]; a.ny = m->g.n[1]; a.nz = m->g.n[2];
    a.sy = m->g.sy; a.sz = m->g.sz;
    a.ix = (float)m->g.idx[0]; a.iy = (float)m->g.idx[1]; a.iz = (float)m->g.idx[2];
    a.dPr = (float)d.dtau_P;
    for (int q = 0; q < 6; ++q) a.dtau[q] = (float)d.dtau_tau[q];
    for (int q = 0; q < 3; ++q) a.dV[q] = (float)d.dtau_V[q];
    a.rh_kind = d.rheology.kind;
    a.rh_has_reg = d.rheology.has_reg;
    a.rh_relax = d.rheology.relax;
    a.rh_eta = (float)d.rheology.eta;
    a.rh_A = (float)d.rheology.A;
    a.rh_n = (float)d.rheology.n;
    a.rh_inv_eta_reg = d.rheology.has_reg ? (float)(1.0 / d.rheology.eta_reg) : 0.0f;
    a.rh_rate = (float)d.rheology.rate;
    a.rh = r->rh;
    return a;
}

BcBatch refine_bc_batch(fi_model_s* m, int group, int d)
{
    Refine* r = m->rf;
    BcBatch b;
    const bool homog = !r->full;
    if (group == FI_BCGROUP_STRESS)
        b = stress_batch_g(m->desc, d, [r](int id) { return (void*)r->cur(id); }, homog, 1);
    else if (group == FI_BCGROUP_VELOCITY) {
        b = velocity_batch_g(m->desc, d, [r](int id) { return (void*)r->cur(id); }, homog, 1);
        if (r->full) add_op(b, r->eta_cur ? r->eta : r->eta2, L_CCC, 1, 0.0, 1); // eta_next Neumann
    } else
        b = residual_batch_g(m->desc, d, [r](int id) { return (void*)r->r[id - FI_F_RP]; }, 1);
    return b;
}

// skip: as run_bc
static void bc32(fi_model_s* m, int group, bool skip = false)
{
    BcBatch b[3];
    for (int d = 0; d < m->g.nd; ++d) b[d] = refine_bc_batch(m, group, d);
    launch_bc_dims(m->g, b, m->stream, skip ? m->crange : nullptr, skip ? m->empty_faces : 0u);
}

StokesArgs<float> refine_args(fi_model_s* m, bool in_place) { return make_args32(m, in_place); }

bool refine_fused(fi_model_s* m, FusedGeom& fg, BoxList& shell)
{
    Refine* r = m->rf;
    const bool fused = m->use_fused && r->has_interior;
    fg = r->fg;
    fg.order = m->forder;
    shell = fused ? r->shell : full_box(m->g);
    return fused;
}

void refine_flip(fi_model_s* m)
{
    Refine* r = m->rf;
    r->sig_cur ^= 1;
    r->v_cur ^= 1;
    if (r->full) r->eta_cur ^= 1;
}

RefineParity refine_parity(fi_model_s* m) { return RefineParity{m->rf->sig_cur, m->rf->v_cur, m->rf->eta_cur}; }

void refine_set_parity(fi_model_s* m, RefineParity p)
{
    m->rf->sig_cur = p.sig;
    m->rf->v_cur = p.v;
    m->rf->eta_cur = p.eta;
}

bool refine_full(fi_model_s* m) { return m->rf->full; }

// one FP32 PT iteration of the correction problem
static void iterate32(fi_model_s* m, hipEvent_t* ev)
{
    Refine* r = m->rf;
    const bool ib = m->desc.immersed_boundary != 0;
    StokesArgs<float> a = make_args32(m, false);
    const int so = 1 - r->sig_cur, vo = 1 - r->v_cur; // the buffers this sweep writes
    a.z_read = m->zv32[r->v_cur];
    a.z_wsig = m->zs32[so];
    a.z_wvel = m->zv32[vo];
    const bool fused = m->use_fused && r->has_interior;
    const BoxList boxes = fused ? r->shell : full_box(m->g);
    // shell boxes in a face layer without a valid location are left out (as in iterate)
    const BoxList sboxes = ib && a.z_wsig ? drop_empty_boxes(m->g, boxes, m->empty_faces) : boxes;
    const BoxList vboxes = ib && a.z_wvel && a.frozen_eta ? drop_empty_boxes(m->g, boxes, m->empty_faces) : boxes;
    if (ev) (void)hipEventRecord(ev[0], m->stream);
    if (fused) {
        FusedGeom fg = r->fg;
        fg.order = m->forder;
        launch_fused(m->g, a, fg, ib, m->ftx32, m->fty32, m->fminw32, m->stream, m->fused_shell ? &sboxes : nullptr);
    }
    if (ev) (void)hipEventRecord(ev[1], m->stream);
    if (!(fused && m->fused_shell)) launch_sigma_boxes(m->g, a, sboxes, ib, m->stream);
    r->sig_cur ^= 1;
    if (ev) (void)hipEventRecord(ev[2], m->stream);
    // the sweep stored nothing at invalid locations (z_wsig): the ghosts the
    // last application on this buffer settled still hold their value
    const bool skip = ib && m->g.nd == 3 && m->crange && a.z_wsig && m->gs32[so];
    bc32(m, FI_BCGROUP_STRESS, skip);
    if (ib && m->g.nd == 3) m->gs32[so] = a.z_wsig != 0;
    if (ev) (void)hipEventRecord(ev[3], m->stream);
    launch_velocity_boxes(m->g, a, vboxes, ib, m->stream);
    r->v_cur ^= 1;
    if (ev) (void)hipEventRecord(ev[4], m->stream);
    bc32(m, FI_BCGROUP_VELOCITY);
    if (ev) (void)hipEventRecord(ev[5], m->stream);
    if (r->full) r->eta_cur ^= 1;
    // every invalid location of the written buffers now holds +0
    m->zs32[so] = true;
    m->zv32[vo] = true;
}

// FP32 residual Inf-norm of the correction state: enqueue, then read once
// the stream has synchronised
static int inner_norm_enqueue(fi_model_s* m)
{
    Refine* r = m->rf;
    const bool ib = m->desc.immersed_boundary != 0;
    StokesArgs<float> a = make_args32(m, true);
    launch_residuals(m->g, a, ib, m->stream);
    bc32(m, FI_BCGROUP_RESIDUAL);
    Norm32Args na;
    const int L[4] = {L_CCC, L_VCC, L_CVC, L_CCV};
    const int nf = m->g.nd == 3 ? 4 : 3;
    for (int q = 0; q < 4; ++q) {
        const int qq = q < nf ? q : 0;
        na.f[q] = r->r[qq];
        for (int d = 0; d < 3; ++d) na.hi[q][d] = (d < m->g.nd) ? m->g.n[d] + ((L[qq] >> d) & 1) : 1;
    }
    hipLaunchKernelGGL(k_zero4, dim3(1), dim3(64), 0, m->stream, r->d_norm);
    long total = (long)na.hi[0][0] * na.hi[0][1] * na.hi[0][2];
    int blocks = (int)((total + 255) / 256);
    blocks = blocks > 1024 ? 1024 : (blocks < 1 ? 1 : blocks);
    hipLaunchKernelGGL(k_norms32, dim3(blocks, nf, 1), dim3(256), 0, m->stream, na, m->g.sy, m->g.sz, m->g.nd, r->d_norm);
    HIPCHK(hipMemcpyAsync(r->h_norm, r->d_norm, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, m->stream));
    return FI_OK;
}

static void inner_norm_read(fi_model_s* m, double* out)
{
    Refine* r = m->rf;
    const int nf = m->g.nd == 3 ? 4 : 3;
    double mx = 0.0;
    for (int q = 0; q < nf; ++q) {
        unsigned b = (unsigned)r->h_norm[q];
        float v;
        std::memcpy(&v, &b, sizeof(v));
        if (!(v == v) || b >= 0x7f800000u) {
            *out = (double)v;
            return;
        }
        mx = (double)v > mx ? (double)v : mx;
    }
    *out = mx;
}

static int inner_norm(fi_model_s* m, double* out)
{
    int rc = inner_norm_enqueue(m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    inner_norm_read(m, out);
    return FI_OK;
}

static void cast(fi_model_s* m, float* dst, const double* src, double scale)
{
    hipLaunchKernelGGL(k_cast, ew_grid(m->g.elems), dim3(256), 0, m->stream, dst - m->g.off, src - m->g.off, scale,
                       m->g.elems);
}

static ColGeom col_geom(const fi_model_s* m)
{
    return ColGeom{m->g.n[0] + 3, m->g.n[1] + 3, m->g.n[2] + 3, m->g.sy, m->g.sz};
}

static dim3 plane_grid(const fi_model_s* m)
{
    return dim3((unsigned)((m->g.n[0] + 3 + 63) / 64), (unsigned)((m->g.n[1] + 3 + 3) / 4), (unsigned)(m->g.n[2] + 3));
}

// The lagged-PT outer step may visit only the column ranges: a 3D immersed
// boundary with traction on every face.  Then the boundary conditions set no
// velocity ghost and set every stress ghost from interior values after each
// update, so ghost positions outside the ranges never need zeroing.
static bool lag_cols(const fi_model_s* m)
{
    if (!(m->refine_cols && m->desc.immersed_boundary && m->g.nd == 3 && m->crange)) return false;
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s)
            if (m->desc.bc_kind[d][s] != FI_BC_TRACTION) return false;
    return true;
}

static void quantise_mask(fi_model_s* m, uint16_t* dst, const double* src)
{
    hipLaunchKernelGGL(k_quantise_mask, ew_grid(m->g.elems), dim3(256), 0, m->stream, dst - m->g.off, src - m->g.off,
                       m->g.elems);
}

// the FP32 masks (and, in 3D, the validity bytes packed with them) from the
// model's FP64 masks, whenever those changed; after prepare()
static void sync_masks(fi_model_s* m)
{
    Refine* r = m->rf;
    if (!m->desc.immersed_boundary || r->mask_gen == m->mask_gen) return;
    if (r->wpk) {
        PackSrc src;
        const int na[4] = {OM_CCC, OM_VVC, OM_VCV, OM_CVV}, ns[3] = {OM_VCC, OM_CVC, OM_CCV};
        for (int q = 0; q < 4; ++q) src.w[q] = m->f[FI_F_OM_NA0 + na[q]] - m->g.off;
        for (int q = 0; q < 3; ++q) src.w[4 + q] = m->f[FI_F_OM_NS0 + ns[q]] - m->g.off;
        hipLaunchKernelGGL(k_pack_masks, ew_grid(m->g.elems), dim3(256), 0, m->stream,
                           reinterpret_cast<uint4*>(r->wpk - 8 * m->g.off), src, m->flags - m->g.off, m->g.elems);
    }
    for (int q = 0; q < 8; ++q) {
        if (r->wna[q]) quantise_mask(m, r->wna[q], m->f[FI_F_OM_NA0 + q]);
        if (r->wns[q]) quantise_mask(m, r->wns[q], m->f[FI_F_OM_NS0 + q]);
    }
    r->mask_gen = m->mask_gen;
}

static void widen(fi_model_s* m, double* dst, const float* src)
{
    hipLaunchKernelGGL(k_widen, ew_grid(m->g.elems), dim3(256), 0, m->stream, dst - m->g.off, src - m->g.off, m->g.elems);
}



static int ensure_anderson(fi_model_s* m, int depth)
{
    Refine* r = m->rf;
    int rc = FI_OK;
    for (int id = FI_F_P; id <= FI_F_VZ && !rc; ++id) {
        if (!m->f[id]) continue;
        if (!r->aa_f[id]) rc = alloc32(m, &r->aa_f[id]);
        for (int sl = 0; sl <= depth && !rc; ++sl) {
            if (!r->aa_x[sl][id]) rc = alloc32(m, &r->aa_x[sl][id]);
            if (!rc && !r->aa_df[sl][id]) rc = alloc32(m, &r->aa_df[sl][id]);
        }
    }
    if (!rc && !r->d_part) {
        const size_t bytes = (size_t)AA_DOT_BLOCKS * AA_NDOT * sizeof(double);
        if (hipMalloc(&r->d_part, bytes) != hipSuccess || hipHostMalloc(&r->h_part, bytes, hipHostMallocDefault) != hipSuccess)
            return fail(FI_ERR_OOM, "Anderson workspace");
        r->allocs.push_back(r->d_part);
    }
    return rc;
}

// Mix the new correction (in the current correction buffers) with the
// history and add the update to the FP64 state.  restart: drop the history.
static int anderson_apply(fi_model_s* m, AaState& st, bool restart)
{
    Refine* r = m->rf;
    const size_t n = m->g.elems;
    const int slots = st.depth + 1;
    if (restart) st.cols = 0;
    // new column: the update of the previous step and f_k - f_{k-1} in its slot
    for (int id = FI_F_P; id <= FI_F_VZ; ++id)
        if (m->f[id])
            hipLaunchKernelGGL(k_aa_diff, ew_grid(n), dim3(256), 0, m->stream, r->aa_df[st.prev_slot][id] - m->g.off,
                               r->aa_f[id] - m->g.off, r->cur(id) - m->g.off, n, st.have_prev);
    if (st.have_prev && !restart) st.cols = st.cols + 1 < st.depth ? st.cols + 1 : st.depth;
    const int c = st.cols;
    int col[AA_MAXC];
    for (int q = 0; q < c; ++q) col[q] = ((st.prev_slot - q) % slots + slots) % slots;
    float gam[AA_MAXC] = {0.0f, 0.0f, 0.0f};
    int used = 0;
    if (c > 0) {
        // weighted least squares over the fields: stress and velocity groups
        // each weighted by the inverse of their |f|^2
        double G[AA_MAXC][AA_MAXC] = {}, b[AA_MAXC] = {}, ff[2] = {0.0, 0.0};
        double Gf[2][AA_MAXC][AA_MAXC] = {}, bf[2][AA_MAXC] = {};
        for (int id = FI_F_P; id <= FI_F_VZ; ++id) {
            if (!m->f[id]) continue;
            AaDotArgs a;
            for (int q = 0; q < AA_MAXC; ++q) a.df[q] = q < c ? r->aa_df[col[q]][id] - m->g.off : r->cur(id) - m->g.off;
            a.f = r->cur(id) - m->g.off;
            a.c = c;
            hipLaunchKernelGGL(k_aa_dots, dim3(AA_DOT_BLOCKS), dim3(256), 0, m->stream, a, n, r->d_part);
            HIPCHK(hipMemcpyAsync(r->h_part, r->d_part, (size_t)AA_DOT_BLOCKS * AA_NDOT * sizeof(double),
                                  hipMemcpyDeviceToHost, m->stream));
            HIPCHK(hipStreamSynchronize(m->stream));
            double dots[AA_NDOT] = {};
            for (int bl = 0; bl < AA_DOT_BLOCKS; ++bl)
                for (int q = 0; q < AA_NDOT; ++q) dots[q] += r->h_part[(size_t)bl * AA_NDOT + q];
            const int grp = fi_model_s::is_vel(id) ? 1 : 0;
            int o = 0;
            for (int p = 0; p < AA_MAXC; ++p)
                for (int q = p; q < AA_MAXC; ++q, ++o)
                    if (p < c && q < c) {
                        Gf[grp][p][q] += dots[o];
                        if (p != q) Gf[grp][q][p] += dots[o];
                    }
            for (int p = 0; p < AA_MAXC; ++p, ++o)
                if (p < c) bf[grp][p] += dots[o];
            ff[grp] += dots[o];
        }
        for (int grp = 0; grp < 2; ++grp) {
            const double w = ff[grp] > 0.0 ? 1.0 / ff[grp] : 0.0;
            for (int p = 0; p < c; ++p) {
                b[p] += w * bf[grp][p];
                for (int q = 0; q < c; ++q) G[p][q] += w * Gf[grp][p][q];
            }
        }
        // Tikhonov-regularised normal equations, Gauss elimination
        double dmax = 0.0;
        for (int p = 0; p < c; ++p) dmax = G[p][p] > dmax ? G[p][p] : dmax;
        for (int p = 0; p < c; ++p) G[p][p] += 1e-10 * dmax;
        double x[AA_MAXC] = {};
        bool ok = dmax > 0.0;
        for (int p = 0; p < c && ok; ++p) {
            int piv = p;
            for (int q = p + 1; q < c; ++q)
                if (std::fabs(G[q][p]) > std::fabs(G[piv][p])) piv = q;
            if (!(std::fabs(G[piv][p]) > 0.0)) { ok = false; break; }
            for (int q = 0; q < c; ++q) std::swap(G[p][q], G[piv][q]);
            std::swap(b[p], b[piv]);
            for (int q = p + 1; q < c; ++q) {
                const double f = G[q][p] / G[p][p];
                for (int t = p; t < c; ++t) G[q][t] -= f * G[p][t];
                b[q] -= f * b[p];
            }
        }
        for (int p = c - 1; p >= 0 && ok; --p) {
            double v = b[p];
            for (int q = p + 1; q < c; ++q) v -= G[p][q] * x[q];
            x[p] = v / G[p][p];
            if (!std::isfinite(x[p]) || std::fabs(x[p]) > 1e3) ok = false;
        }
        if (ok) {
            used = c;
            for (int p = 0; p < c; ++p) gam[p] = (float)x[p];
        } else {
            st.cols = 0;
        }
    }
    const int uslot = (st.prev_slot + 1) % slots;
    for (int id = FI_F_P; id <= FI_F_VZ; ++id) {
        if (!m->f[id]) continue;
        AaMixArgs a;
        for (int q = 0; q < AA_MAXC; ++q) {
            const int sl = q < used ? col[q] : uslot;
            a.dx[q] = r->aa_x[sl][id] - m->g.off;
            a.df[q] = r->aa_df[sl][id] - m->g.off;
            a.g[q] = q < used ? gam[q] : 0.0f;
        }
        a.c = used;
        hipLaunchKernelGGL(k_aa_mix, ew_grid(n), dim3(256), 0, m->stream, a, r->cur(id) - m->g.off,
                           r->aa_x[uslot][id] - m->g.off, m->buf(id) - m->g.off, n);
    }
    st.have_prev = 1;
    st.prev_slot = uslot;
    HIPCHK(hipGetLastError());
    return FI_OK;
}

// Start of phase 1 of a nonlinear refined solve: the full nonlinear PT
// iteration (eta updated every sweep, real boundary values and forcing) on
// the whole state in FP32, from the FP64 state.  It runs until the FP32
// residual has dropped by rtol or stops decreasing; the result then becomes
// the FP64 state (phase1_end).
static void phase1_begin(fi_model_s* m)
{
    Refine* r = m->rf;
    r->full = true;
    for (int id = FI_F_P; id <= FI_F_VZ; ++id)
        if (m->f[id]) {
            cast(m, r->d[id][0], m->buf(id), 1.0);
            cast(m, r->d[id][1], m->buf(id), 1.0);
        }
    r->sig_cur = r->v_cur = 0;
    m->zs32[0] = m->zs32[1] = m->zv32[0] = m->zv32[1] = false;
    m->gs32[0] = m->gs32[1] = false;
    cast(m, r->eta, m->eta(), 1.0);
    cast(m, r->eta2, m->eta_next(), 1.0);
    r->eta_cur = 0;
    if (r->rh) cast(m, r->rh, m->f[FI_F_RHEO], 1.0);
    for (int q = 0; q < 3; ++q)
        if (r->gf[q]) cast(m, r->gf[q], m->f[FI_F_GX + q], 1.0);
}

static void phase1_end(fi_model_s* m)
{
    Refine* r = m->rf;
    for (int id = FI_F_P; id <= FI_F_VZ; ++id)
        if (m->f[id]) widen(m, m->buf(id), r->cur(id));
    widen(m, m->eta(), r->eta_cur ? r->eta2 : r->eta);
    m->state_written();
    r->full = false;
}

static void fill32(fi_model_s* m, float* p, float v)
{
    hipLaunchKernelGGL(k_fill32, ew_grid(m->g.elems), dim3(256), 0, m->stream, p - m->g.off, m->g.elems, v);
}

// The refined solve as a resumable job: its FP32 sweeps come in chunks of
// `chunk` sweeps (run by fi_refined_solve or packed with other models by a
// batch), and after each chunk job_after_chunk checks the FP32 residual and
// moves on (more sweeps, end of phase 1, an FP64 outer step, a new
// correction problem, or done).



static void next_chunk(RefineJob& j)
{
    if (j.state == JOB_INNER && j.mode == 1) {
        j.chunk = j.ncheck; // lagged PT: one chunk between outer steps
        return;
    }
    j.chunk = j.ncheck < j.max_inner - j.done ? j.ncheck : j.max_inner - j.done;
}

// top of an outer step: eta (Picard, relaxed), tau, residuals; then either
// done or the next correction problem
static int job_outer_lag(fi_model_s* m, RefineJob& j);

static int job_outer(fi_model_s* m, RefineJob& j)
{
    if (j.mode == 1) return job_outer_lag(m, j);
    Refine* r = m->rf;
    const bool ib = m->desc.immersed_boundary != 0;
    const int nd = m->g.nd;
    if (j.nonlinear && (j.outer > 0 || m->refine_full_phase)) {
        // Picard update of the frozen viscosity, relaxed in log space
        StokesArgs<double> a = make_args(m, true);
        a.rh_relax = FI_RELAX_LOG;
        a.rh_rate = j.omega;
        launch_eta_update(m->g, a, m->stream);
        m->eta_cur ^= 1;
        BcBatch b[3];
        for (int d = 0; d < nd; ++d) b[d] = neumann_batch(m->eta(), L_CCC, d);
        launch_bc_dims(m->g, b, m->stream);
    }
    {
        // tau <- 2 eta edot'(V), then the stress boundary conditions the PT
        // iteration applies after every stress update (boundary_conditions.jl:26-100)
        StokesArgs<double> a = make_args(m, true);
        launch_tau(m->g, a, ib, m->stream);
        run_bc(m, FI_BCGROUP_STRESS);
        m->state_written();
    }
    int rc = residual_norms(m, j.norms, false);
    if (rc) return rc;
    double r0 = 0.0;
    for (int q = 0; q < 4; ++q) {
        if (!(j.norms[q] == j.norms[q]) || j.norms[q] > 1.79e308) {
            j.state = JOB_DONE;
            return fail(FI_ERR_NONFINITE, "simulation failed at outer step %d: non-finite residual", j.outer);
        }
        r0 = j.norms[q] > r0 ? j.norms[q] : r0;
    }
    if (j.trace)
        std::fprintf(stderr, "[fastice] refined: outer %d residual %.3e eta rate %.3f FP32 iterations so far %d\n",
                     j.outer, r0, j.omega, j.inner_total);
    // outer viscosity relaxation: halved whenever an outer step fails to
    // reduce the residual by 10 % (with Anderson mixing the rate stays >= 1/2:
    // a frozen viscosity update would stall the mixed iteration)
    if (j.r_prev > 0.0 && r0 > 0.9 * j.r_prev && j.omega > (j.aa.depth > 0 ? 0.5 : 1.0 / 64)) j.omega *= 0.5;
    j.aa_restart = j.r_prev > 0.0 && r0 > j.r_prev; // the last mixed update did not help
    j.r_prev = r0;
    j.r0 = r0;
    if (r0 < j.tol || j.outer >= j.max_outer) {
        j.converged = r0 < j.tol;
        j.state = JOB_DONE;
        j.chunk = 0;
        return FI_OK;
    }
    // correction problem: frozen eta, forcing from the FP64 defect, zero state
    cast(m, r->eta, m->eta(), 1.0);
    cast(m, r->fP, m->f[FI_F_RP], 1.0);
    const int idR[3] = {FI_F_RVX, FI_F_RVY, FI_F_RVZ};
    for (int q = 0; q < nd; ++q) cast(m, r->g[q], m->f[idR[q]], -1.0);
    for (int id = FI_F_P; id <= FI_F_VZ; ++id)
        if (m->f[id]) {
            fill32(m, r->d[id][0], 0.0f);
            fill32(m, r->d[id][1], 0.0f);
        }
    r->sig_cur = r->v_cur = 0;
    m->zs32[0] = m->zs32[1] = m->zv32[0] = m->zv32[1] = true;
    m->gs32[0] = m->gs32[1] = false;
    j.state = JOB_INNER;
    j.done = 0;
    next_chunk(j);
    return FI_OK;
}

// Lagged-PT outer step.  The FP64 state (P, tau, V, eta) is the PT iterate
// itself, tau included.  Write T = 2 eta edot'(V) (compute_tau!) over tau and
// let the FP32 stress correction start from the lag tau - T: the correction
// sweeps are then, to rounding, the plain PT sweeps from the FP64 state with
// eta frozen (the residuals evaluated with T plus the lag give the same
// updates).  eta itself is updated here, between chunks, by the relaxation
// the PT iteration applies every sweep, compounded over the chunk:
// rate_C = 1 - (1 - rate)^C.
static int job_outer_lag_enqueue(fi_model_s* m, RefineJob& j)
{
    Refine* r = m->rf;
    const bool ib = m->desc.immersed_boundary != 0;
    const int nd = m->g.nd;
    if (j.nonlinear && j.since_eta > 0) {
        StokesArgs<double> a = make_args(m, true);
        // after a column-range outer step, the stress is 0 at every invalid
        // location: outside erange the update needs no stress
        a.skip_empty = j.cols_clean;
        const fi_rheology& rh = m->desc.rheology;
        if (rh.relax != FI_RELAX_NONE) a.rh_rate = 1.0 - std::pow(1.0 - rh.rate, (double)j.since_eta);
        launch_eta_update(m->g, a, m->stream);
        m->eta_cur ^= 1;
        BcBatch b[3];
        for (int d = 0; d < nd; ++d) b[d] = neumann_batch(m->eta(), L_CCC, d);
        launch_bc_dims(m->g, b, m->stream);
    }
    j.since_eta = 0;
    // column ranges only: the last chunk's k_lag_apply has zeroed the FP32
    // buffers there, and they hold 0 elsewhere
    const bool cols = j.cols_clean;
    if (!cols)
        for (int id = FI_F_P; id <= FI_F_VZ; ++id)
            if (m->f[id]) {
                fill32(m, r->d[id][0], 0.0f);
                fill32(m, r->d[id][1], 0.0f);
            }
    r->sig_cur = r->v_cur = 0;
    // the lag written below is +0 at every invalid location, like the fill
    m->zs32[0] = m->zs32[1] = m->zv32[0] = m->zv32[1] = true;
    m->gs32[0] = m->gs32[1] = false;
    {
        StokesArgs<double> a = make_args(m, true);
        a.skip_empty = cols;
        const int tid[6] = {FI_F_TXX, FI_F_TYY, FI_F_TZZ, FI_F_TXY, FI_F_TXZ, FI_F_TYZ};
        TauLag lg;
        for (int q = 0; q < 6; ++q) lg.t[q] = m->f[tid[q]] ? r->d[tid[q]][0] : nullptr;
        launch_tau(m->g, a, ib, m->stream, &lg);
        run_bc(m, FI_BCGROUP_STRESS);
        m->state_written();
    }
    bc32(m, FI_BCGROUP_STRESS); // homogeneous boundary values of the lag
    int rc = residual_norms_enqueue(m, false);
    if (rc) return rc;
    // forcing and frozen viscosity of the next chunk (unused if the solve ends
    // here); with column ranges, the forcing where a valid location reads it
    // and the viscosity on the planes a valid location's averages reach
    const int idR[3] = {FI_F_RVX, FI_F_RVY, FI_F_RVZ};
    if (cols) {
        CastOuter ca;
        ca.eta = r->eta;
        ca.eta64 = m->eta();
        ca.f[0] = r->fP;
        ca.r[0] = m->f[FI_F_RP];
        ca.scale[0] = 1.0;
        for (int q = 0; q < 3; ++q) {
            ca.f[1 + q] = r->g[q];
            ca.r[1 + q] = m->f[idR[q]];
            ca.scale[1 + q] = -1.0;
        }
        const dim3 grd((unsigned)((m->g.n[0] + 3 + 63) / 64), (unsigned)((m->g.n[1] + 3 + 3) / 4),
                       (unsigned)((m->g.n[2] + 3 + CAST_KCH - 1) / CAST_KCH));
        hipLaunchKernelGGL(k_cast_outer, grd, dim3(64, 4), 0, m->stream, ca, m->crange, m->erange, col_geom(m));
    } else {
        cast(m, r->eta, m->eta(), 1.0);
        cast(m, r->fP, m->f[FI_F_RP], 1.0);
        for (int q = 0; q < nd; ++q) cast(m, r->g[q], m->f[idR[q]], -1.0);
    }
    j.cols_clean = lag_cols(m);
    return FI_OK;
}

// after the stream has synchronised: the outer residual decides
static int job_outer_lag_finish(fi_model_s* m, RefineJob& j)
{
    residual_norms_read(m, j.norms);
    double r0 = 0.0;
    for (int q = 0; q < 4; ++q) {
        if (!(j.norms[q] == j.norms[q]) || j.norms[q] > 1.79e308) {
            j.state = JOB_DONE;
            return fail(FI_ERR_NONFINITE, "simulation failed at outer step %d: non-finite residual", j.outer);
        }
        r0 = j.norms[q] > r0 ? j.norms[q] : r0;
    }
    if (j.trace)
        std::fprintf(stderr, "[fastice] refined (lagged PT): outer %d residual %.3e FP32 iterations so far %d\n",
                     j.outer, r0, j.inner_total);
    j.r0 = r0;
    if (r0 < j.tol || j.inner_total >= j.sweep_cap) {
        j.converged = r0 < j.tol;
        j.state = JOB_DONE;
        j.chunk = 0;
        // the outer step wrote T = 2 eta edot'(V) over tau and kept the PT
        // stress memory tau - T in the FP32 lag: give tau back its PT value,
        // so that a later solve or iteration continues the same iterate
        for (int id = FI_F_TXX; id <= FI_F_TYZ; ++id)
            if (m->f[id])
                hipLaunchKernelGGL(k_add_correction, ew_grid(m->g.elems), dim3(256), 0, m->stream,
                                   m->buf(id) - m->g.off, m->rf->cur(id) - m->g.off, m->g.elems);
        m->state_written();
        return FI_OK;
    }
    // a viscosity lag of ncheck sweeps can hold a stiff problem in a limit
    // cycle: when the residual has not halved for lag_window sweeps, halve
    // the chunk (the lag) and start watching again
    if (j.lag_best < 0.0 || r0 < 0.5 * j.lag_best) {
        j.lag_best = r0;
        j.lag_stall = 0;
    } else if ((j.lag_stall += j.last) >= j.lag_window && j.ncheck > 8) {
        j.ncheck = j.ncheck / 2 > 8 ? j.ncheck / 2 : 8;
        j.lag_best = r0;
        j.lag_stall = 0;
        if (j.trace) std::fprintf(stderr, "[fastice] refined (lagged PT): chunk now %d sweeps\n", j.ncheck);
    }
    j.state = JOB_INNER;
    j.done = 0;
    next_chunk(j);
    return FI_OK;
}

static int job_outer_lag(fi_model_s* m, RefineJob& j)
{
    int rc = job_outer_lag_enqueue(m, j);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return job_outer_lag_finish(m, j);
}

int job_begin(fi_model_s* m, RefineJob& j, int max_outer, int max_inner, int ncheck, double inner_rtol,
                     double tol)
{
    if (max_outer < 1 || max_inner < 1 || ncheck < 1 || !(inner_rtol > 0.0) || !(tol > 0.0))
        return fail(FI_ERR_ARG, "bad refinement parameters");
    int rc = ensure_refine(m);
    if (rc) return rc;
    rc = prepare(m, false);
    if (rc) return rc;
    sync_masks(m);
    j = RefineJob();
    j.max_outer = max_outer;
    j.max_inner = max_inner;
    j.ncheck = ncheck;
    j.inner_rtol = inner_rtol;
    j.tol = tol;
    j.trace = std::getenv("FASTICE_TRACE") != nullptr;
    const fi_rheology& rh = m->desc.rheology;
    j.nonlinear = rh.kind == FI_RHEO_POWERLAW || rh.relax != FI_RELAX_NONE;
    j.mode = m->refine_mode;
    j.lag_window = 8 * std::max(m->g.n[0], std::max(m->g.n[1], m->g.nd == 3 ? m->g.n[2] : 0));
    j.sweep_cap = (long)max_outer * ncheck;
    j.omega = m->refine_eta_rate;
    j.aa.depth = j.nonlinear && j.mode == 0 ? m->refine_anderson : 0;
    if (j.aa.depth > 0) {
        rc = ensure_anderson(m, j.aa.depth);
        if (rc) return rc;
    }
    if (j.nonlinear && m->refine_full_phase) {
        phase1_begin(m);
        // PT residuals are not monotone and rise first: phase 1 stops after a
        // drop by 10^-digits from the peak residual, or when the best residual
        // has not halved over 4 n_max sweeps since (the FP32 floor; a PT
        // transient spans about n_max sweeps, whatever the check cadence)
        j.p1_rtol = std::pow(10.0, -m->refine_phase1_digits);
        j.stall_window = 4 * std::max(m->g.n[0], std::max(m->g.n[1], m->g.nd == 3 ? m->g.n[2] : 0));
        j.state = JOB_PHASE1;
        j.done = 0;
        next_chunk(j);
        return FI_OK;
    }
    return job_outer(m, j);
}

// after the pending chunk of sweeps has run
int job_chunk_enqueue(fi_model_s* m, RefineJob& j)
{
    j.done += j.chunk;
    j.last = j.chunk;
    j.chunk = 0;
    if (j.state == JOB_INNER && j.mode == 1) {
        // lagged PT: the end of every ncheck sweeps is an outer step
        j.inner_total += j.done;
        j.since_eta += j.done;
        if (j.cols_clean) {
            // corrections are non-zero only on the column ranges: add them
            // there and clear the FP32 buffers the next chunk starts from in
            // the same pass (lag_cols: 3D, all ten fields)
            LagApply la;
            std::memset(&la, 0, sizeof(la));
            for (int id = FI_F_P; id <= FI_F_VZ; ++id) {
                la.x[id - FI_F_P] = m->buf(id);
                la.d[id - FI_F_P] = m->rf->cur(id);
            }
            la.z[0] = m->rf->d[FI_F_P][0];
            for (int q = 0; q < 3; ++q) la.z[1 + q] = m->rf->d[FI_F_VX + q][0];
            hipLaunchKernelGGL(k_lag_apply, plane_grid(m), dim3(64, 4), 0, m->stream, la, m->crange, col_geom(m));
        } else {
            for (int id = FI_F_P; id <= FI_F_VZ; ++id)
                if (m->f[id])
                    hipLaunchKernelGGL(k_add_correction, ew_grid(m->g.elems), dim3(256), 0, m->stream,
                                       m->buf(id) - m->g.off, m->rf->cur(id) - m->g.off, m->g.elems);
        }
        m->state_written();
        ++j.outer;
        j.pending_outer = true;
        return job_outer_lag_enqueue(m, j);
    }
    j.pending_outer = false;
    return inner_norm_enqueue(m);
}

int job_chunk_finish(fi_model_s* m, RefineJob& j)
{
    if (j.pending_outer) return job_outer_lag_finish(m, j);
    double rin = 0.0;
    inner_norm_read(m, &rin);
    int rc = FI_OK;
    if (j.state == JOB_PHASE1) {
        if (!(rin == rin) || rin > 3.4e38) {
            j.inner_total += j.done;
            j.phase1 = j.done;
            m->rf->full = false;
            j.state = JOB_DONE;
            return fail(FI_ERR_NONFINITE, "FP32 nonlinear phase diverged after %d iterations", j.done);
        }
        bool stop = j.done >= j.max_inner;
        if (rin > j.first) {
            // still in the start-up transient: measure from the peak
            j.first = j.best = rin;
            j.stall = 0;
        } else if (rin <= j.p1_rtol * j.first) {
            stop = true;
        } else if (rin < 0.5 * j.best) {
            j.best = rin;
            j.stall = 0;
        } else if ((j.stall += j.last) >= j.stall_window) {
            stop = true;
        }
        if (!stop) {
            next_chunk(j);
            return FI_OK;
        }
        phase1_end(m);
        j.phase1 = j.done;
        j.inner_total += j.done;
        if (j.trace) std::fprintf(stderr, "[fastice] refined: FP32 nonlinear phase %d iterations\n", j.done);
        return job_outer(m, j);
    }
    // correction solve
    if (j.trace)
        std::fprintf(stderr, "[fastice] refined: outer %d correction sweeps %d inner residual %.3e (%.3e of the defect)\n",
                     j.outer, j.done, rin, rin / j.r0);
    if (!(rin == rin) || rin > 3.4e38) {
        j.inner_total += j.done;
        j.state = JOB_DONE;
        return fail(FI_ERR_NONFINITE, "FP32 correction diverged at outer step %d", j.outer);
    }
    if (!(rin <= j.inner_rtol * j.r0) && j.done < j.max_inner) {
        next_chunk(j);
        return FI_OK;
    }
    j.inner_total += j.done;
    Refine* r = m->rf;
    if (j.aa.depth > 0) {
        rc = anderson_apply(m, j.aa, j.aa_restart);
        if (rc) return rc;
    } else {
        for (int id = FI_F_P; id <= FI_F_VZ; ++id)
            if (m->f[id])
                hipLaunchKernelGGL(k_add_correction, ew_grid(m->g.elems), dim3(256), 0, m->stream, m->buf(id) - m->g.off,
                                   r->cur(id) - m->g.off, m->g.elems);
    }
    m->state_written();
    ++j.outer;
    return job_outer(m, j);
}

int job_after_chunk(fi_model_s* m, RefineJob& j)
{
    int rc = job_chunk_enqueue(m, j);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return job_chunk_finish(m, j);
}

} // namespace fi

using namespace fi;

extern "C" {

int fi_refined_solve(fi_model_t m, int max_outer, int max_inner, int ncheck, double inner_rtol, double tol,
                     int counts_out[3], double norms_inf[4])
{
    if (!m || !norms_inf) return fail(FI_ERR_ARG, "null argument");
    if (counts_out) counts_out[0] = counts_out[1] = counts_out[2] = 0;
    RefineJob j;
    int rc = job_begin(m, j, max_outer, max_inner, ncheck, inner_rtol, tol);
    while (!rc && j.state != JOB_DONE) {
        for (int it = 0; it < j.chunk; ++it) iterate32(m, nullptr);
        rc = job_after_chunk(m, j);
    }
    if (counts_out) {
        counts_out[0] = j.outer;
        counts_out[1] = j.inner_total;
        counts_out[2] = j.phase1;
    }
    for (int q = 0; q < 4; ++q) norms_inf[q] = j.norms[q];
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    return FI_OK;
}

int fi_refined_time_inner(fi_model_t m, int n_iter, double out_ms[4])
{
    if (!m || !out_ms || n_iter < 1) return fail(FI_ERR_ARG, "bad arguments");
    int rc = ensure_refine(m);
    if (rc) return rc;
    rc = prepare(m, false);
    if (rc) return rc;
    sync_masks(m);
    hipEvent_t ev[6];
    for (int q = 0; q < 6; ++q) HIPCHK(hipEventCreate(&ev[q]));
    double acc[4] = {0, 0, 0, 0};
    for (int it = 0; it < n_iter; ++it) {
        iterate32(m, ev);
        HIPCHK(hipEventSynchronize(ev[5]));
        float t[5];
        for (int q = 0; q < 5; ++q) HIPCHK(hipEventElapsedTime(&t[q], ev[q], ev[q + 1]));
        acc[0] += t[0] + t[1] + t[2] + t[3] + t[4];
        acc[1] += t[0];
        acc[2] += t[1] + t[3];
        acc[3] += t[2] + t[4];
    }
    for (int q = 0; q < 6; ++q) (void)hipEventDestroy(ev[q]);
    for (int q = 0; q < 4; ++q) out_ms[q] = acc[q] / n_iter;
    return FI_OK;
}

int fi_refined_advance_inner(fi_model_t m, int n_iter)
{
    if (!m || n_iter < 0) return fail(FI_ERR_ARG, "bad arguments");
    int rc = ensure_refine(m);
    if (rc) return rc;
    rc = prepare(m, false);
    if (rc) return rc;
    sync_masks(m);
    for (int it = 0; it < n_iter; ++it) iterate32(m, nullptr);
    HIPCHK(hipGetLastError());
    return FI_OK;
}

} // extern "C"
