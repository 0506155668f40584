// heat.hip -- PT heat diffusion (src/Models/Diffusion/Heat): the model
// handle behind fi_heat_t, one fused kernel per PT iteration and the C ABI.
//
// One iteration of the reference is two launches with boundary conditions
// after each (Heat.jl:42-58): update_q! then the Flux conditions (the face
// value of q), update_T! then the Value conditions (the ghost of T).  Here
// one kernel does both updates (in 3D marching up z, k_heat_iter_z).  Each
// thread owns a cell and the three faces below it (x, y, z); it recomputes
// the new flux of its three upper faces from the same inputs their owners
// use, so every flux is evaluated
// by the same expression wherever it is needed, and writes the new flux of
// its own faces and the new temperature.  Flux conditions are applied in the
// kernel (the face takes the prescribed value before the divergence uses
// it), Value conditions by the boundary kernels afterwards.  T and q are
// double buffered; T_o is read only.  Traffic per cell: T, T_o, q read, T
// and q written (9 arrays in 3D), against 13 for the two launches.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "model.h"

using namespace fi;

struct fi_heat_s {
    fi_heat_desc desc;
    Geo g;
    hipStream_t stream = nullptr;
    double* base[2][4] = {};  // [buffer][T, qx, qy, qz] allocations
    double* f[2][4] = {};     // index (0,0,0) of each
    double* to_base = nullptr;
    double* to = nullptr;     // T_o
    int cur = 0;              // buffer holding the current T and q
    bool shadow_dirty = true; // the other buffer must first copy positions no kernel writes
    unsigned long long* d_norm = nullptr;
    unsigned long long* h_norm = nullptr;
    double* d_stage = nullptr;
    size_t stage_elems = 0;
};

namespace fi {

struct HeatArgs {
    const double* T;
    const double* To;
    const double* q[3];
    double* Tn;
    double* qn[3];
    int n[3];
    long st[3];
    double dx[3];
    double lam, dtq, dtT, dt;
    int flux_lo[3], flux_hi[3];
    double v_lo[3], v_hi[3];
};

// new flux of dimension d at face index idx (linear index c): the prescribed
// value on a Flux face, else q -= (q + lam * dv(T) / dx) * dtq (kernels.jl:5-17)
__device__ __forceinline__ double heat_flux(const HeatArgs& a, int d, long c, int idx)
{
    if (idx == 1 && a.flux_lo[d]) return a.v_lo[d];
    if (idx == a.n[d] + 1 && a.flux_hi[d]) return a.v_hi[d];
    const double q = a.q[d][c];
    const double dT = a.lam * (a.T[c] - a.T[c - a.st[d]]) / a.dx[d];
    return q - (q + dT) * a.dtq;
}

template <int ND>
__global__ __launch_bounds__(256) void k_heat_iter(HeatArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = ND == 3 ? (int)blockIdx.z + 1 : 0;
    if (i > a.n[0] + 1 || j > a.n[1] + 1) return;
    const int id[3] = {i, j, k};
    const long c = (long)i + a.st[1] * j + a.st[2] * k;
    bool cell = true;
    int beyond = 0; // dimensions in which this index is past the last cell
    for (int d = 0; d < ND; ++d)
        if (id[d] > a.n[d]) {
            cell = false;
            ++beyond;
        }
    // own faces: face d of this index exists when no other dimension is past its cells
    for (int d = 0; d < ND; ++d) {
        const int other = beyond - (id[d] > a.n[d] ? 1 : 0);
        if (other == 0) a.qn[d][c] = heat_flux(a, d, c, id[d]);
    }
    if (!cell) return;
    // update_T! (kernels.jl:19-29)
    double divq = (heat_flux(a, 0, c + 1, i + 1) - heat_flux(a, 0, c, i)) / a.dx[0];
    divq = divq + (heat_flux(a, 1, c + a.st[1], j + 1) - heat_flux(a, 1, c, j)) / a.dx[1];
    if (ND == 3) divq = divq + (heat_flux(a, 2, c + a.st[2], k + 1) - heat_flux(a, 2, c, k)) / a.dx[2];
    const double dTdt = (a.T[c] - a.To[c]) / a.dt;
    a.Tn[c] = a.T[c] - (dTdt + divq) * a.dtT;
}

// 3D: the same iteration marching up a chunk of HEAT_KC z planes per thread
// column.  T of the planes below, at and above k and the new flux of the z
// face below k stay in registers, so each plane of T and q_z is read once per
// column; every flux is still computed from the same values by the same
// expression as in k_heat_iter.
constexpr int HEAT_KC = 32;

__device__ __forceinline__ double heat_zflux(const HeatArgs& a, long c, int idx, double Tup, double Tlo)
{
    if (idx == 1 && a.flux_lo[2]) return a.v_lo[2];
    if (idx == a.n[2] + 1 && a.flux_hi[2]) return a.v_hi[2];
    const double q = a.q[2][c];
    const double dT = a.lam * (Tup - Tlo) / a.dx[2];
    return q - (q + dT) * a.dtq;
}

__global__ __launch_bounds__(256) void k_heat_iter_z(HeatArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > a.n[0] + 1 || j > a.n[1] + 1) return;
    const int k0 = (int)blockIdx.z * HEAT_KC + 1;
    const int k1 = min(k0 + HEAT_KC - 1, a.n[2] + 1);
    const long sy = a.st[1], sz = a.st[2];
    long c = (long)i + sy * j + sz * k0;
    const bool xin = i <= a.n[0], yin = j <= a.n[1];
    if (!(xin && yin)) {
        // the x faces at i = nx + 1 or the y faces at j = ny + 1 of this column
        for (int k = k0; k <= k1 && k <= a.n[2]; ++k, c += sz) {
            if (yin) a.qn[0][c] = heat_flux(a, 0, c, i);
            if (xin) a.qn[1][c] = heat_flux(a, 1, c, j);
        }
        return;
    }
    double Tc = a.T[c];
    double qz_lo = heat_zflux(a, c, k0, Tc, a.T[c - sz]);
    for (int k = k0; k <= k1; ++k, c += sz) {
        a.qn[2][c] = qz_lo;
        if (k > a.n[2]) break; // the top z face of the column
        const double qx_lo = heat_flux(a, 0, c, i);
        const double qy_lo = heat_flux(a, 1, c, j);
        a.qn[0][c] = qx_lo;
        a.qn[1][c] = qy_lo;
        const double Tp = a.T[c + sz];
        const double qz_hi = heat_zflux(a, c + sz, k + 1, Tp, Tc);
        double divq = (heat_flux(a, 0, c + 1, i + 1) - qx_lo) / a.dx[0];
        divq = divq + (heat_flux(a, 1, c + sy, j + 1) - qy_lo) / a.dx[1];
        divq = divq + (qz_hi - qz_lo) / a.dx[2];
        const double dTdt = (Tc - a.To[c]) / a.dt;
        a.Tn[c] = Tc - (dTdt + divq) * a.dtT;
        Tc = Tp;
        qz_lo = qz_hi;
    }
}

// Inf-norms of the T and q residuals as |x| bit patterns (NaN above +Inf)
template <int ND>
__global__ __launch_bounds__(256) void k_heat_residual(HeatArgs a, unsigned long long* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = ND == 3 ? (int)blockIdx.z + 1 : 0;
    unsigned long long rT = 0ull, rq = 0ull;
    if (i <= a.n[0] + 1 && j <= a.n[1] + 1) {
        const int id[3] = {i, j, k};
        const long c = (long)i + a.st[1] * j + a.st[2] * k;
        int beyond = 0;
        for (int d = 0; d < ND; ++d) beyond += id[d] > a.n[d];
        for (int d = 0; d < ND; ++d) {
            if (beyond - (id[d] > a.n[d] ? 1 : 0) != 0) continue;
            if ((id[d] == 1 && a.flux_lo[d]) || (id[d] == a.n[d] + 1 && a.flux_hi[d])) continue;
            const double r = a.q[d][c] + a.lam * (a.T[c] - a.T[c - a.st[d]]) / a.dx[d];
            const unsigned long long b = (unsigned long long)__double_as_longlong(r) & 0x7fffffffffffffffull;
            rq = b > rq ? b : rq;
        }
        if (beyond == 0) {
            double divq = (a.q[0][c + 1] - a.q[0][c]) / a.dx[0];
            divq = divq + (a.q[1][c + a.st[1]] - a.q[1][c]) / a.dx[1];
            if (ND == 3) divq = divq + (a.q[2][c + a.st[2]] - a.q[2][c]) / a.dx[2];
            const double r = (a.T[c] - a.To[c]) / a.dt + divq;
            rT = (unsigned long long)__double_as_longlong(r) & 0x7fffffffffffffffull;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(rT, o), y = __shfl_xor(rq, o);
        rT = x > rT ? x : rT;
        rq = y > rq ? y : rq;
    }
    __shared__ unsigned long long red[2][4];
    const int lane = threadIdx.x + blockDim.x * threadIdx.y;
    if ((lane & 63) == 0) {
        red[0][lane >> 6] = rT;
        red[1][lane >> 6] = rq;
    }
    __syncthreads();
    if (lane == 0) {
        for (int w = 1; w < 4; ++w) {
            rT = red[0][w] > rT ? red[0][w] : rT;
            rq = red[1][w] > rq ? red[1][w] : rq;
        }
        atomicMax(out, rT);
        atomicMax(out + 1, rq);
    }
}

__global__ void k_heat_zero2(unsigned long long* p)
{
    if (threadIdx.x < 2) p[threadIdx.x] = 0ull;
}

static dim3 heat_grid(const Geo& g) { return dim3((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.nd == 3 ? g.n[2] + 1 : 1); }
static dim3 heat_grid_z(const Geo& g) { return dim3((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, (g.n[2] + HEAT_KC) / HEAT_KC); }

static void launch_heat_iter(fi_heat_s* h, const HeatArgs& a)
{
    if (h->g.nd == 3) hipLaunchKernelGGL(k_heat_iter_z, heat_grid_z(h->g), dim3(64, 4, 1), 0, h->stream, a);
    else hipLaunchKernelGGL(k_heat_iter<2>, heat_grid(h->g), dim3(64, 4, 1), 0, h->stream, a);
}

static int heat_loc(int nd, int id)
{
    (void)nd;
    return id == FI_HF_QX ? 1 : (id == FI_HF_QY ? 2 : (id == FI_HF_QZ ? 4 : 0));
}

static HeatArgs heat_args(fi_heat_s* h, double dt)
{
    const fi_heat_desc& d = h->desc;
    const Geo& g = h->g;
    HeatArgs a;
    std::memset(&a, 0, sizeof(a));
    const int p = h->cur;
    a.T = h->f[p][0];
    a.To = h->to;
    a.Tn = h->f[1 - p][0];
    for (int q = 0; q < 3; ++q) {
        a.q[q] = q < g.nd ? h->f[p][1 + q] : nullptr;
        a.qn[q] = q < g.nd ? h->f[1 - p][1 + q] : nullptr;
        a.n[q] = g.n[q];
        a.dx[q] = g.dx[q];
        a.flux_lo[q] = q < g.nd && d.bc_kind[q][0] == FI_HEAT_BC_FLUX;
        a.flux_hi[q] = q < g.nd && d.bc_kind[q][1] == FI_HEAT_BC_FLUX;
        a.v_lo[q] = d.bc_value[q][0];
        a.v_hi[q] = d.bc_value[q][1];
    }
    a.st[0] = 1;
    a.st[1] = g.sy;
    a.st[2] = g.sz;
    a.lam = d.lambda_rho_cp;
    a.dtq = d.dtau_q;
    a.dtT = d.dtau_T;
    a.dt = dt;
    return a;
}

// the Value conditions on the temperature of buffer p (DirichletBC{HalfCell}:
// ghost = 2 v - interior), x then y then z
static void heat_value_bc(fi_heat_s* h, int p)
{
    const Geo& g = h->g;
    BcBatch b[3];
    for (int d = 0; d < g.nd; ++d) {
        b[d].nop = 0;
        b[d].d = d;
        for (int s = 0; s < 2; ++s)
            if (h->desc.bc_kind[d][s] == FI_HEAT_BC_VALUE)
                add_side_op(b[d], h->f[p][0], 0, 0, h->desc.bc_value[d][s], s);
    }
    launch_bc_dims(g, b, h->stream);
}

static int heat_prepare(fi_heat_s* h)
{
    if (!h->shadow_dirty) return FI_OK;
    const size_t bytes = h->g.elems * sizeof(double);
    for (int q = 0; q <= h->g.nd; ++q)
        HIPCHK(hipMemcpyAsync(h->base[1 - h->cur][q], h->base[h->cur][q], bytes, hipMemcpyDeviceToDevice, h->stream));
    h->shadow_dirty = false;
    return FI_OK;
}

static int heat_iterate(fi_heat_s* h, double dt)
{
    launch_heat_iter(h, heat_args(h, dt));
    h->cur ^= 1;
    heat_value_bc(h, h->cur);
    return FI_OK;
}

static int heat_norms(fi_heat_s* h, double dt, double out[2])
{
    HeatArgs a = heat_args(h, dt);
    hipLaunchKernelGGL(k_heat_zero2, dim3(1), dim3(64), 0, h->stream, h->d_norm);
    if (h->g.nd == 3) hipLaunchKernelGGL(k_heat_residual<3>, heat_grid(h->g), dim3(64, 4, 1), 0, h->stream, a, h->d_norm);
    else hipLaunchKernelGGL(k_heat_residual<2>, heat_grid(h->g), dim3(64, 4, 1), 0, h->stream, a, h->d_norm);
    HIPCHK(hipMemcpyAsync(h->h_norm, h->d_norm, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int q = 0; q < 2; ++q) out[q] = bits_to_double(h->h_norm[q]);
    return FI_OK;
}

static void heat_free(fi_heat_s* h)
{
    for (int b = 0; b < 2; ++b)
        for (int q = 0; q < 4; ++q)
            if (h->base[b][q]) (void)hipFree(h->base[b][q]);
    if (h->to_base) (void)hipFree(h->to_base);
    if (h->d_norm) (void)hipFree(h->d_norm);
    if (h->h_norm) (void)hipHostFree(h->h_norm);
    if (h->d_stage) (void)hipFree(h->d_stage);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

} // namespace fi

extern "C" {

int fi_heat_create(const fi_heat_desc* desc, fi_heat_t* out)
{
    if (!out) return fail(FI_ERR_ARG, "null output handle");
    *out = nullptr;
    if (!desc) return fail(FI_ERR_ARG, "null descriptor");
    if (desc->ndims != 2 && desc->ndims != 3) return fail(FI_ERR_ARG, "ndims must be 2 or 3 (got %d)", desc->ndims);
    for (int d = 0; d < desc->ndims; ++d) {
        if (desc->n[d] < 1) return fail(FI_ERR_ARG, "n[%d] must be >= 1", d);
        if (!(desc->extent[d] > 0.0)) return fail(FI_ERR_ARG, "extent[%d] must be > 0", d);
        for (int s = 0; s < 2; ++s)
            if (desc->bc_kind[d][s] < FI_HEAT_BC_NONE || desc->bc_kind[d][s] > FI_HEAT_BC_FLUX)
                return fail(FI_ERR_ARG, "bad boundary condition kind %d", desc->bc_kind[d][s]);
    }
    fi_heat_s* h = new (std::nothrow) fi_heat_s();
    if (!h) return fail(FI_ERR_OOM, "host allocation failed");
    h->desc = *desc;
    h->g = geo_for(desc->ndims, desc->n, desc->extent);
    const size_t bytes = h->g.elems * sizeof(double);
    bool ok = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess;
    for (int b = 0; b < 2 && ok; ++b)
        for (int q = 0; q <= h->g.nd && ok; ++q) {
            ok = hipMalloc(&h->base[b][q], bytes) == hipSuccess &&
                 hipMemsetAsync(h->base[b][q], 0, bytes, h->stream) == hipSuccess;
            if (ok) h->f[b][q] = h->base[b][q] + h->g.off;
        }
    ok = ok && hipMalloc(&h->to_base, bytes) == hipSuccess && hipMemsetAsync(h->to_base, 0, bytes, h->stream) == hipSuccess;
    ok = ok && hipMalloc(&h->d_norm, 2 * sizeof(unsigned long long)) == hipSuccess &&
         hipHostMalloc(&h->h_norm, 2 * sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        heat_free(h);
        return fail(FI_ERR_OOM, "device allocation failed (%zu bytes per field)", bytes);
    }
    h->to = h->to_base + h->g.off;
    *out = h;
    return FI_OK;
}

int fi_heat_destroy(fi_heat_t h)
{
    if (!h) return FI_OK;
    (void)hipStreamSynchronize(h->stream);
    heat_free(h);
    return FI_OK;
}

static int heat_copy(fi_heat_t h, int id, double* host, size_t count, int layout, bool up)
{
    if (!h) return fail(FI_ERR_ARG, "null model");
    if (!host) return fail(FI_ERR_ARG, "null host pointer");
    if (id < 0 || id >= FI_HF_COUNT || (id >= FI_HF_QX + h->g.nd))
        return fail(FI_ERR_ARG, "field %d not present in this model", id);
    if (layout != FI_LAYOUT_PARENT && layout != FI_LAYOUT_INTERIOR) return fail(FI_ERR_ARG, "bad layout %d", layout);
    const Geo& g = h->g;
    const int L = heat_loc(g.nd, id);
    size_t need = 1;
    for (int d = 0; d < g.nd; ++d) need *= (size_t)(layout == FI_LAYOUT_PARENT ? g.n[d] + 3 : g.n[d] + ((L >> d) & 1));
    if (count != need) return fail(FI_ERR_ARG, "field %d needs %zu doubles in this layout, got %zu", id, need, count);
    double* dev = id == FI_HF_T_OLD ? h->to : h->f[h->cur][id == FI_HF_T ? 0 : id - FI_HF_QX + 1];
    if (up && id != FI_HF_T_OLD) h->shadow_dirty = true;
    return field_transfer(g, dev, L, host, count, layout, up, h->stream, &h->d_stage, &h->stage_elems);
}

int fi_heat_field_upload(fi_heat_t h, int id, const double* host, size_t count, int layout)
{
    return heat_copy(h, id, const_cast<double*>(host), count, layout, true);
}

int fi_heat_field_download(fi_heat_t h, int id, double* host, size_t count, int layout)
{
    return heat_copy(h, id, host, count, layout, false);
}

int fi_heat_apply_bc(fi_heat_t h)
{
    if (!h) return fail(FI_ERR_ARG, "null model");
    heat_value_bc(h, h->cur);
    h->shadow_dirty = true;
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_heat_advance_iterations(fi_heat_t h, double dt, int n_iter)
{
    if (!h) return fail(FI_ERR_ARG, "null model");
    if (n_iter < 0 || !(dt > 0.0)) return fail(FI_ERR_ARG, "need n_iter >= 0 and dt > 0");
    int rc = heat_prepare(h);
    if (rc) return rc;
    for (int it = 0; it < n_iter; ++it) heat_iterate(h, dt);
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_heat_residual_norms(fi_heat_t h, double dt, double norms_inf[2])
{
    if (!h || !norms_inf) return fail(FI_ERR_ARG, "null argument");
    if (!(dt > 0.0)) return fail(FI_ERR_ARG, "dt must be > 0");
    return heat_norms(h, dt, norms_inf);
}

int fi_heat_step(fi_heat_t h, double dt, int max_iter, int ncheck, double tol, int* iters_out, double norms_inf[2])
{
    if (!h || !norms_inf) return fail(FI_ERR_ARG, "null argument");
    if (max_iter < 0 || ncheck < 1 || !(dt > 0.0)) return fail(FI_ERR_ARG, "bad step parameters");
    int rc = heat_prepare(h);
    if (rc) return rc;
    int it = 0;
    for (;;) {
        rc = heat_norms(h, dt, norms_inf);
        if (rc) return rc;
        for (int q = 0; q < 2; ++q)
            if (!(norms_inf[q] == norms_inf[q]) || norms_inf[q] > 1.79e308) {
                if (iters_out) *iters_out = it;
                return fail(FI_ERR_NONFINITE, "heat diffusion failed at iteration %d: non-finite residual", it);
            }
        if ((norms_inf[0] < tol && norms_inf[1] < tol) || it >= max_iter) break;
        const int nrun = std::min(ncheck, max_iter - it);
        for (int q = 0; q < nrun; ++q) heat_iterate(h, dt);
        it += nrun;
    }
    // T_o <- T
    HIPCHK(hipMemcpyAsync(h->to_base, h->base[h->cur][0], h->g.elems * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (iters_out) *iters_out = it;
    return FI_OK;
}

int fi_heat_time_iterations(fi_heat_t h, double dt, int n_iter, double out_ms[2])
{
    if (!h || !out_ms || n_iter < 1 || !(dt > 0.0)) return fail(FI_ERR_ARG, "bad arguments");
    int rc = heat_prepare(h);
    if (rc) return rc;
    // back to back, as a solve runs them: events around every kernel, one
    // synchronisation at the end
    std::vector<hipEvent_t> ev(2 * (size_t)n_iter + 1);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(ev[0], h->stream));
    for (int it = 0; it < n_iter; ++it) {
        if (it > 0) HIPCHK(hipEventRecord(ev[2 * it], h->stream));
        launch_heat_iter(h, heat_args(h, dt));
        HIPCHK(hipEventRecord(ev[2 * it + 1], h->stream));
        h->cur ^= 1;
        heat_value_bc(h, h->cur);
    }
    HIPCHK(hipEventRecord(ev[2 * (size_t)n_iter], h->stream));
    HIPCHK(hipEventSynchronize(ev[2 * (size_t)n_iter]));
    double kernel = 0.0;
    for (int it = 0; it < n_iter; ++it) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, ev[2 * it], ev[2 * it + 1]));
        kernel += t;
    }
    float total = 0.0f;
    HIPCHK(hipEventElapsedTime(&total, ev[0], ev[2 * (size_t)n_iter]));
    for (auto& e : ev) (void)hipEventDestroy(e);
    out_ms[0] = total / n_iter;
    out_ms[1] = kernel / n_iter;
    return FI_OK;
}

} // extern "C"
