// api.hip -- C ABI of libfasticehip.so: model handles, device fields, the PT
// iteration driver, residual norms, and the level-set set-up calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include <algorithm>
#include <mutex>
#include <vector>

#include "model.h"

using namespace fi;

namespace fi {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int field_loc(int nd, int id)
{
    switch (id) {
    case FI_F_P: case FI_F_TXX: case FI_F_TYY: case FI_F_TZZ: case FI_F_ETA: case FI_F_ETA_NEXT: case FI_F_RP:
    case FI_F_RHEO: return L_CCC;
    case FI_F_TXY: return L_VVC;
    case FI_F_TXZ: return L_VCV;
    case FI_F_TYZ: return L_CVV;
    case FI_F_VX: case FI_F_RVX: case FI_F_GX: return L_VCC;
    case FI_F_VY: case FI_F_RVY: case FI_F_GY: return L_CVC;
    case FI_F_VZ: case FI_F_RVZ: case FI_F_GZ: return L_CCV;
    case FI_F_PSI_NA: case FI_F_PSI_NS: return (nd == 3) ? L_VVV : L_VVC;
    default: break;
    }
    if (id >= FI_F_OM_NA0 && id < FI_F_COUNT) {
        int q = (id - FI_F_OM_NA0) % 8;
        if (nd == 3) {
            const int L3[8] = {L_CCC, L_VVV, L_VCC, L_CVC, L_CCV, L_VVC, L_VCV, L_CVV};
            return L3[q];
        }
        const int L2[4] = {L_CCC, L_VVC, L_VCC, L_CVC};
        return q < 4 ? L2[q] : -1;
    }
    return -1;
}

bool field_used(const fi_model_desc& d, int id)
{
    if (id == FI_F_TZZ || id == FI_F_TXZ || id == FI_F_TYZ || id == FI_F_VZ || id == FI_F_RVZ) return d.ndims == 3;
    if (id == FI_F_GX || id == FI_F_GY) return d.gravity_is_field != 0;
    if (id == FI_F_GZ) return d.ndims == 3 && d.gravity_is_field != 0;
    if (id == FI_F_RHEO)
        return d.rheology.kind == FI_RHEO_LINEAR_FIELD || (d.rheology.kind == FI_RHEO_POWERLAW && d.rheology.A_is_field);
    if (id == FI_F_PSI_NA || id == FI_F_PSI_NS) return d.immersed_boundary != 0;
    if (id >= FI_F_OM_NA0) {
        if (!d.immersed_boundary) return false;
        int q = (id - FI_F_OM_NA0) % 8;
        return d.ndims == 3 || q < 4;
    }
    return true;
}

// in_place: kernels read and write the current buffers (residuals, compute_tau);
// otherwise they read the current and write the other buffer.
StokesArgs<double> make_args(fi_model_s* m, bool in_place)
{
    StokesArgs<double> a;
    std::memset(&a, 0, sizeof(a));
    const fi_model_desc& d = m->desc;
    a.P = m->buf(FI_F_P);
    a.txx = m->buf(FI_F_TXX);
    a.tyy = m->buf(FI_F_TYY);
    a.tzz = m->buf(FI_F_TZZ);
    a.txy = m->buf(FI_F_TXY);
    a.txz = m->buf(FI_F_TXZ);
    a.tyz = m->buf(FI_F_TYZ);
    a.Vx = m->buf(FI_F_VX);
    a.Vy = m->buf(FI_F_VY);
    a.Vz = m->buf(FI_F_VZ);
    auto out = [&](int id) { return in_place ? m->buf(id) : m->alt(id); };
    a.Pn = out(FI_F_P);
    a.txxn = out(FI_F_TXX);
    a.tyyn = out(FI_F_TYY);
    a.tzzn = out(FI_F_TZZ);
    a.txyn = out(FI_F_TXY);
    a.txzn = out(FI_F_TXZ);
    a.tyzn = out(FI_F_TYZ);
    a.Vxn = out(FI_F_VX);
    a.Vyn = out(FI_F_VY);
    a.Vzn = out(FI_F_VZ);
    a.eta = m->eta();
    a.eta_next = m->eta_next();
    a.gx = m->f[FI_F_GX];
    a.gy = m->f[FI_F_GY];
    a.gz = m->f[FI_F_GZ];
    a.rh = m->f[FI_F_RHEO];
    a.rP = m->f[FI_F_RP];
    a.rVx = m->f[FI_F_RVX];
    a.rVy = m->f[FI_F_RVY];
    a.rVz = m->f[FI_F_RVZ];
    for (int q = 0; q < 8; ++q) {
        a.na[q] = m->f[FI_F_OM_NA0 + q];
        a.ns[q] = m->f[FI_F_OM_NS0 + q];
    }
    a.flags = m->flags;
    a.crange = m->crange;
    a.erange = m->erange;
    a.tmarch = m->tmarch;
    a.nx = m->g.n[0];
    a.ny = m->g.n[1];
    a.nz = m->g.n[2];
    a.sy = m->g.sy;
    a.sz = m->g.sz;
    a.ix = m->g.idx[0];
    a.iy = m->g.idx[1];
    a.iz = m->g.idx[2];
    a.dPr = d.dtau_P;
    for (int q = 0; q < 6; ++q) a.dtau[q] = d.dtau_tau[q];
    for (int q = 0; q < 3; ++q) {
        a.dV[q] = d.dtau_V[q];
        a.grav[q] = d.gravity[q];
    }
    a.grav_field = d.gravity_is_field;
    a.rh_kind = d.rheology.kind;
    a.rh_has_reg = d.rheology.has_reg;
    a.rh_relax = d.rheology.relax;
    a.rh_eta = d.rheology.eta;
    a.rh_A = d.rheology.A;
    a.rh_n = d.rheology.n;
    a.rh_inv_eta_reg = d.rheology.has_reg ? 1.0 / d.rheology.eta_reg : 0.0;
    a.rh_rate = d.rheology.rate;
    if (!(d.rheology.kind == FI_RHEO_POWERLAW && d.rheology.A_is_field) && d.rheology.kind != FI_RHEO_LINEAR_FIELD)
        a.rh = nullptr;
    return a;
}

BcBatch stress_batch(fi_model_s* m, int d)
{
    return stress_batch_g(m->desc, d, [m](int id) { return (void*)m->buf(id); }, false, 0);
}

BcBatch velocity_batch(fi_model_s* m, int d, bool with_eta)
{
    BcBatch b = velocity_batch_g(m->desc, d, [m](int id) { return (void*)m->buf(id); }, false, 0);
    if (with_eta) add_op(b, m->eta_next(), L_CCC, 1, 0.0);
    return b;
}

BcBatch residual_batch(fi_model_s* m, int d)
{
    return residual_batch_g(m->desc, d, [m](int id) { return (void*)m->f[id]; }, 0);
}

BcBatch neumann_batch(double* p, int L, int d)
{
    BcBatch b;
    b.nop = 0;
    b.d = d;
    add_op(b, p, L, 1, 0.0);
    return b;
}

void run_bc(fi_model_s* m, int group, bool skip)
{
    BcBatch b[3];
    for (int d = 0; d < m->g.nd; ++d) {
        if (group == FI_BCGROUP_STRESS) b[d] = stress_batch(m, d);
        else if (group == FI_BCGROUP_VELOCITY) b[d] = velocity_batch(m, d, true);
        else if (group == FI_BCGROUP_ETA) b[d] = neumann_batch(m->eta(), L_CCC, d);
        else b[d] = residual_batch(m, d);
    }
    launch_bc_dims(m->g, b, m->stream, skip ? m->crange : nullptr, skip ? m->empty_faces : 0u);
}

int prepare(fi_model_s* m, bool sync_shadow)
{
    if (sync_shadow && m->shadow_dirty) {
        // positions no kernel or boundary condition writes must hold the same
        // value in both buffers
        const size_t bytes = m->g.elems * sizeof(double);
        for (int id = FI_F_P; id <= FI_F_VZ; ++id)
            if (m->f[id]) HIPCHK(hipMemcpyAsync(m->alt(id) - m->g.off, m->buf(id) - m->g.off, bytes,
                                                hipMemcpyDeviceToDevice, m->stream));
        m->zs[1 - m->sig_cur] = m->zs[m->sig_cur];
        m->zv[1 - m->v_cur] = m->zv[m->v_cur];
        m->gs[0] = m->gs[1] = false;
        m->shadow_dirty = false;
    }
    if (m->desc.immersed_boundary && m->flags_dirty) {
        const double* na[8];
        const double* ns[8];
        for (int q = 0; q < 8; ++q) {
            na[q] = m->f[FI_F_OM_NA0 + q];
            ns[q] = m->f[FI_F_OM_NS0 + q];
        }
        launch_ib_flags(m->g, na, ns, m->flags, m->stream);
        if (m->crange) {
            launch_col_ranges(m->g, m->flags, m->crange, m->erange, m->stream);
            std::vector<int2> r(col_ranges_count(m->g));
            HIPCHK(hipMemcpyAsync(r.data(), m->crange, r.size() * sizeof(int2), hipMemcpyDeviceToHost, m->stream));
            HIPCHK(hipStreamSynchronize(m->stream));
            m->empty_faces = face_layers_empty(m->g, r.data());
        }
        // locations may have changed validity: no buffer is known to be clean
        m->resid_clean = false;
        m->zs[0] = m->zs[1] = m->zv[0] = m->zv[1] = false;
        m->zs32[0] = m->zs32[1] = m->zv32[0] = m->zv32[1] = false;
        m->gs[0] = m->gs[1] = m->gs32[0] = m->gs32[1] = false;
        m->flags_dirty = false;
    }
    return FI_OK;
}

// One PT iteration: stress update, stress BCs, velocity + viscosity update,
// velocity and viscosity BCs, eta swap (Isothermal.jl:91-112).  On 3D grids
// the interior goes through the fused sweep and the boundary shell through
// the pointwise kernels; both write the other state buffer.
void iterate(fi_model_s* m, hipEvent_t* ev)
{
    const bool ib = m->desc.immersed_boundary != 0;
    StokesArgs<double> a = make_args(m, false);
    const int so = 1 - m->sig_cur, vo = 1 - m->v_cur; // the buffers this iteration writes
    a.z_read = m->zv[m->v_cur];
    a.z_wsig = m->zs[so];
    a.z_wvel = m->zv[vo];
    const bool fused = m->use_fused && m->has_interior;
    const BoxList boxes = fused ? m->shell : full_box(m->g);
    // shell boxes in a face layer without a valid location: every point of
    // theirs returns at once (sigma_point3, velocity_point3), so they are
    // left out of the launches
    const BoxList sboxes = ib && a.z_wsig ? drop_empty_boxes(m->g, boxes, m->empty_faces) : boxes;
    const BoxList vboxes = ib && a.z_wvel && a.frozen_eta ? drop_empty_boxes(m->g, boxes, m->empty_faces) : boxes;
    if (ev) (void)hipEventRecord(ev[0], m->stream);
    if (fused) {
        FusedGeom fg = m->fg;
        fg.order = m->forder;
        launch_fused(m->g, a, fg, ib, m->ftx, m->fty, m->fminw, m->stream, m->fused_shell ? &sboxes : nullptr);
    }
    if (ev) (void)hipEventRecord(ev[1], m->stream);
    if (!(fused && m->fused_shell)) launch_sigma_boxes(m->g, a, sboxes, ib, m->stream);
    m->sig_cur ^= 1;
    if (ev) (void)hipEventRecord(ev[2], m->stream);
    // the sweep stored nothing at invalid locations (z_wsig): the ghosts the
    // last application on this buffer settled still hold their value
    const bool skip = ib && m->g.nd == 3 && m->crange && a.z_wsig && m->gs[so];
    run_bc(m, FI_BCGROUP_STRESS, skip);
    if (ib && m->g.nd == 3) m->gs[so] = a.z_wsig != 0;
    if (ev) (void)hipEventRecord(ev[3], m->stream);
    launch_velocity_boxes(m->g, a, vboxes, ib, m->stream);
    m->v_cur ^= 1;
    if (ev) (void)hipEventRecord(ev[4], m->stream);
    run_bc(m, FI_BCGROUP_VELOCITY);
    if (ev) (void)hipEventRecord(ev[5], m->stream);
    m->eta_cur ^= 1;
    // every invalid location of the written buffers now holds +0
    m->zs[so] = true;
    m->zv[vo] = true;
}

int advance(fi_model_s* m, int niter)
{
    if (m->batch && niter > 0) batch_touch(m); // the batch's tables assume the buffer parity
    int rc = prepare(m);
    if (rc) return rc;
    for (int it = 0; it < niter; ++it) iterate(m, nullptr);
    return FI_OK;
}

double bits_to_double(unsigned long long b)
{
    double v;
    std::memcpy(&v, &b, sizeof(v));
    return v;
}

// enqueue compute_residuals! and the Inf-norms into the pinned h_norm
int residual_norms_enqueue(fi_model_s* m, bool sync_shadow)
{
    const bool ib = m->desc.immersed_boundary != 0;
    if (sync_shadow) {
        int rc = prepare(m);
        if (rc) return rc;
    }
    StokesArgs<double> a = make_args(m, true);
    // positions outside the column plane ranges hold only invalid locations,
    // whose residual is 0: once the fields hold it there, neither the
    // residual kernel nor the norms visit them
    const bool cols = ib && m->g.nd == 3 && m->crange;
    a.skip_empty = cols && m->resid_clean;
    launch_residuals(m->g, a, ib, m->stream);
    m->resid_clean = cols;
    run_bc(m, FI_BCGROUP_RESIDUAL);
    const double* f[4] = {m->f[FI_F_RP], m->f[FI_F_RVX], m->f[FI_F_RVY], m->f[FI_F_RVZ]};
    const int L[4] = {L_CCC, L_VCC, L_CVC, L_CCV};
    const int nf = m->g.nd == 3 ? 4 : 3;
    launch_norms(m->g, f, L, nf, m->d_norm, m->stream, cols ? m->crange : nullptr);
    HIPCHK(hipMemcpyAsync(m->h_norm, m->d_norm, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, m->stream));
    return FI_OK;
}

// the norms of the last residual_norms_enqueue, once its stream has synchronised
void residual_norms_read(fi_model_s* m, double out[4])
{
    const int nf = m->g.nd == 3 ? 4 : 3;
    for (int q = 0; q < 4; ++q) out[q] = (q < nf) ? bits_to_double(m->h_norm[q]) : 0.0;
}

int residual_norms(fi_model_s* m, double out[4], bool sync_shadow)
{
    int rc = residual_norms_enqueue(m, sync_shadow);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    residual_norms_read(m, out);
    return FI_OK;
}

int check_desc(const fi_model_desc* d)
{
    if (!d) return fail(FI_ERR_ARG, "null descriptor");
    if (d->ndims != 2 && d->ndims != 3) return fail(FI_ERR_ARG, "ndims must be 2 or 3 (got %d)", d->ndims);
    for (int q = 0; q < d->ndims; ++q) {
        if (d->n[q] < 1) return fail(FI_ERR_ARG, "n[%d] must be >= 1 (got %d)", q, d->n[q]);
        if (!(d->extent[q] > 0.0)) return fail(FI_ERR_ARG, "extent[%d] must be > 0", q);
        for (int s = 0; s < 2; ++s)
            if (d->bc_kind[q][s] < FI_BC_TRACTION || d->bc_kind[q][s] > FI_BC_SLIP)
                return fail(FI_ERR_ARG, "bad boundary-condition kind on dim %d side %d", q, s);
    }
    if (d->rheology.kind < FI_RHEO_LINEAR_CONST || d->rheology.kind > FI_RHEO_POWERLAW)
        return fail(FI_ERR_ARG, "bad rheology kind %d", d->rheology.kind);
    if (d->rheology.relax < FI_RELAX_NONE || d->rheology.relax > FI_RELAX_LOG10)
        return fail(FI_ERR_ARG, "bad relaxation scale %d", d->rheology.relax);
    if (d->rheology.has_reg && !(d->rheology.eta_reg > 0.0)) return fail(FI_ERR_ARG, "eta_reg must be > 0");
    if (d->precision != FI_PREC_FP64 && d->precision != FI_PREC_FP32_REFINED)
        return fail(FI_ERR_ARG, "bad precision %d", d->precision);
    return FI_OK;
}

void free_model(fi_model_s* m)
{
    free_refine(m);
    for (int q = 0; q < FI_F_COUNT; ++q)
        if (m->base[q]) (void)hipFree(m->base[q]);
    if (m->flags_base) (void)hipFree(m->flags_base);
    if (m->crange) (void)hipFree(m->crange);
    if (m->erange) (void)hipFree(m->erange);
    if (m->tmarch) (void)hipFree(m->tmarch);
    if (m->d_norm) (void)hipFree(m->d_norm);
    if (m->h_norm) (void)hipHostFree(m->h_norm);
    if (m->d_dem) (void)hipFree(m->d_dem);
    if (m->d_stage) (void)hipFree(m->d_stage);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

size_t interior_count(const Geo& g, int L)
{
    size_t c = 1;
    for (int d = 0; d < g.nd; ++d) c *= (size_t)(g.n[d] + ((L >> d) & 1));
    return c;
}

size_t parent_count(const Geo& g)
{
    size_t c = 1;
    for (int d = 0; d < g.nd; ++d) c *= (size_t)(g.n[d] + 3);
    return c;
}

Geo geo_for(int nd, const int n[3], const double extent[3])
{
    Geo g;
    g.nd = nd;
    for (int d = 0; d < 3; ++d) {
        if (d < g.nd) {
            g.n[d] = n[d];
            g.dx[d] = extent[d] / (double)n[d];
            g.idx[d] = 1.0 / g.dx[d];
        } else {
            g.n[d] = 0;
            g.dx[d] = 1.0;
            g.idx[d] = 1.0;
        }
    }
    // rows padded to 32 elements and 31 leading elements: Chmy index 1 of
    // every row starts a 128-byte line in FP64 and in FP32 storage
    const long align = 32;
    g.sy = ((long)(g.n[0] + 3) + align - 1) / align * align;
    g.sz = (g.nd == 3) ? g.sy * (g.n[1] + 3) : 0;
    g.off = align - 1;
    const long rows = (g.nd == 3) ? (long)(g.n[1] + 3) * (g.n[2] + 3) : (long)(g.n[1] + 3);
    g.elems = (size_t)(g.off + g.sy * rows + align);
    return g;
}

// interior of a field <-> contiguous x-fastest array: stage[t] holds interior
// element u = t0 + t, u = i + mx (j + my k), the field at Chmy index
// (i+1, j+1, k+1) (sz = 0 in 2D); t in [0, n)
__global__ __launch_bounds__(256) void k_interior_stage(double* stage, double* field, int mx, int my, long t0, long n,
                                                        long sy, long sz, int unpack)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long u = t0 + t;
        const long i = u % mx, r = u / mx, j = r % my, k = r / my;
        const long c = 1 + i + sy * (1 + j) + sz * (1 + k);
        if (unpack) field[c] = stage[t];
        else stage[t] = field[c];
    }
}

int field_transfer(const Geo& g, double* dev, int L, double* host, size_t count, int layout, bool up, hipStream_t stream,
                   double** stage, size_t* stage_elems)
{
    HIPCHK(hipStreamSynchronize(stream));
    if (layout == FI_LAYOUT_PARENT) {
        const size_t w = (size_t)(g.n[0] + 3) * sizeof(double);
        const size_t rows = (g.nd == 3) ? (size_t)(g.n[1] + 3) * (g.n[2] + 3) : (size_t)(g.n[1] + 3);
        if (up) HIPCHK(hipMemcpy2D(dev, g.sy * sizeof(double), host, w, w, rows, hipMemcpyHostToDevice));
        else HIPCHK(hipMemcpy2D(host, w, dev, g.sy * sizeof(double), w, rows, hipMemcpyDeviceToHost));
        return FI_OK;
    }
    const int mx = g.n[0] + (L & 1), my = g.n[1] + ((L >> 1) & 1);
    const int mz = (g.nd == 3) ? g.n[2] + ((L >> 2) & 1) : 1;
    const size_t n = (size_t)mx * my * mz;
    if (n != count) return fail(FI_ERR_ARG, "interior transfer of %zu doubles, got %zu", n, count);
    if (n > *stage_elems) {
        if (*stage) HIPCHK(hipFree(*stage));
        *stage = nullptr;
        *stage_elems = 0;
        if (hipMalloc(stage, n * sizeof(double)) != hipSuccess)
            return fail(FI_ERR_OOM, "transfer staging (%zu bytes)", n * sizeof(double));
        *stage_elems = n;
    }
    const long sz = g.nd == 3 ? g.sz : 0;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (up) {
        HIPCHK(hipMemcpyAsync(*stage, host, n * sizeof(double), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_interior_stage, dim3((unsigned)blocks), dim3(256), 0, stream, *stage, dev, mx, my, 0L,
                           (long)n, g.sy, sz, 1);
    } else {
        hipLaunchKernelGGL(k_interior_stage, dim3((unsigned)blocks), dim3(256), 0, stream, *stage, dev, mx, my, 0L,
                           (long)n, g.sy, sz, 0);
        HIPCHK(hipMemcpyAsync(host, *stage, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    return FI_OK;
}

// Staging area of fi_field_download_noalloc: 16 MiB of device memory per GPU,
// held until the process ends.  Host code never uses this name as an
// address: the device address of each GPU's copy comes from
// hipGetSymbolAddress (scratch_address below).
constexpr long kNoAllocStage = 2L << 20; // doubles
static __device__ __align__(256) double scratch[kNoAllocStage];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {};

// device address of the current device's staging area, looked up once per device
static int scratch_address(double** out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(FI_ERR_UNSUPPORTED, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(FI_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return FI_OK;
}

// interior download through the staging area, a run of whole x rows at a time
static int download_noalloc(const Geo& g, double* dev, int L, double* host, size_t count, hipStream_t stream)
{
    const int mx = g.n[0] + (L & 1), my = g.n[1] + ((L >> 1) & 1);
    const int mz = (g.nd == 3) ? g.n[2] + ((L >> 2) & 1) : 1;
    const size_t n = (size_t)mx * my * mz;
    if (n != count) return fail(FI_ERR_ARG, "interior transfer of %zu doubles, got %zu", n, count);
    const long rows_per_chunk = kNoAllocStage / mx;
    if (rows_per_chunk < 1)
        return fail(FI_ERR_ARG, "an x row of %d doubles exceeds the %ld-double staging area", mx, kNoAllocStage);
    double* stage = nullptr;
    int rc = scratch_address(&stage);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    const long rows = (long)my * mz, sz = g.nd == 3 ? g.sz : 0;
    for (long r0 = 0; r0 < rows; r0 += rows_per_chunk) {
        const long nr = std::min(rows_per_chunk, rows - r0);
        const long t0 = r0 * mx, nt = nr * mx;
        long blocks = (nt + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        // stream order: this chunk's packing follows the copy-out of the previous one
        hipLaunchKernelGGL(k_interior_stage, dim3((unsigned)blocks), dim3(256), 0, stream, stage, dev, mx, my, t0, nt,
                           g.sy, sz, 0);
        HIPCHK(hipMemcpyAsync(host + t0, stage, (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    return FI_OK;
}

} // namespace fi

extern "C" {

int fi_abi_version(void) { return FI_ABI_VERSION; }

const char* fi_last_error(void) { return g_err; }

// device layout: padded rows, Chmy index 1 of every row 128-byte aligned
static Geo make_geo(const fi_model_desc& desc) { return geo_for(desc.ndims, desc.n, desc.extent); }


size_t fi_model_bytes(const fi_model_desc* desc)
{
    if (check_desc(desc) != FI_OK) return 0;
    const Geo g = make_geo(*desc);
    size_t bytes = 4 * sizeof(unsigned long long);
    for (int id = 0; id < FI_F_COUNT; ++id)
        if (field_used(*desc, id))
            bytes += ((id >= FI_F_P && id <= FI_F_VZ) ? 2 * pair_stride(g) : g.elems) * sizeof(double);
    if (desc->immersed_boundary) bytes += g.elems;
    if (desc->immersed_boundary && g.nd == 3)
        bytes += (2 * col_ranges_count(g) + tile_march_count(g) + 3) * sizeof(int2);
    return bytes;
}

int fi_model_create(const fi_model_desc* desc, fi_model_t* out)
{
    if (!out) return fail(FI_ERR_ARG, "null output handle");
    *out = nullptr;
    int rc = check_desc(desc);
    if (rc) return rc;
    if (desc->precision == FI_PREC_FP32_REFINED)
        return fail(FI_ERR_UNSUPPORTED, "FP32-refined models are created through fi_refined_* entry points");
    fi_model_s* m = new (std::nothrow) fi_model_s();
    if (!m) return fail(FI_ERR_OOM, "host allocation failed");
    m->desc = *desc;
    m->g = make_geo(*desc);
    Geo& g = m->g;
    if (hipGetDevice(&m->device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        free_model(m);
        return fail(FI_ERR_HIP, "hipStreamCreate failed");
    }
    // the stress and velocity fields and their second buffers are one
    // allocation each, the second buffer pair_stride(g) elements after the
    // first: the fused immersed-boundary sweep addresses both through one
    // base pointer
    for (int id = 0; id < FI_F_COUNT; ++id) {
        if (!field_used(*desc, id)) continue;
        const bool pair = id >= FI_F_P && id <= FI_F_VZ;
        const size_t n = pair ? 2 * pair_stride(g) : g.elems;
        if (hipMalloc(&m->base[id], n * sizeof(double)) != hipSuccess) {
            free_model(m);
            return fail(FI_ERR_OOM, "device allocation of field %d (%zu bytes) failed", id, n * sizeof(double));
        }
        m->f[id] = m->base[id] + g.off;
        if (pair) m->sf[id] = m->base[id] + pair_stride(g) + g.off;
        if (hipMemsetAsync(m->base[id], 0, n * sizeof(double), m->stream) != hipSuccess) {
            free_model(m);
            return fail(FI_ERR_HIP, "hipMemsetAsync failed");
        }
    }
    if (desc->immersed_boundary) {
        if (hipMalloc(&m->flags_base, g.elems) != hipSuccess) {
            free_model(m);
            return fail(FI_ERR_OOM, "device allocation of the validity flags failed");
        }
        m->flags = m->flags_base + g.off;
        if (hipMemsetAsync(m->flags_base, 0, g.elems, m->stream) != hipSuccess) {
            free_model(m);
            return fail(FI_ERR_HIP, "hipMemsetAsync failed");
        }
        if (g.nd == 3) {
            const size_t nc = col_ranges_count(g);
            if (hipMalloc(&m->crange, nc * sizeof(int2)) != hipSuccess ||
                hipMalloc(&m->erange, nc * sizeof(int2)) != hipSuccess ||
                hipMalloc(&m->tmarch, (tile_march_count(g) + 3) * sizeof(int2)) != hipSuccess) {
                free_model(m);
                return fail(FI_ERR_OOM, "device allocation of the column ranges failed");
            }
        }
    }
    if (fused_fits(g)) m->has_interior = fused_split(g, 8, m->kch, m->fg, m->shell);
    if (hipMalloc(&m->d_norm, 4 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(&m->h_norm, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        free_model(m);
        return fail(FI_ERR_OOM, "norm buffers");
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess) {
        free_model(m);
        return fail(FI_ERR_HIP, "stream sync after allocation failed");
    }
    *out = m;
    return FI_OK;
}

int fi_model_destroy(fi_model_t m)
{
    if (!m) return FI_OK;
    if (m->batch) batch_detach(m);
    (void)hipStreamSynchronize(m->stream);
    free_model(m);
    return FI_OK;
}

size_t fi_field_size(fi_model_t m, int id, int layout)
{
    if (!m || id < 0 || id >= FI_F_COUNT || !m->f[id]) return 0;
    if (layout == FI_LAYOUT_PARENT) return parent_count(m->g);
    int L = field_loc(m->g.nd, id);
    return L < 0 ? 0 : interior_count(m->g, L);
}

static void mark_dirty(fi_model_t m, int id)
{
    if (id >= FI_F_P && id <= FI_F_VZ) m->state_written();
    if (id >= FI_F_RP && id <= FI_F_RVZ) m->resid_clean = false;
    if (id >= FI_F_OM_NA0 && id < FI_F_COUNT) {
        m->flags_dirty = true;
        ++m->mask_gen;
    }
    if (m->batch) batch_touch(m); // the batch re-prepares the model before its next step
}

static int field_copy(fi_model_t m, int id, double* host, size_t count, int layout, bool up)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (id < 0 || id >= FI_F_COUNT || !m->f[id]) return fail(FI_ERR_ARG, "field %d not present in this model", id);
    if (!host) return fail(FI_ERR_ARG, "null host pointer");
    size_t need = fi_field_size(m, id, layout);
    if (layout != FI_LAYOUT_PARENT && layout != FI_LAYOUT_INTERIOR) return fail(FI_ERR_ARG, "bad layout %d", layout);
    if (count != need) return fail(FI_ERR_ARG, "field %d needs %zu doubles in this layout, got %zu", id, need, count);
    if (up) mark_dirty(m, id);
    return field_transfer(m->g, m->buf(id), field_loc(m->g.nd, id), host, count, layout, up, m->stream, &m->d_stage,
                          &m->stage_elems);
}

int fi_field_upload(fi_model_t m, int id, const double* host, size_t count, int layout)
{
    return field_copy(m, id, const_cast<double*>(host), count, layout, true);
}

int fi_field_download(fi_model_t m, int id, double* host, size_t count, int layout)
{
    return field_copy(m, id, host, count, layout, false);
}

int fi_field_download_noalloc(fi_model_t m, int id, double* host, size_t count)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (id < 0 || id >= FI_F_COUNT || !m->f[id]) return fail(FI_ERR_ARG, "field %d not present in this model", id);
    if (!host) return fail(FI_ERR_ARG, "null host pointer");
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != m->device) HIPCHK(hipSetDevice(m->device));
    int rc = download_noalloc(m->g, m->buf(id), field_loc(m->g.nd, id), host, count, m->stream);
    if (cur != m->device) {
        hipError_t e = hipSetDevice(cur);
        if (!rc && e != hipSuccess) rc = fail(FI_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    }
    return rc;
}

int fi_field_fill(fi_model_t m, int id, double value)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (id < 0 || id >= FI_F_COUNT || !m->f[id]) return fail(FI_ERR_ARG, "field %d not present in this model", id);
    double* p = m->buf(id);
    mark_dirty(m, id);
    // the whole allocation, i.e. the parent box including halos and row padding
    launch_fill(p - m->g.off, m->g.elems, value, m->stream);
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_apply_bc(fi_model_t m, int group)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (group < FI_BCGROUP_STRESS || group > FI_BCGROUP_RESIDUAL) return fail(FI_ERR_ARG, "bad bc group %d", group);
    if (group == FI_BCGROUP_ETA) {
        // bc!(eta => Neumann()) and bc!(eta_next => Neumann()) (tm_stokes_chmy.jl:93-94)
        BcBatch b[3];
        for (int d = 0; d < m->g.nd; ++d) {
            b[d] = neumann_batch(m->eta(), L_CCC, d);
            add_op(b[d], m->eta_next(), L_CCC, 1, 0.0);
        }
        launch_bc_dims(m->g, b, m->stream);
    } else if (group == FI_BCGROUP_VELOCITY) {
        BcBatch b[3];
        for (int d = 0; d < m->g.nd; ++d) b[d] = velocity_batch(m, d, false);
        launch_bc_dims(m->g, b, m->stream);
        m->state_written();
    } else {
        run_bc(m, group);
        if (group == FI_BCGROUP_STRESS) m->state_written();
    }
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_advance_iterations(fi_model_t m, int n_iter)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (n_iter < 0) return fail(FI_ERR_ARG, "negative iteration count");
    int rc = advance(m, n_iter);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_compute_residuals(fi_model_t m, double norms_inf[4])
{
    if (!m || !norms_inf) return fail(FI_ERR_ARG, "null argument");
    int rc = residual_norms(m, norms_inf);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_compute_tau(fi_model_t m)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    int rc = prepare(m);
    if (rc) return rc;
    StokesArgs<double> a = make_args(m, true);
    launch_tau(m->g, a, m->desc.immersed_boundary != 0, m->stream);
    m->state_written();
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_solve(fi_model_t m, int max_iter, int ncheck, double tol, int* iters_out, double norms_inf[4])
{
    if (!m || !norms_inf) return fail(FI_ERR_ARG, "null argument");
    if (max_iter < 0 || ncheck < 1) return fail(FI_ERR_ARG, "bad max_iter/ncheck");
    int done = 0;
    for (int q = 0; q < 4; ++q) norms_inf[q] = 0.0;
    while (done < max_iter) {
        int chunk = ncheck - (done % ncheck);
        if (done + chunk > max_iter) chunk = max_iter - done;
        int rc = advance(m, chunk);
        if (rc) return rc;
        done += chunk;
        if (done % ncheck == 0) {
            int rc = residual_norms(m, norms_inf);
            if (rc) return rc;
            bool conv = true;
            for (int q = 0; q < 4; ++q) {
                if (!std::isfinite(norms_inf[q])) {
                    if (iters_out) *iters_out = done;
                    return fail(FI_ERR_NONFINITE, "simulation failed at iteration %d: non-finite residual", done);
                }
                if (!(norms_inf[q] < tol)) conv = false;
            }
            if (conv) break;
        }
    }
    if (iters_out) *iters_out = done;
    HIPCHK(hipStreamSynchronize(m->stream));
    return FI_OK;
}

int fi_time_iterations(fi_model_t m, int n_iter, double out_ms[4])
{
    if (!m || !out_ms || n_iter < 1) return fail(FI_ERR_ARG, "bad arguments");
    int rc = prepare(m);
    if (rc) return rc;
    hipEvent_t ev[6];
    for (int q = 0; q < 6; ++q) HIPCHK(hipEventCreate(&ev[q]));
    double acc[4] = {0, 0, 0, 0};
    for (int it = 0; it < n_iter; ++it) {
        iterate(m, ev);
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

int fi_fused_region(fi_model_t m, int lo[3], int hi[3])
{
    if (!m || !lo || !hi) return fail(FI_ERR_ARG, "null argument");
    const bool on = m->use_fused && m->has_interior;
    for (int d = 0; d < 3; ++d) {
        lo[d] = on ? m->fg.lo[d] : 0;
        hi[d] = on ? m->fg.hi[d] : -1;
    }
    return on ? 1 : 0;
}

int fi_fused_active_cells(fi_model_t m, long long out[2])
{
    if (!m || !out) return fail(FI_ERR_ARG, "null argument");
    out[0] = out[1] = 0;
    if (!(m->use_fused && m->has_interior)) return FI_OK;
    const FusedGeom& fg = m->fg;
    const long long nz = fg.hi[2] - fg.lo[2] + 1;
    out[0] = (long long)(fg.hi[0] - fg.lo[0] + 1) * (fg.hi[1] - fg.lo[1] + 1) * nz;
    if (!m->crange) {
        out[1] = out[0];
        return FI_OK;
    }
    int rc = prepare(m, false);
    if (rc) return rc;
    std::vector<int2> r(col_ranges_count(m->g));
    HIPCHK(hipMemcpyAsync(r.data(), m->crange, r.size() * sizeof(int2), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    for (int j = fg.lo[1]; j <= fg.hi[1]; ++j)
        for (int i = fg.lo[0]; i <= fg.hi[0]; ++i) {
            const int2 c = r[(size_t)i + (size_t)m->g.sy * j];
            const long long lo = std::max(c.x, fg.lo[2]), hi = std::min(c.y, fg.hi[2]);
            if (hi >= lo) out[1] += hi - lo + 1;
        }
    return FI_OK;
}

int fi_ib_validity(fi_model_t m, uint8_t* host, size_t count)
{
    if (!m || !host) return fail(FI_ERR_ARG, "null argument");
    if (!m->desc.immersed_boundary) return fail(FI_ERR_ARG, "model has no immersed boundary");
    const Geo& g = m->g;
    const size_t w = (size_t)(g.n[0] + 3);
    const size_t rows = (g.nd == 3) ? (size_t)(g.n[1] + 3) * (g.n[2] + 3) : (size_t)(g.n[1] + 3);
    if (count != w * rows) return fail(FI_ERR_ARG, "validity bits need %zu bytes, got %zu", w * rows, count);
    int rc = prepare(m, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy2D(host, w, m->flags, (size_t)g.sy, w, rows, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < count; ++q) host[q] &= (uint8_t)~F_FULL; // validity bits only
    return FI_OK;
}

int fi_set_option(fi_model_t m, int key, int value)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (key == FI_OPT_FUSED) {
        m->use_fused = value ? 1 : 0;
    } else if (key == FI_OPT_ZCHUNK) {
        if (value < 0) return fail(FI_ERR_ARG, "z chunk must be >= 1 (0: sized to the grid)");
        m->kch = value;
        if (fused_fits(m->g)) m->has_interior = fused_split(m->g, 8, m->kch, m->fg, m->shell);
    } else if (key == FI_OPT_REFINE_FULL_PHASE) {
        m->refine_full_phase = value ? 1 : 0;
    } else if (key == FI_OPT_REFINE_ETA_RATE) {
        if (value < 1 || value > 1000) return fail(FI_ERR_ARG, "viscosity relaxation rate must be in 1..1000 per mille");
        m->refine_eta_rate = value / 1000.0;
    } else if (key == FI_OPT_FUSED_VARIANT || key == FI_OPT_FUSED_VARIANT_FP32) {
        const bool f32 = key == FI_OPT_FUSED_VARIANT_FP32;
        int cfg[16][3];
        const int n = fused_configs(f32, cfg, 16);
        if (value < 0 || value >= n) return fail(FI_ERR_ARG, "fused variant must be in 0..%d", n - 1);
        if (f32) {
            m->ftx32 = cfg[value][0];
            m->fty32 = cfg[value][1];
            m->fminw32 = cfg[value][2];
        } else {
            m->ftx = cfg[value][0];
            m->fty = cfg[value][1];
            m->fminw = cfg[value][2];
        }
    } else if (key == FI_OPT_FUSED_SHELL) {
        m->fused_shell = value != 0;
    } else if (key == FI_OPT_REFINE_COLS) {
        m->refine_cols = value != 0;
    } else if (key == FI_OPT_REFINE_MODE) {
        if (value < 0 || value > 1) return fail(FI_ERR_ARG, "refinement mode must be 0 or 1");
        m->refine_mode = value;
    } else if (key == FI_OPT_REFINE_ANDERSON) {
        if (value < 0 || value > 3) return fail(FI_ERR_ARG, "Anderson depth must be in 0..3");
        m->refine_anderson = value;
    } else if (key == FI_OPT_TILE_ORDER) {
        if (value < 0 || value > 2) return fail(FI_ERR_ARG, "tile order must be 0, 1 or 2");
        m->forder = value;
    } else if (key == FI_OPT_REFINE_PHASE1_DIGITS) {
        if (value < 1 || value > 7) return fail(FI_ERR_ARG, "nonlinear phase digits must be in 1..7");
        m->refine_phase1_digits = value;
    } else {
        return fail(FI_ERR_ARG, "unknown option %d", key);
    }
    return FI_OK;
}

int fi_fused_variants(int fp32, int* tx_ty_minw, int max_variants)
{
    int cfg[16][3];
    const int n = fused_configs(fp32 != 0, cfg, 16);
    if (tx_ty_minw)
        for (int q = 0; q < n && q < max_variants; ++q)
            for (int k = 0; k < 3; ++k) tx_ty_minw[3 * q + k] = cfg[q][k];
    return n;
}

int fi_synchronize(fi_model_t m)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    HIPCHK(hipStreamSynchronize(m->stream));
    return FI_OK;
}

int fi_levelset_from_dem(fi_model_t m, int phase, const double* dem_parent, int dem_nx, int dem_ny,
                         const double dem_origin[2], const double dem_extent[2], const double R[9])
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (!m->desc.immersed_boundary) return fail(FI_ERR_STATE, "model has no immersed boundary");
    if (m->g.nd != 3) return fail(FI_ERR_UNSUPPORTED, "level sets from a DEM are defined for 3D grids");
    if (phase != FI_PHASE_NA && phase != FI_PHASE_NS) return fail(FI_ERR_ARG, "bad phase %d", phase);
    if (!dem_parent || dem_nx < 1 || dem_ny < 1 || !dem_origin || !dem_extent) return fail(FI_ERR_ARG, "bad DEM");
    if (!(dem_extent[0] > 0.0) || !(dem_extent[1] > 0.0)) return fail(FI_ERR_ARG, "bad DEM extent");
    const size_t elems = (size_t)(dem_nx + 3) * (dem_ny + 3);
    if (elems > m->dem_elems) {
        if (m->d_dem) HIPCHK(hipFree(m->d_dem));
        m->d_dem = nullptr;
        m->dem_elems = 0;
        HIPCHK(hipMalloc(&m->d_dem, elems * sizeof(double)));
        m->dem_elems = elems;
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(m->d_dem, dem_parent, elems * sizeof(double), hipMemcpyHostToDevice));
    DemArgs a;
    a.dem = m->d_dem;
    a.nv[0] = dem_nx + 1;
    a.nv[1] = dem_ny + 1;
    a.sy = dem_nx + 3;
    for (int d = 0; d < 2; ++d) {
        a.o[d] = dem_origin[d];
        a.ext[d] = dem_extent[d];
        a.d[d] = dem_extent[d] / (double)(d == 0 ? dem_nx : dem_ny);
    }
    a.has_R = R ? 1 : 0;
    for (int q = 0; q < 9; ++q) a.R[q] = R ? R[q] : (q % 4 == 0 ? 1.0 : 0.0);
    double* psi = m->f[phase == FI_PHASE_NA ? FI_F_PSI_NA : FI_F_PSI_NS];
    launch_levelset_from_dem(m->g, m->desc.origin, psi, a, m->stream);
    for (int d = 0; d < 3; ++d) launch_bc_dim(m->g, neumann_batch(psi, L_VVV, d), m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    return FI_OK;
}

int fi_levelset_from_polygon(fi_model_t m, int phase, const double* xy, int npts)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (!m->desc.immersed_boundary) return fail(FI_ERR_STATE, "model has no immersed boundary");
    if (m->g.nd != 2) return fail(FI_ERR_UNSUPPORTED, "polygon level sets are defined for 2D grids");
    if (phase != FI_PHASE_NA && phase != FI_PHASE_NS) return fail(FI_ERR_ARG, "bad phase %d", phase);
    if (!xy || npts < 3) return fail(FI_ERR_ARG, "a polygon needs at least 3 points");
    const size_t elems = 2 * (size_t)npts;
    for (size_t q = 0; q < elems; ++q)
        if (!std::isfinite(xy[q])) return fail(FI_ERR_ARG, "polygon coordinates must be finite");
    for (int q = 0; q < npts; ++q) {
        const int r = (q + 1) % npts;
        if (xy[2 * q] == xy[2 * r] && xy[2 * q + 1] == xy[2 * r + 1])
            return fail(FI_ERR_ARG, "polygon has a zero-length edge at point %d", q);
    }
    if (elems > m->dem_elems) {
        if (m->d_dem) HIPCHK(hipFree(m->d_dem));
        m->d_dem = nullptr;
        m->dem_elems = 0;
        HIPCHK(hipMalloc(&m->d_dem, elems * sizeof(double)));
        m->dem_elems = elems;
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(m->d_dem, xy, elems * sizeof(double), hipMemcpyHostToDevice));
    double* psi = m->f[phase == FI_PHASE_NA ? FI_F_PSI_NA : FI_F_PSI_NS];
    launch_levelset_from_polygon(m->g, m->desc.origin, psi, m->d_dem, npts, m->stream);
    for (int d = 0; d < 2; ++d) launch_bc_dim(m->g, neumann_batch(psi, L_VVV, d), m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    return FI_OK;
}

int fi_invert_levelset(fi_model_t m, int phase)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (!m->desc.immersed_boundary) return fail(FI_ERR_STATE, "model has no immersed boundary");
    if (phase != FI_PHASE_NA && phase != FI_PHASE_NS) return fail(FI_ERR_ARG, "bad phase %d", phase);
    double* psi = m->f[phase == FI_PHASE_NA ? FI_F_PSI_NA : FI_F_PSI_NS];
    launch_negate_vertex(m->g, psi, m->stream);
    const int L = m->g.nd == 3 ? L_VVV : L_VVC;
    for (int d = 0; d < m->g.nd; ++d) launch_bc_dim(m->g, neumann_batch(psi, L, d), m->stream);
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_omega_from_psi(fi_model_t m, int phase)
{
    if (!m) return fail(FI_ERR_ARG, "null model");
    if (!m->desc.immersed_boundary) return fail(FI_ERR_STATE, "model has no immersed boundary");
    if (phase != FI_PHASE_NA && phase != FI_PHASE_NS) return fail(FI_ERR_ARG, "bad phase %d", phase);
    const double* psi = m->f[phase == FI_PHASE_NA ? FI_F_PSI_NA : FI_F_PSI_NS];
    double* om[8];
    const int base = phase == FI_PHASE_NA ? FI_F_OM_NA0 : FI_F_OM_NS0;
    for (int q = 0; q < 8; ++q) om[q] = m->f[base + q];
    launch_omega_from_psi(m->g, psi, om, m->stream);
    m->flags_dirty = true;
    ++m->mask_gen;
    HIPCHK(hipGetLastError());
    return FI_OK;
}

int fi_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return FI_OK;
}

} // extern "C"
