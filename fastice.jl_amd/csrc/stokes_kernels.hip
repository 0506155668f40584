// stokes_kernels.hip -- PT-Stokes sweep kernels for gfx950.
//
// Every expression keeps the operation order of the reference source line it
// cites (Julia evaluates left to right) and the library is compiled with
// -ffp-contract=off, so the FP64 kernels round exactly like the CPU
// restatement of the same formulas.  The arithmetic of one staggered point
// lives in the inline helpers below and is shared by the pointwise box
// kernels and the fused 2.5D sweep.
//
// Discrete operators (Chmy semantics, fixed by this library):
//   d/dx of a Vertex-in-x field at a Center index: (f[i+1]-f[i]) * (1/dx)
//   d/dx of a Center-in-x field at a Vertex index: (f[i]-f[i-1]) * (1/dx)
//   lerp Center->Vertex 0.5*(f[i-1]+f[i]); Vertex->Center 0.5*(f[i]+f[i+1]),
//   nested x innermost, then y, then z.
//   masked d/dx(f, w): (w[r]*f[r] - w[l]*f[l]) * (1/dx), w at f's location.
#include <algorithm>

#include "fi_internal.h"

namespace fi {

// empty plane range of a column (lo > hi), see k_col_range
constexpr int RANGE_NONE = 1 << 20;

// ---------------------------------------------------------------------------
// point arithmetic
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T dif(T r, T l, T inv)
{
    return (r - l) * inv;
}
template <typename T>
__device__ __forceinline__ T lerp4(T a, T b, T c, T d)
{
    return T(0.5) * (T(0.5) * (a + b) + T(0.5) * (c + d));
}
// P -= divV * eta * dtau.Pr                             (kernels_3d.jl:16)
template <typename T>
__device__ __forceinline__ T upd_P(T p, T divV, T eta, T dPr)
{
    return p - divV * eta * dPr;
}
// tau_ii -= (tau_ii - 2 eta (e_ii - divV/3)) * dtau      (kernels_3d.jl:18-20)
template <typename T>
__device__ __forceinline__ T upd_td(T t, T eta, T e, T divV, T dt)
{
    return t - (t - T(2) * eta * (e - divV / T(3))) * dt;
}
// tau_ij -= (tau_ij - 2 eta_ij e_ij) * dtau               (kernels_3d.jl:22-24)
template <typename T>
__device__ __forceinline__ T upd_to(T t, T eta, T e, T dt)
{
    return t - (t - T(2) * eta * e) * dt;
}
// V += (-dP + dtau_a + dtau_b + dtau_c - g) / eta_V * dtau.V   (kernels_3d.jl:31-33)
template <typename T>
__device__ __forceinline__ T upd_V(T v, T dP, T da, T db, T dc, T g, T ev, T dv)
{
    return v + (-dP + da + db + dc - g) / ev * dv;
}
// IB variant: ... / eta_V / omega_ns * dtau.V             (IB kernels_3d.jl:51-52)
template <typename T>
__device__ __forceinline__ T upd_V_ib(T v, T dP, T da, T db, T dc, T g, T ev, T w, T dv)
{
    return v + (-dP + da + db + dc - g) / ev / w * dv;
}
// The FP32 sweeps (corrections of the refined solver, whose accuracy the FP64
// outer steps set) divide with the hardware reciprocal (v_rcp_f32, within 1
// ulp) instead of the correctly rounded division sequence: the three
// divisions by 3 become products with 1/3 and the velocity update one
// reciprocal of eta_V * omega_ns.  The FP64 iteration keeps the reference's
// correctly rounded divisions.
template <>
__device__ __forceinline__ float upd_td<float>(float t, float eta, float e, float divV, float dt)
{
    return t - (t - 2.0f * eta * (e - divV * (1.0f / 3.0f))) * dt;
}
template <>
__device__ __forceinline__ float upd_V<float>(float v, float dP, float da, float db, float dc, float g, float ev, float dv)
{
    return v + (-dP + da + db + dc - g) * __builtin_amdgcn_rcpf(ev) * dv;
}
template <>
__device__ __forceinline__ float upd_V_ib<float>(float v, float dP, float da, float db, float dc, float g, float ev,
                                                 float w, float dv)
{
    return v + (-dP + da + db + dc - g) * __builtin_amdgcn_rcpf(ev * w) * dv;
}

// tau_II (kernels_3d.jl:35-38)
template <typename T>
__device__ __forceinline__ T tauII(T xx, T yy, T zz, T lxy, T lxz, T lyz)
{
    return sqrt(T(0.5) * (xx * xx + yy * yy + zz * zz) + lxy * lxy + lxz * lxz + lyz * lyz);
}

template <typename T>
__device__ __forceinline__ T ipow_(T x, T p)
{
    if (p == T(0)) return T(1);
    if (p == T(1)) return x;
    if (p == T(2)) return x * x;
    if (p == T(3)) return x * x * x;
    return pow(x, p);
}

// Physics.jl:11,18,26,46-47 (the relaxed field is the lagged viscosity eta)
template <typename T>
__device__ __forceinline__ T rheology(const StokesArgs<T>& a, long c, T tII)
{
    T eta;
    if (a.rh_kind == FI_RHEO_LINEAR_CONST) eta = a.rh_eta;
    else if (a.rh_kind == FI_RHEO_LINEAR_FIELD) eta = a.rh[c];
    else {
        T A = a.rh ? a.rh[c] : a.rh_A;
        eta = T(0.5) / (A * ipow_(tII, a.rh_n - T(1)));
    }
    if (a.rh_has_reg) eta = T(1) / (T(1) / eta + a.rh_inv_eta_reg);
    if (a.rh_relax != FI_RELAX_NONE) {
        T old = a.eta[c];
        T w = T(1) - a.rh_rate;
        if (a.rh_relax == FI_RELAX_LIN) eta = old * w + eta * a.rh_rate;
        else if (a.rh_relax == FI_RELAX_LOG) eta = exp(log(old) * w + log(eta) * a.rh_rate);
        else eta = pow(T(10), log10(old) * w + log10(eta) * a.rh_rate);
    }
    return eta;
}

// The same at byte offset bo with the lagged viscosity at bo in a register,
// split into the target viscosity, the target in the relaxation's scale (its
// log or log10, for a constant-coefficient rheology the same at every cell
// when tII is) and the relaxation towards it.
template <typename T>
__device__ __forceinline__ T rheo_target(const StokesArgs<T>& a, uint32_t bo, T tII)
{
    T eta;
    if (a.rh_kind == FI_RHEO_LINEAR_CONST) eta = a.rh_eta;
    else if (a.rh_kind == FI_RHEO_LINEAR_FIELD) eta = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a.rh) + bo);
    else {
        T A = a.rh ? *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a.rh) + bo) : a.rh_A;
        eta = T(0.5) / (A * ipow_(tII, a.rh_n - T(1)));
    }
    if (a.rh_has_reg) eta = T(1) / (T(1) / eta + a.rh_inv_eta_reg);
    return eta;
}
template <typename T>
__device__ __forceinline__ T rheo_scaled(const StokesArgs<T>& a, T eta)
{
    if (a.rh_relax == FI_RELAX_LOG) return log(eta);
    if (a.rh_relax == FI_RELAX_LOG10) return log10(eta);
    return eta;
}
template <typename T>
__device__ __forceinline__ T rheo_relax(const StokesArgs<T>& a, T st, T old)
{
    if (a.rh_relax == FI_RELAX_NONE) return st;
    const T w = T(1) - a.rh_rate;
    if (a.rh_relax == FI_RELAX_LIN) return old * w + st * a.rh_rate;
    if (a.rh_relax == FI_RELAX_LOG) return exp(log(old) * w + st * a.rh_rate);
    return pow(T(10), log10(old) * w + st * a.rh_rate);
}

__device__ __forceinline__ bool same_bits(double x, double y) { return __double_as_longlong(x) == __double_as_longlong(y); }
__device__ __forceinline__ bool same_bits(float x, float y) { return __float_as_uint(x) == __float_as_uint(y); }

template <typename T>
__device__ __forceinline__ T rheology_o(const StokesArgs<T>& a, uint32_t bo, T tII, T old)
{
    T eta;
    if (a.rh_kind == FI_RHEO_LINEAR_CONST) eta = a.rh_eta;
    else if (a.rh_kind == FI_RHEO_LINEAR_FIELD) eta = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a.rh) + bo);
    else {
        T A = a.rh ? *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a.rh) + bo) : a.rh_A;
        eta = T(0.5) / (A * ipow_(tII, a.rh_n - T(1)));
    }
    if (a.rh_has_reg) eta = T(1) / (T(1) / eta + a.rh_inv_eta_reg);
    if (a.rh_relax != FI_RELAX_NONE) {
        T w = T(1) - a.rh_rate;
        if (a.rh_relax == FI_RELAX_LIN) eta = old * w + eta * a.rh_rate;
        else if (a.rh_relax == FI_RELAX_LOG) eta = exp(log(old) * w + log(eta) * a.rh_rate);
        else eta = pow(T(10), log10(old) * w + log10(eta) * a.rh_rate);
    }
    return eta;
}

// ---------------------------------------------------------------------------
// 3D pointwise sweeps (used on the boundary shell and for small grids)
// ---------------------------------------------------------------------------
// update_sigma!: FullStokes kernels_3d.jl:3-25, IB kernels_3d.jl:3-44
// With an immersed boundary, a point whose stress locations are all invalid
// is left alone once the output buffers are known to hold 0 there (z_wsig,
// z_wvel: see fused3_ib_tile); so are the zero stores of invalid locations.
template <typename T, bool IB>
__device__ __forceinline__ void sigma_point3(const StokesArgs<T>& a, long c, long sy, long sz)
{
    if (IB && a.z_wsig && !(a.flags[c] & (F_CCC | F_VVC | F_VCV | F_CVV))) return;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz, *eta = a.eta;
    const T ec = eta[c];
    const T exy_eta = lerp4(eta[c - 1 - sy], eta[c - sy], eta[c - 1], eta[c]);
    const T exz_eta = lerp4(eta[c - 1 - sz], eta[c - sz], eta[c - 1], eta[c]);
    const T eyz_eta = lerp4(eta[c - sy - sz], eta[c - sz], eta[c - sy], eta[c]);
    if (!IB) {
        T exx = dif(Vx[c + 1], Vx[c], ix);
        T eyy = dif(Vy[c + sy], Vy[c], iy);
        T ezz = dif(Vz[c + sz], Vz[c], iz);
        T exy = T(0.5) * (dif(Vy[c], Vy[c - 1], ix) + dif(Vx[c], Vx[c - sy], iy));
        T exz = T(0.5) * (dif(Vz[c], Vz[c - 1], ix) + dif(Vx[c], Vx[c - sz], iz));
        T eyz = T(0.5) * (dif(Vz[c], Vz[c - sy], iy) + dif(Vy[c], Vy[c - sz], iz));
        T divV = exx + eyy + ezz;
        a.Pn[c] = upd_P(a.P[c], a.fP ? divV + a.fP[c] : divV, ec, a.dPr);
        a.txxn[c] = upd_td(a.txx[c], ec, exx, divV, a.dtau[0]);
        a.tyyn[c] = upd_td(a.tyy[c], ec, eyy, divV, a.dtau[1]);
        a.tzzn[c] = upd_td(a.tzz[c], ec, ezz, divV, a.dtau[2]);
        a.txyn[c] = upd_to(a.txy[c], exy_eta, exy, a.dtau[3]);
        a.txzn[c] = upd_to(a.txz[c], exz_eta, exz, a.dtau[4]);
        a.tyzn[c] = upd_to(a.tyz[c], eyz_eta, eyz, a.dtau[5]);
    } else {
        const auto wx = a.ns[OM_VCC], wy = a.ns[OM_CVC], wz = a.ns[OM_CCV];
        const unsigned fl = a.flags[c];
        if (fl & F_CCC) {
            T exx = dif(wx[c + 1] * Vx[c + 1], wx[c] * Vx[c], ix);
            T eyy = dif(wy[c + sy] * Vy[c + sy], wy[c] * Vy[c], iy);
            T ezz = dif(wz[c + sz] * Vz[c + sz], wz[c] * Vz[c], iz);
            T divV = exx + eyy + ezz;
            a.Pn[c] = upd_P(a.P[c], a.fP ? divV + a.fP[c] : divV, ec, a.dPr);
            a.txxn[c] = upd_td(a.txx[c], ec, exx, divV, a.dtau[0]);
            a.tyyn[c] = upd_td(a.tyy[c], ec, eyy, divV, a.dtau[1]);
            a.tzzn[c] = upd_td(a.tzz[c], ec, ezz, divV, a.dtau[2]);
        } else if (!a.z_wsig) {
            a.Pn[c] = T(0); a.txxn[c] = T(0); a.tyyn[c] = T(0); a.tzzn[c] = T(0);
        }
        if (fl & F_VVC) {
            T exy = T(0.5) * (dif(wy[c] * Vy[c], wy[c - 1] * Vy[c - 1], ix) + dif(wx[c] * Vx[c], wx[c - sy] * Vx[c - sy], iy));
            a.txyn[c] = upd_to(a.txy[c], exy_eta, exy, a.dtau[3]);
        } else if (!a.z_wsig) a.txyn[c] = T(0);
        if (fl & F_VCV) {
            T exz = T(0.5) * (dif(wz[c] * Vz[c], wz[c - 1] * Vz[c - 1], ix) + dif(wx[c] * Vx[c], wx[c - sz] * Vx[c - sz], iz));
            a.txzn[c] = upd_to(a.txz[c], exz_eta, exz, a.dtau[4]);
        } else if (!a.z_wsig) a.txzn[c] = T(0);
        if (fl & F_CVV) {
            T eyz = T(0.5) * (dif(wz[c] * Vz[c], wz[c - sy] * Vz[c - sy], iy) + dif(wy[c] * Vy[c], wy[c - sz] * Vy[c - sz], iz));
            a.tyzn[c] = upd_to(a.tyz[c], eyz_eta, eyz, a.dtau[5]);
        } else if (!a.z_wsig) a.tyzn[c] = T(0);
    }
}

// update_V!: FullStokes kernels_3d.jl:27-40, IB kernels_3d.jl:46-74 (reads the new stress)
template <typename T, bool IB>
__device__ __forceinline__ void velocity_point3(const StokesArgs<T>& a, long c, long sy, long sz)
{
    if (IB && a.z_wvel && a.frozen_eta && !(a.flags[c] & (F_VCC | F_CVC | F_CCV))) return;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *P = a.Pn, *txx = a.txxn, *tyy = a.tyyn, *tzz = a.tzzn, *txy = a.txyn, *txz = a.txzn, *tyz = a.tyzn;
    const T* eta = a.eta;
    const T gx = a.grav_field ? a.gx[c] : a.grav[0];
    const T gy = a.grav_field ? a.gy[c] : a.grav[1];
    const T gz = a.grav_field ? a.gz[c] : a.grav[2];
    const T ex = T(0.5) * (eta[c - 1] + eta[c]);
    const T ey = T(0.5) * (eta[c - sy] + eta[c]);
    const T ez = T(0.5) * (eta[c - sz] + eta[c]);
    if (!IB) {
        a.Vxn[c] = upd_V(a.Vx[c], dif(P[c], P[c - 1], ix), dif(txx[c], txx[c - 1], ix), dif(txy[c + sy], txy[c], iy),
                         dif(txz[c + sz], txz[c], iz), gx, ex, a.dV[0]);
        a.Vyn[c] = upd_V(a.Vy[c], dif(P[c], P[c - sy], iy), dif(tyy[c], tyy[c - sy], iy), dif(txy[c + 1], txy[c], ix),
                         dif(tyz[c + sz], tyz[c], iz), gy, ey, a.dV[1]);
        a.Vzn[c] = upd_V(a.Vz[c], dif(P[c], P[c - sz], iz), dif(tzz[c], tzz[c - sz], iz), dif(txz[c + 1], txz[c], ix),
                         dif(tyz[c + sy], tyz[c], iy), gz, ez, a.dV[2]);
    } else {
        const auto wc = a.na[OM_CCC], wxy = a.na[OM_VVC], wxz = a.na[OM_VCV], wyz = a.na[OM_CVV];
        const unsigned fl = a.flags[c];
        if (fl & F_VCC)
            a.Vxn[c] = upd_V_ib(a.Vx[c], dif(wc[c] * P[c], wc[c - 1] * P[c - 1], ix), dif(wc[c] * txx[c], wc[c - 1] * txx[c - 1], ix),
                                dif(wxy[c + sy] * txy[c + sy], wxy[c] * txy[c], iy), dif(wxz[c + sz] * txz[c + sz], wxz[c] * txz[c], iz),
                                gx, ex, a.ns[OM_VCC][c], a.dV[0]);
        else if (!a.z_wvel)
            a.Vxn[c] = T(0);
        if (fl & F_CVC)
            a.Vyn[c] = upd_V_ib(a.Vy[c], dif(wc[c] * P[c], wc[c - sy] * P[c - sy], iy), dif(wc[c] * tyy[c], wc[c - sy] * tyy[c - sy], iy),
                                dif(wxy[c + 1] * txy[c + 1], wxy[c] * txy[c], ix), dif(wyz[c + sz] * tyz[c + sz], wyz[c] * tyz[c], iz),
                                gy, ey, a.ns[OM_CVC][c], a.dV[1]);
        else if (!a.z_wvel)
            a.Vyn[c] = T(0);
        if (fl & F_CCV)
            a.Vzn[c] = upd_V_ib(a.Vz[c], dif(wc[c] * P[c], wc[c - sz] * P[c - sz], iz), dif(wc[c] * tzz[c], wc[c - sz] * tzz[c - sz], iz),
                                dif(wxz[c + 1] * txz[c + 1], wxz[c] * txz[c], ix), dif(wyz[c + sy] * tyz[c + sy], wyz[c] * tyz[c], iy),
                                gz, ez, a.ns[OM_CCV][c], a.dV[2]);
        else if (!a.z_wvel)
            a.Vzn[c] = T(0);
    }
    if (a.frozen_eta) return;
    const T t = tauII(txx[c], tyy[c], tzz[c], lerp4(txy[c], txy[c + 1], txy[c + sy], txy[c + 1 + sy]),
                      lerp4(txz[c], txz[c + 1], txz[c + sz], txz[c + 1 + sz]),
                      lerp4(tyz[c], tyz[c + sy], tyz[c + sz], tyz[c + sy + sz]));
    a.eta_next[c] = rheology(a, c, t);
}

// point t of box b in flat (i fastest) order; false past its end (32-bit
// index arithmetic where the box holds fewer than 2^32 points)
__device__ __forceinline__ bool box_point(const Box& b, long t, int& i, int& j, int& k)
{
    const long nx = b.i1 - b.i0 + 1, ny = b.j1 - b.j0 + 1, nz = b.k1 - b.k0 + 1;
    const long n = nx * ny * nz;
    if (t >= n) return false;
    if (n <= 0xffffffffL) {
        const unsigned tu = (unsigned)t, nxu = (unsigned)nx, nyu = (unsigned)ny;
        const unsigned r = tu / nxu;
        i = b.i0 + (int)(tu - r * nxu);
        const unsigned s = r / nyu;
        j = b.j0 + (int)(r - s * nyu);
        k = b.k0 + (int)s;
    } else {
        i = b.i0 + (int)(t % nx);
        const long r = t / nx;
        j = b.j0 + (int)(r % ny);
        k = b.k0 + (int)(r / ny);
    }
    return true;
}

// flat (i fastest) indexing over the boxes of a list: every lane of a block
// is busy however thin the box, and wide boxes keep coalesced x rows
__device__ __forceinline__ bool box_index(const BoxList& bl, int& i, int& j, int& k)
{
    const int bid = blockIdx.x;
    int q = 0;
    while (q + 1 < bl.nb && bid >= bl.start[q + 1]) ++q;
    const long t = (long)(bid - bl.start[q]) * blockDim.x + threadIdx.x;
    return box_point(bl.b[q], t, i, j, k);
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_sigma_box(StokesArgs<T> a, BoxList bl)
{
    int i, j, k;
    if (!box_index(bl, i, j, k)) return;
    sigma_point3<T, IB>(a, (long)i + a.sy * j + a.sz * k, a.sy, a.sz);
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_velocity_box(StokesArgs<T> a, BoxList bl)
{
    int i, j, k;
    if (!box_index(bl, i, j, k)) return;
    velocity_point3<T, IB>(a, (long)i + a.sy * j + a.sz * k, a.sy, a.sz);
}

// ---------------------------------------------------------------------------
// fused 2.5D sweep (3D interior)
// ---------------------------------------------------------------------------
// A workgroup owns a (64-2) x (TY-2) column of the interior and marches up a
// chunk of z planes.  For plane kk it publishes the previous-iteration
// velocity and viscosity of its 64 x TY footprint to LDS, computes the new
// stress there (each thread only the components whose stencil lies inside
// the footprint: the footprint's outer ring supplies exactly the stress the
// inner velocity update needs), and updates velocity and viscosity at plane
// kk-1 of the inner threads.  Stress and velocity are double buffered, so
// every global read is of the previous iteration and the workgroups are
// independent.  Each field is read and written about once per iteration:
// z neighbours come from registers, x/y neighbours from LDS.
// Global accesses of the fused sweep use 32-bit byte offsets from the
// uniform field base (SGPR), which maps to gfx950's saddr + voffset
// addressing and keeps one VGPR per offset instead of a 64-bit address per
// field.  The launcher only takes this path when every field is < 4 GiB.
template <typename T>
__device__ __forceinline__ T ldo(const T* base, uint32_t bo)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + bo);
}
// a mask read at the byte offset of the FP32 fields it weights
__device__ __forceinline__ float ldo(H16Mask base, uint32_t bo) { return base[bo / sizeof(float)]; }
__device__ __forceinline__ uint4 ldo(const uint4* base, uint32_t bo)
{
    return *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(base) + bo);
}
template <typename T>
__device__ __forceinline__ void sto(T* base, uint32_t bo, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + bo) = v;
}

// x placement of the fused tiles.  A tile of TX lanes sweeps TX-2 inner
// columns with one halo column on each side.  With fg.nx0 = 1 the first tile
// covers the columns [lo, xa-1] and every further tile starts at
// xa + (bx-1)(TX-2); xa is chosen so that those tiles start on a 128-byte
// line (whole-line stores, no line shared by two tiles' stores).
__host__ __device__ __forceinline__ void fused_xrange(const FusedGeom& fg, int tx, int bx, int& xs, int& xe)
{
    if (bx < fg.nx0) {
        xs = fg.lo[0];
        xe = fg.xa - 1;
    } else {
        xs = fg.xa + (bx - fg.nx0) * (tx - 2);
        xe = min(xs + tx - 3, fg.hi[0]);
    }
}

// VISC false: a sweep with frozen viscosity (the FP32 correction sweeps),
// compiled without the viscosity update and its registers
template <typename T, bool IB, int TX, int TY, bool VISC>
__device__ __forceinline__ void fused3_tile(const StokesArgs<T>& a, const FusedGeom& fg, int bx, int by, int bz)
{
    __shared__ T sVx[TY][TX], sVy[TY][TX], sVz[TY][TX];
    __shared__ T sE[3][TY][TX];
    __shared__ T sS[2][6][TY][TX]; // P txx tyy txy txz tyz

    const int tx = threadIdx.x, ty = threadIdx.y;
    int xs, xe; // inner x range of this tile
    fused_xrange(fg, TX, bx, xs, xe);
    const int i = xs - 1 + tx;
    const int j = fg.lo[1] - 1 + by * (TY - 2) + ty;
    const int ks = fg.lo[2] + bz * fg.kch;
    const int ke = min(ks + fg.kch - 1, fg.hi[2]);
    const bool inner = tx >= 1 && tx <= TX - 2 && ty >= 1 && ty <= TY - 2 && i <= xe && j <= fg.hi[1];
    const bool can_c = tx < TX - 1 && ty < TY - 1;
    const bool can_xy = tx > 0 && ty > 0;
    const bool can_xz = tx > 0;
    const bool can_yz = ty > 0;
    constexpr uint32_t E = sizeof(T);
    const uint32_t bsy = (uint32_t)a.sy * E, bsz = (uint32_t)a.sz * E;
    // footprint positions past the inner range + 1 are never read by inner threads: clamp their addresses
    const int ic = min(i, xe + 1), jc = min(j, fg.hi[1] + 1);
    const uint32_t col = ((uint32_t)ic + (uint32_t)a.sy * (uint32_t)jc) * E;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    // old-stress components this thread's row needs: the halo rows only feed
    // the inner rows P, tyy (row 0) or txy, tyz (row TY-1); the halo planes
    // ks-1 and ke+1 only P, tzz and txz, tyz (bits P xx yy zz xy xz yz)
    const unsigned rneed = ty == 0 ? 0x05u : ty == TY - 1 ? 0x50u : 0x7fu;

    // registers of the own column
    const uint32_t c2 = col + bsz * (uint32_t)(ks - 2), c1 = c2 + bsz, c3 = c2 - bsz;
    T vxm = ldo(a.Vx, c2), vym = ldo(a.Vy, c2), vzm = ldo(a.Vz, c2);
    T vx0 = ldo(a.Vx, c1), vy0 = ldo(a.Vy, c1), vz0 = ldo(a.Vz, c1);
    T vzp = ldo(a.Vz, c1 + bsz);
    T emm = ldo(a.eta, c3), em = ldo(a.eta, c2), e0 = ldo(a.eta, c1);
    const unsigned n0 = rneed & 0x09u;
    T oP = (n0 & 1u) ? ldo(a.P, c1) : T(0), oxx = T(0), oyy = T(0), ozz = (n0 & 8u) ? ldo(a.tzz, c1) : T(0),
      oxy = T(0), oxz = T(0), oyz = T(0);
    T gx = T(0), gy = T(0), gz = T(0);
    T ofP = a.fP ? ldo(a.fP, c1) : T(0);
    T Pk = T(0), xxk = T(0), yyk = T(0), zzk = T(0), Pkm = T(0), zzkm = T(0);
    sE[(ks - 2) % 3][ty][tx] = em;

    for (int kk = ks - 1; kk <= ke + 1; ++kk) {
        const uint32_t c = col + bsz * (uint32_t)kk;
        // A: publish the plane-kk inputs
        sVx[ty][tx] = vx0;
        sVy[ty][tx] = vy0;
        sVz[ty][tx] = vz0;
        sE[kk % 3][ty][tx] = e0;
        // prefetch plane kk+1 (velocity z at kk+2, gravity at kk)
        T nvx = T(0), nvy = T(0), nvz2 = T(0), ne = T(0), nP = T(0), nxx = T(0), nyy = T(0), nzz = T(0), nxy = T(0),
          nxz = T(0), nyz = T(0), ngx = T(0), ngy = T(0), ngz = T(0), nfP = T(0);
        if (kk <= ke) {
            const uint32_t cn = c + bsz;
            if (a.fP) nfP = ldo(a.fP, cn);
            nvx = ldo(a.Vx, cn);
            nvy = ldo(a.Vy, cn);
            nvz2 = ldo(a.Vz, cn + bsz);
            ne = ldo(a.eta, cn);
            // unneeded components read the line of a needed one (no branch, no
            // extra HBM traffic, value never used)
            const unsigned nd = rneed & (kk + 1 == ke + 1 ? 0x60u : 0x7fu);
            const T* fb = (nd & 1u) ? a.P : (nd & 16u) ? a.txy : (nd & 32u) ? a.txz : (nd & 64u) ? a.tyz : a.P;
            nP = ldo((nd & 1u) ? a.P : fb, cn);
            nxx = ldo((nd & 2u) ? a.txx : fb, cn);
            nyy = ldo((nd & 4u) ? a.tyy : fb, cn);
            nzz = ldo((nd & 8u) ? a.tzz : fb, cn);
            nxy = ldo((nd & 16u) ? a.txy : fb, cn);
            nxz = ldo((nd & 32u) ? a.txz : fb, cn);
            nyz = ldo((nd & 64u) ? a.tyz : fb, cn);
        }
        if (a.grav_field) {
            ngx = ldo(a.gx, c);
            ngy = ldo(a.gy, c);
            ngz = ldo(a.gz, c);
        }
        __syncthreads();

        // B: new stress at plane kk
        const int e_k = kk % 3, e_km = (kk + 2) % 3;
        T Pc = T(0), xxc = T(0), yyc = T(0), zzc = T(0), xyc = T(0), xzc = T(0), yzc = T(0);
        if (!IB) {
            if (can_c) {
                T exx = dif(sVx[ty][tx + 1], vx0, ix);
                T eyy = dif(sVy[ty + 1][tx], vy0, iy);
                T ezz = dif(vzp, vz0, iz);
                T divV = exx + eyy + ezz;
                Pc = upd_P(oP, a.fP ? divV + ofP : divV, e0, a.dPr);
                xxc = upd_td(oxx, e0, exx, divV, a.dtau[0]);
                yyc = upd_td(oyy, e0, eyy, divV, a.dtau[1]);
                zzc = upd_td(ozz, e0, ezz, divV, a.dtau[2]);
            }
            if (can_xy) {
                T exy = T(0.5) * (dif(vy0, sVy[ty][tx - 1], ix) + dif(vx0, sVx[ty - 1][tx], iy));
                T el = lerp4(sE[e_k][ty - 1][tx - 1], sE[e_k][ty - 1][tx], sE[e_k][ty][tx - 1], e0);
                xyc = upd_to(oxy, el, exy, a.dtau[3]);
            }
            if (can_xz) {
                T exz = T(0.5) * (dif(vz0, sVz[ty][tx - 1], ix) + dif(vx0, vxm, iz));
                T el = lerp4(sE[e_km][ty][tx - 1], em, sE[e_k][ty][tx - 1], e0);
                xzc = upd_to(oxz, el, exz, a.dtau[4]);
            }
            if (can_yz) {
                T eyz = T(0.5) * (dif(vz0, sVz[ty - 1][tx], iy) + dif(vy0, vym, iz));
                T el = lerp4(sE[e_km][ty - 1][tx], em, sE[e_k][ty - 1][tx], e0);
                yzc = upd_to(oyz, el, eyz, a.dtau[5]);
            }
        } else {
            const auto wx = a.ns[OM_VCC], wy = a.ns[OM_CVC], wz = a.ns[OM_CCV];
            const unsigned fl = ldo(a.flags, c / E);
            const T wx0 = ldo(wx, c), wy0 = ldo(wy, c), wz0 = ldo(wz, c);
            if (can_c && (fl & F_CCC)) {
                T exx = dif(ldo(wx, c + E) * sVx[ty][tx + 1], wx0 * vx0, ix);
                T eyy = dif(ldo(wy, c + bsy) * sVy[ty + 1][tx], wy0 * vy0, iy);
                T ezz = dif(ldo(wz, c + bsz) * vzp, wz0 * vz0, iz);
                T divV = exx + eyy + ezz;
                Pc = upd_P(oP, a.fP ? divV + ofP : divV, e0, a.dPr);
                xxc = upd_td(oxx, e0, exx, divV, a.dtau[0]);
                yyc = upd_td(oyy, e0, eyy, divV, a.dtau[1]);
                zzc = upd_td(ozz, e0, ezz, divV, a.dtau[2]);
            }
            if (can_xy && (fl & F_VVC)) {
                T exy = T(0.5) * (dif(wy0 * vy0, ldo(wy, c - E) * sVy[ty][tx - 1], ix) + dif(wx0 * vx0, ldo(wx, c - bsy) * sVx[ty - 1][tx], iy));
                T el = lerp4(sE[e_k][ty - 1][tx - 1], sE[e_k][ty - 1][tx], sE[e_k][ty][tx - 1], e0);
                xyc = upd_to(oxy, el, exy, a.dtau[3]);
            }
            if (can_xz && (fl & F_VCV)) {
                T exz = T(0.5) * (dif(wz0 * vz0, ldo(wz, c - E) * sVz[ty][tx - 1], ix) + dif(wx0 * vx0, ldo(wx, c - bsz) * vxm, iz));
                T el = lerp4(sE[e_km][ty][tx - 1], em, sE[e_k][ty][tx - 1], e0);
                xzc = upd_to(oxz, el, exz, a.dtau[4]);
            }
            if (can_yz && (fl & F_CVV)) {
                T eyz = T(0.5) * (dif(wz0 * vz0, ldo(wz, c - bsy) * sVz[ty - 1][tx], iy) + dif(wy0 * vy0, ldo(wy, c - bsz) * vym, iz));
                T el = lerp4(sE[e_km][ty - 1][tx], em, sE[e_k][ty - 1][tx], e0);
                yzc = upd_to(oyz, el, eyz, a.dtau[5]);
            }
        }
        const int b = kk & 1;
        sS[b][0][ty][tx] = Pc;
        sS[b][1][ty][tx] = xxc;
        sS[b][2][ty][tx] = yyc;
        sS[b][3][ty][tx] = xyc;
        sS[b][4][ty][tx] = xzc;
        sS[b][5][ty][tx] = yzc;
        if (inner && kk >= ks && kk <= ke) {
            sto(a.Pn, c, Pc);
            sto(a.txxn, c, xxc);
            sto(a.tyyn, c, yyc);
            sto(a.tzzn, c, zzc);
            sto(a.txyn, c, xyc);
            sto(a.txzn, c, xzc);
            sto(a.tyzn, c, yzc);
        }
        __syncthreads();

        // C: velocity and viscosity at plane k = kk-1
        const int k = kk - 1;
        if (inner && k >= ks) {
            const uint32_t cv = c - bsz;
            const int bo = k & 1, bn = kk & 1, ek = k % 3;
            const T P_ = Pk, P_im = sS[bo][0][ty][tx - 1], P_jm = sS[bo][0][ty - 1][tx];
            const T xx_im = sS[bo][1][ty][tx - 1], yy_jm = sS[bo][2][ty - 1][tx];
            const T xy = sS[bo][3][ty][tx], xy_ip = sS[bo][3][ty][tx + 1], xy_jp = sS[bo][3][ty + 1][tx],
                    xy_ijp = sS[bo][3][ty + 1][tx + 1];
            const T xz = sS[bo][4][ty][tx], xz_ip = sS[bo][4][ty][tx + 1], xz_kp = sS[bn][4][ty][tx],
                    xz_ikp = sS[bn][4][ty][tx + 1];
            const T yz = sS[bo][5][ty][tx], yz_jp = sS[bo][5][ty + 1][tx], yz_kp = sS[bn][5][ty][tx],
                    yz_jkp = sS[bn][5][ty + 1][tx];
            const T ex = T(0.5) * (sE[ek][ty][tx - 1] + em);
            const T ey = T(0.5) * (sE[ek][ty - 1][tx] + em);
            const T ez = T(0.5) * (emm + em);
            const T g0 = a.grav_field ? gx : a.grav[0];
            const T g1 = a.grav_field ? gy : a.grav[1];
            const T g2 = a.grav_field ? gz : a.grav[2];
            if (!IB) {
                sto(a.Vxn, cv, upd_V(vxm, dif(P_, P_im, ix), dif(xxk, xx_im, ix), dif(xy_jp, xy, iy), dif(xz_kp, xz, iz), g0, ex, a.dV[0]));
                sto(a.Vyn, cv, upd_V(vym, dif(P_, P_jm, iy), dif(yyk, yy_jm, iy), dif(xy_ip, xy, ix), dif(yz_kp, yz, iz), g1, ey, a.dV[1]));
                sto(a.Vzn, cv, upd_V(vzm, dif(P_, Pkm, iz), dif(zzk, zzkm, iz), dif(xz_ip, xz, ix), dif(yz_jp, yz, iy), g2, ez, a.dV[2]));
            } else {
                const auto wc = a.na[OM_CCC], wxy = a.na[OM_VVC], wxz = a.na[OM_VCV], wyz = a.na[OM_CVV];
                const unsigned fl = ldo(a.flags, cv / E);
                const T wc0 = ldo(wc, cv), wxy0 = ldo(wxy, cv), wxz0 = ldo(wxz, cv), wyz0 = ldo(wyz, cv);
                sto(a.Vxn, cv, (fl & F_VCC) ? upd_V_ib(vxm, dif(wc0 * P_, ldo(wc, cv - E) * P_im, ix), dif(wc0 * xxk, ldo(wc, cv - E) * xx_im, ix),
                                                      dif(ldo(wxy, cv + bsy) * xy_jp, wxy0 * xy, iy), dif(ldo(wxz, cv + bsz) * xz_kp, wxz0 * xz, iz),
                                                      g0, ex, ldo(a.ns[OM_VCC], cv), a.dV[0])
                                           : T(0));
                sto(a.Vyn, cv, (fl & F_CVC) ? upd_V_ib(vym, dif(wc0 * P_, ldo(wc, cv - bsy) * P_jm, iy), dif(wc0 * yyk, ldo(wc, cv - bsy) * yy_jm, iy),
                                                      dif(ldo(wxy, cv + E) * xy_ip, wxy0 * xy, ix), dif(ldo(wyz, cv + bsz) * yz_kp, wyz0 * yz, iz),
                                                      g1, ey, ldo(a.ns[OM_CVC], cv), a.dV[1])
                                           : T(0));
                sto(a.Vzn, cv, (fl & F_CCV) ? upd_V_ib(vzm, dif(wc0 * P_, ldo(wc, cv - bsz) * Pkm, iz), dif(wc0 * zzk, ldo(wc, cv - bsz) * zzkm, iz),
                                                      dif(ldo(wxz, cv + E) * xz_ip, wxz0 * xz, ix), dif(ldo(wyz, cv + bsy) * yz_jp, wyz0 * yz, iy),
                                                      g2, ez, ldo(a.ns[OM_CCV], cv), a.dV[2])
                                           : T(0));
            }
            if (VISC && !a.frozen_eta) {
                const T t = tauII(xxk, yyk, zzk, lerp4(xy, xy_ip, xy_jp, xy_ijp), lerp4(xz, xz_ip, xz_kp, xz_ikp),
                                  lerp4(yz, yz_jp, yz_kp, yz_jkp));
                sto(a.eta_next, cv, rheology_o(a, cv, t, em));
            }
        }

        // rotate the column registers
        Pkm = Pk;
        zzkm = zzk;
        Pk = Pc;
        xxk = xxc;
        yyk = yyc;
        zzk = zzc;
        vxm = vx0;
        vym = vy0;
        vzm = vz0;
        vx0 = nvx;
        vy0 = nvy;
        vz0 = vzp;
        vzp = nvz2;
        emm = em;
        em = e0;
        e0 = ne;
        oP = nP;
        oxx = nxx;
        oyy = nyy;
        ozz = nzz;
        oxy = nxy;
        oxz = nxz;
        oyz = nyz;
        ofP = nfP;
        gx = ngx;
        gy = ngy;
        gz = ngz;
    }
}

// Immersed-boundary fused sweep.  Same marching scheme as k_fused3; every
// thread loads the seven mask weights of its own position once per plane and
// publishes weighted values to LDS: ns-weighted velocity for the masked strain
// rates and na-weighted stress for the masked momentum fluxes (the products
// are the operands the pointwise kernels form, so results are identical).
// Raw off-diagonal stress is kept beside it for tau_II when eta is updated.
//
// Empty space is not read.  A thread reads the old stress of a location only
// where the location is valid (elsewhere the update is 0 whatever it holds),
// and the flags and na weights of a cell only on the planes of its column
// that hold a valid location (crange: elsewhere both are unused, every new
// stress there being 0).  With frozen eta (FP32 correction sweeps) the
// viscosity is read only on the planes a valid location may need (erange).
// Velocity and its ns weight are skipped outside crange only when z_read
// says that the current velocity is +0 at every invalid location (then the
// weighted value is +0 either way), and the new state at invalid locations is
// stored only when the output buffers are not known to hold +0 there already
// (z_wsig, z_wvel).  Every value the sweep produces is therefore the one of
// the full sweep, bit for bit.
template <typename P>
__device__ __forceinline__ auto ldif(bool p, P base, uint32_t bo) -> decltype(ldo(base, bo))
{
    using R = decltype(ldo(base, bo));
    return p ? ldo(base, bo) : R(0);
}

__device__ __forceinline__ uint4 ldw(bool p, const uint4* base, uint32_t bo)
{
    return p ? *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(base) + bo) : make_uint4(0u, 0u, 0u, 0u);
}

// ETA_HERE: packed launches, which keep no march table, update the viscosity
// of the planes a narrowed sweep does not march inside the sweep
template <typename T, int TX, int TY, bool VISC, bool ETA_HERE>
__device__ __forceinline__ void fused3_ib_tile(const StokesArgs<T>& a, const FusedGeom& fg, int bx, int by, int bz)
{
    // FP32 sweeps read masks and flags from the packed words (a.mw)
    constexpr bool PACKED = sizeof(T) == 4;
    // the FP32 correction sweep (no viscosity update) reads the old stress
    // and forcing a plane ahead and needs one barrier per plane
    constexpr bool AHEAD = PACKED && !VISC;
    constexpr int NV = AHEAD ? 2 : 1;
    __shared__ T sVx[NV][TY][TX], sVy[NV][TY][TX], sVz[NV][TY][TX]; // by plane parity (AHEAD)
    __shared__ T sE[3][TY][TX];
    __shared__ T sS[2][VISC ? 9 : 6][TY][TX]; // na-weighted P txx tyy txy txz tyz; raw txy txz tyz

    const int tx = threadIdx.x, ty = threadIdx.y;
    int xs, xe; // inner x range of this tile
    fused_xrange(fg, TX, bx, xs, xe);
    const int i = xs - 1 + tx;
    const int j = fg.lo[1] - 1 + by * (TY - 2) + ty;
    const int ks0 = fg.lo[2] + bz * fg.kch;
    const int ke0 = min(ks0 + fg.kch - 1, fg.hi[2]);
    const bool inner = tx >= 1 && tx <= TX - 2 && ty >= 1 && ty <= TY - 2 && i <= xe && j <= fg.hi[1];
    const bool can_c = tx < TX - 1 && ty < TY - 1;
    const bool can_xy = tx > 0 && ty > 0;
    const bool can_xz = tx > 0;
    const bool can_yz = ty > 0;
    const bool need_raw = VISC && !a.frozen_eta;
    constexpr uint32_t E = sizeof(T);
    const uint32_t bsz = (uint32_t)a.sz * E;
    const int ic = min(i, xe + 1), jc = min(j, fg.hi[1] + 1);
    const uint32_t col = ((uint32_t)ic + (uint32_t)a.sy * (uint32_t)jc) * E;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const auto wvx = a.ns[OM_VCC], wvy = a.ns[OM_CVC], wvz = a.ns[OM_CCV];
    const auto wc = a.na[OM_CCC], wxy = a.na[OM_VVC], wxz = a.na[OM_VCV], wyz = a.na[OM_CVV];
    // old-stress components this thread's row needs (as in fused3_tile); with
    // 64 lanes a row is one wavefront, and the mask is wavefront-uniform
    const int tyw = TX == 64 ? __builtin_amdgcn_readfirstlane(ty) : ty;
    const unsigned rneed = tyw == 0 ? 0x05u : tyw == TY - 1 ? 0x50u : 0x7fu;
    // planes of this column with a valid location, and with a needed viscosity
    const int2 rg = a.crange[(long)ic + a.sy * (long)jc];
    const int2 re = a.erange[(long)ic + a.sy * (long)jc];
    const bool zr = a.z_read != 0, ws = a.z_wsig == 0, wv = a.z_wvel == 0;
    auto inr = [&](int p) { return p >= rg.x && p <= rg.y; };
    auto inv = [&](int p) { return !zr || inr(p); };
    auto ine = [&](int p) { return need_raw || (p >= re.x && p <= re.y); };

    // Planes of the chunk whose new state can differ from 0: the tile's
    // footprint holds valid locations only on [tlo, thi].  Once the output
    // buffers are known to be +0 at invalid locations, the march covers
    // [max(ks0, tlo - 1), min(ke0, thi)] only; the planes of the chunk outside
    // it keep their stress and velocity (0), and with a viscosity update their
    // viscosity goes through the rheology at tau_II = 0 (every stress around
    // such a cell, on its plane and the next, is at an invalid location).
    __shared__ int s_lo, s_hi;
    if (tx == 0 && ty == 0) {
        s_lo = RANGE_NONE;
        s_hi = -RANGE_NONE;
    }
    __syncthreads();
    if (rg.x <= rg.y) {
        atomicMin(&s_lo, rg.x);
        atomicMax(&s_hi, rg.y);
    }
    __syncthreads();
    // With a viscosity update, the planes outside the march are left to
    // k_eta_rest, which reads the march of every tile from a.tmarch (single
    // models); packed launches (ETA_HERE) update them here.
    const bool narrow = !ws && !wv && (!need_raw || ETA_HERE || a.tmarch != nullptr);
    const int ks = narrow ? max(ks0, s_lo - 1) : ks0;
    const int ke = narrow ? min(ke0, s_hi) : ke0;
    if constexpr (VISC && !ETA_HERE) {
        if (narrow && need_raw && tx == 0 && ty == 0) {
            const int nby = (fg.hi[1] - fg.lo[1] + 1 + TY - 3) / (TY - 2);
            a.tmarch[bx + fg.nbx * (by + nby * bz)] = make_int2(ks, ke);
        }
    }
    if constexpr (VISC && ETA_HERE) {
        if (narrow && need_raw) {
            if (inner) {
                // With a constant-coefficient rheology the result depends on
                // the old viscosity alone, and the unmarched planes of a
                // column (empty space) mostly hold the same one: the last
                // input and result are kept, and a cell holding the same bits
                // takes the result without a log and an exp.
                const bool memo = a.rh == nullptr;
                T m_in = T(0), m_out = T(0);
                bool m_set = false;
                auto eta_only = [&](int k0, int k1) {
                    for (int k = k0; k <= k1; k += 4) {
                        T o[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            o[u] = k + u <= k1 ? ldo(a.eta, col + bsz * (uint32_t)(k + u)) : T(1);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (k + u <= k1) {
                                const uint32_t cu = col + bsz * (uint32_t)(k + u);
                                T en;
                                if (memo && m_set && same_bits(o[u], m_in)) {
                                    en = m_out;
                                } else {
                                    en = rheology_o(a, cu, T(0), o[u]);
                                    m_in = o[u];
                                    m_out = en;
                                    m_set = true;
                                }
                                sto(a.eta_next, cu, en);
                            }
                    }
                };
                if (ks > ke) {
                    eta_only(ks0, ke0);
                } else {
                    eta_only(ks0, ks - 1);
                    eta_only(ke + 1, ke0);
                }
            }
        }
    }
    if (ks > ke) return;

    const uint32_t c2 = col + bsz * (uint32_t)(ks - 2), c1 = c2 + bsz, c0 = c1 + bsz, c3 = c2 - bsz;
    // The two buffers of a stress or velocity field are one allocation, the
    // second a fixed number of bytes after the first (the same for every
    // field), so the sweep reads and writes a field through one base pointer,
    // with one input and one output byte offset for the stress fields and
    // for the velocity fields: fewer pointers held across the march.
    auto lo = [](const T* x, const T* y) { return x < y ? x : y; };
    const T *bP = lo(a.P, a.Pn), *bxx = lo(a.txx, a.txxn), *byy = lo(a.tyy, a.tyyn), *bzz = lo(a.tzz, a.tzzn);
    const T *bxy = lo(a.txy, a.txyn), *bxz = lo(a.txz, a.txzn), *byz = lo(a.tyz, a.tyzn);
    const T *bvx = lo(a.Vx, a.Vxn), *bvy = lo(a.Vy, a.Vyn), *bvz = lo(a.Vz, a.Vzn);
    T *uP = const_cast<T*>(bP), *uxx = const_cast<T*>(bxx), *uyy = const_cast<T*>(byy), *uzz = const_cast<T*>(bzz);
    T *uxy = const_cast<T*>(bxy), *uxz = const_cast<T*>(bxz), *uyz = const_cast<T*>(byz);
    T *uvx = const_cast<T*>(bvx), *uvy = const_cast<T*>(bvy), *uvz = const_cast<T*>(bvz);
    auto db = [](const T* x, const T* b) { return (uint32_t)((const char*)x - (const char*)b); };
    const uint32_t osi = db(a.P, bP), oso = db(a.Pn, bP), ovi = db(a.Vx, bvx), ovo = db(a.Vxn, bvx);
    // raw velocity and ns weights of the own column
    T vxm = ldif(inv(ks - 2), bvx, c2 + ovi), vym = ldif(inv(ks - 2), bvy, c2 + ovi), vzm = ldif(inv(ks - 2), bvz, c2 + ovi);
    T vx0 = ldif(inv(ks - 1), bvx, c1 + ovi), vy0 = ldif(inv(ks - 1), bvy, c1 + ovi), vz0 = ldif(inv(ks - 1), bvz, c1 + ovi);
    T vzp = ldif(inv(ks), bvz, c0 + ovi);
    T mxm, mym, mzm, mx0, my0, mz0, mzp, mc0, mxy0, mxz0, myz0;
    // flags of planes k-1 and k
    unsigned flm, fl0;
    // FP32: the mask word of the plane after the current one (read a plane ahead)
    uint4 wn1 = make_uint4(0u, 0u, 0u, 0u);
    if constexpr (PACKED) {
        // a word read where the velocity is (a superset of crange): outside
        // crange its flags are 0 and its na weights multiply 0, as above
        const uint4 w2 = ldw(inv(ks - 2), a.mw, 4 * c2), w1 = ldw(inv(ks - 1), a.mw, 4 * c1);
        wn1 = ldw(inv(ks), a.mw, 4 * c0);
        mxm = mw_mask(w2, MW_NS_VCC), mym = mw_mask(w2, MW_NS_CVC), mzm = mw_mask(w2, MW_NS_CCV);
        flm = mw_flags(w2);
        mx0 = mw_mask(w1, MW_NS_VCC), my0 = mw_mask(w1, MW_NS_CVC), mz0 = mw_mask(w1, MW_NS_CCV);
        mzp = mw_mask(wn1, MW_NS_CCV);
        mc0 = mw_mask(w1, MW_NA_CCC), mxy0 = mw_mask(w1, MW_NA_VVC), mxz0 = mw_mask(w1, MW_NA_VCV),
        myz0 = mw_mask(w1, MW_NA_CVV);
        fl0 = mw_flags(w1);
    } else {
        mxm = ldif(inv(ks - 2), wvx, c2), mym = ldif(inv(ks - 2), wvy, c2), mzm = ldif(inv(ks - 2), wvz, c2);
        flm = ldif(inr(ks - 2), a.flags, c2 / E);
        mx0 = ldif(inv(ks - 1), wvx, c1), my0 = ldif(inv(ks - 1), wvy, c1), mz0 = ldif(inv(ks - 1), wvz, c1);
        mzp = ldif(inv(ks), wvz, c0);
        mc0 = ldif(inr(ks - 1), wc, c1), mxy0 = ldif(inr(ks - 1), wxy, c1), mxz0 = ldif(inr(ks - 1), wxz, c1),
        myz0 = ldif(inr(ks - 1), wyz, c1);
        fl0 = ldif(inr(ks - 1), a.flags, c1 / E);
    }
    // FP32 correction sweep (AHEAD): the old stress and the continuity forcing
    // of a plane are read one plane ahead of the plane's update (otherwise in
    // the plane's own iteration), and the gravity of a velocity location one
    // plane ahead of the velocity update, so that their latency overlaps the
    // plane in between.  Stress components are read where the tile needs them
    // at a valid location (the halo planes ks-1 and ke+1 only feed P, tzz and
    // txz, tyz): one branch per location class (the needed components are
    // uniform over the row).
    const bool gf = a.grav_field != 0 && inner;
    T oP = T(0), oxx = T(0), oyy = T(0), ozz = T(0), oxy = T(0), oxz = T(0), oyz = T(0), ofP = T(0);
    auto ld_stress = [&](unsigned need, unsigned f, uint32_t cq, T& P_, T& xx_, T& yy_, T& zz_, T& xy_, T& xz_, T& yz_,
                         T& fP_) {
        if (can_c && (f & F_CCC)) {
            if (need & 1u) P_ = ldo(bP, cq + osi);
            if (need & 2u) xx_ = ldo(bxx, cq + osi);
            if (need & 4u) yy_ = ldo(byy, cq + osi);
            if (need & 8u) zz_ = ldo(bzz, cq + osi);
            if (a.fP) fP_ = ldo(a.fP, cq);
        }
        if ((need & 16u) && can_xy && (f & F_VVC)) xy_ = ldo(bxy, cq + osi);
        if ((need & 32u) && can_xz && (f & F_VCV)) xz_ = ldo(bxz, cq + osi);
        if ((need & 64u) && can_yz && (f & F_CVV)) yz_ = ldo(byz, cq + osi);
    };
    if constexpr (AHEAD) ld_stress(rneed & 0x09u, fl0, c1, oP, oxx, oyy, ozz, oxy, oxz, oyz, ofP);
    T gvx = T(0), gvy = T(0), gvz = T(0);
    T qxm = mxm * vxm, qym = mym * vym;
    T qx0 = mx0 * vx0, qy0 = my0 * vy0, qz0 = mz0 * vz0;
    T qzp = mzp * vzp;
    T emm = ldif(ine(ks - 3), a.eta, c3), em = ldif(ine(ks - 2), a.eta, c2), e0 = ldif(ine(ks - 1), a.eta, c1);
    // own new stress: weighted at k and k-1, raw diagonal at k
    T Pw = T(0), Pwm = T(0), xxw = T(0), yyw = T(0), zzw = T(0), zzwm = T(0), xxr = T(0), yyr = T(0), zzr = T(0);
    sE[(ks - 2) % 3][ty][tx] = em;

    for (int kk = ks - 1; kk <= ke + 1; ++kk) {
        const uint32_t c = col + bsz * (uint32_t)kk;
        // plane kk+1: its old stress and forcing; plane kk: the gravity of
        // its velocity locations (updated in the next iteration)
        T nP = T(0), nxx = T(0), nyy = T(0), nzz = T(0), nxy = T(0), nxz = T(0), nyz = T(0), nfP = T(0);
        if constexpr (AHEAD) {
            if (kk <= ke)
                ld_stress(rneed & (kk == ke ? 0x60u : 0x7fu), mw_flags(wn1), c + bsz, nP, nxx, nyy, nzz, nxy, nxz, nyz,
                          nfP);
        } else {
            // the components this thread computes at a valid location that the
            // tile needs, one predicated load each
            const unsigned rcan =
                (can_c ? 0x0fu : 0u) | (can_xy ? 0x10u : 0u) | (can_xz ? 0x20u : 0u) | (can_yz ? 0x40u : 0u);
            const unsigned valid = (fl0 & F_CCC ? 0x0fu : 0u) | (fl0 & F_VVC ? 0x10u : 0u) |
                                   (fl0 & F_VCV ? 0x20u : 0u) | (fl0 & F_CVV ? 0x40u : 0u);
            const unsigned nd = rneed & rcan & valid & (kk == ks - 1 ? 0x09u : kk == ke + 1 ? 0x60u : 0x7fu);
            const uint32_t cs = c + osi;
            oP = ldif(nd & 1u, bP, cs), oxx = ldif(nd & 2u, bxx, cs), oyy = ldif(nd & 4u, byy, cs);
            ozz = ldif(nd & 8u, bzz, cs), oxy = ldif(nd & 16u, bxy, cs), oxz = ldif(nd & 32u, bxz, cs);
            oyz = ldif(nd & 64u, byz, cs);
            ofP = ldif(a.fP != nullptr && (nd & 1u), a.fP, c);
        }
        const bool gk = gf && kk >= ks && kk <= ke;
        const T ngx = ldif(gk && (fl0 & F_VCC), a.gx, c), ngy = ldif(gk && (fl0 & F_CVC), a.gy, c),
                ngz = ldif(gk && (fl0 & F_CCV), a.gz, c);
        const int pv = AHEAD ? (kk & 1) : 0; // velocity plane buffer
        sVx[pv][ty][tx] = qx0;
        sVy[pv][ty][tx] = qy0;
        sVz[pv][ty][tx] = qz0;
        sE[kk % 3][ty][tx] = e0;
        T nvx = T(0), nvy = T(0), nvz2 = T(0), nmx = T(0), nmy = T(0), nmz2 = T(0), nmc = T(0), nmxy = T(0),
          nmxz = T(0), nmyz = T(0), ne = T(0);
        unsigned nfl = 0;
        if (kk <= ke) {
            const uint32_t cn = c + bsz;
            const bool v1 = inv(kk + 1), v2 = inv(kk + 2);
            if (v1) {
                nvx = ldo(bvx, cn + ovi);
                nvy = ldo(bvy, cn + ovi);
            }
            if constexpr (PACKED) {
                uint4 wn2 = make_uint4(0u, 0u, 0u, 0u);
                if (v2) {
                    nvz2 = ldo(bvz, cn + bsz + ovi);
                    wn2 = ldo(a.mw, 4 * (cn + bsz));
                }
                nmx = mw_mask(wn1, MW_NS_VCC);
                nmy = mw_mask(wn1, MW_NS_CVC);
                nmc = mw_mask(wn1, MW_NA_CCC);
                nmxy = mw_mask(wn1, MW_NA_VVC);
                nmxz = mw_mask(wn1, MW_NA_VCV);
                nmyz = mw_mask(wn1, MW_NA_CVV);
                nfl = mw_flags(wn1);
                nmz2 = mw_mask(wn2, MW_NS_CCV);
                wn1 = wn2;
            } else {
                const bool r1 = inr(kk + 1);
                if (v1) {
                    nmx = ldo(wvx, cn);
                    nmy = ldo(wvy, cn);
                }
                if (v2) {
                    nvz2 = ldo(bvz, cn + bsz + ovi);
                    nmz2 = ldo(wvz, cn + bsz);
                }
                if (r1) {
                    nmc = ldo(wc, cn);
                    nmxy = ldo(wxy, cn);
                    nmxz = ldo(wxz, cn);
                    nmyz = ldo(wyz, cn);
                    nfl = ldo(a.flags, cn / E);
                }
            }
            ne = ldif(ine(kk + 1), a.eta, cn);
        }
        __syncthreads();

        // new stress at plane kk (IB kernels_3d.jl:6-43)
        const int e_k = kk % 3, e_km = (kk + 2) % 3;
        T Pc = T(0), xxc = T(0), yyc = T(0), zzc = T(0), xyc = T(0), xzc = T(0), yzc = T(0);
        if (can_c && (fl0 & F_CCC)) {
            T exx = dif(sVx[pv][ty][tx + 1], qx0, ix);
            T eyy = dif(sVy[pv][ty + 1][tx], qy0, iy);
            T ezz = dif(qzp, qz0, iz);
            T divV = exx + eyy + ezz;
            Pc = upd_P(oP, a.fP ? divV + ofP : divV, e0, a.dPr);
            xxc = upd_td(oxx, e0, exx, divV, a.dtau[0]);
            yyc = upd_td(oyy, e0, eyy, divV, a.dtau[1]);
            zzc = upd_td(ozz, e0, ezz, divV, a.dtau[2]);
        }
        if (can_xy && (fl0 & F_VVC)) {
            T exy = T(0.5) * (dif(qy0, sVy[pv][ty][tx - 1], ix) + dif(qx0, sVx[pv][ty - 1][tx], iy));
            T el = lerp4(sE[e_k][ty - 1][tx - 1], sE[e_k][ty - 1][tx], sE[e_k][ty][tx - 1], e0);
            xyc = upd_to(oxy, el, exy, a.dtau[3]);
        }
        if (can_xz && (fl0 & F_VCV)) {
            T exz = T(0.5) * (dif(qz0, sVz[pv][ty][tx - 1], ix) + dif(qx0, qxm, iz));
            T el = lerp4(sE[e_km][ty][tx - 1], em, sE[e_k][ty][tx - 1], e0);
            xzc = upd_to(oxz, el, exz, a.dtau[4]);
        }
        if (can_yz && (fl0 & F_CVV)) {
            T eyz = T(0.5) * (dif(qz0, sVz[pv][ty - 1][tx], iy) + dif(qy0, qym, iz));
            T el = lerp4(sE[e_km][ty - 1][tx], em, sE[e_k][ty - 1][tx], e0);
            yzc = upd_to(oyz, el, eyz, a.dtau[5]);
        }
        const int b = kk & 1;
        const T Pwc = mc0 * Pc, zzwc = mc0 * zzc;
        sS[b][0][ty][tx] = Pwc;
        sS[b][1][ty][tx] = mc0 * xxc;
        sS[b][2][ty][tx] = mc0 * yyc;
        sS[b][3][ty][tx] = mxy0 * xyc;
        sS[b][4][ty][tx] = mxz0 * xzc;
        sS[b][5][ty][tx] = myz0 * yzc;
        if constexpr (VISC) {
            if (need_raw) {
                sS[b][6][ty][tx] = xyc;
                sS[b][7][ty][tx] = xzc;
                sS[b][8][ty][tx] = yzc;
            }
        }
        if (inner && kk >= ks && kk <= ke) {
            if (ws || (fl0 & F_CCC)) {
                sto(uP, c + oso, Pc);
                sto(uxx, c + oso, xxc);
                sto(uyy, c + oso, yyc);
                sto(uzz, c + oso, zzc);
            }
            if (ws || (fl0 & F_VVC)) sto(uxy, c + oso, xyc);
            if (ws || (fl0 & F_VCV)) sto(uxz, c + oso, xzc);
            if (ws || (fl0 & F_CVV)) sto(uyz, c + oso, yzc);
        }
        // The velocity at plane k = kk-1 reads the neighbours' stress of plane
        // k, published before this iteration's barrier, and its own stress of
        // plane kk; only tau_II (viscosity update) reads neighbours' stress of
        // plane kk.  Without it (AHEAD) the iteration needs one barrier: the
        // stress buffer of a parity is written after the barrier that follows
        // its last read (by the velocity of the plane before), and the
        // velocity planes are double buffered the same way.
        if constexpr (!AHEAD) __syncthreads();

        // velocity (IB kernels_3d.jl:50-67) and viscosity at plane k = kk-1
        const int k = kk - 1;
        if (inner && k >= ks) {
            const uint32_t cv = c - bsz;
            const int bo = k & 1, bn = kk & 1, ek = k % 3;
            const T ex = T(0.5) * (sE[ek][ty][tx - 1] + em);
            const T ey = T(0.5) * (sE[ek][ty - 1][tx] + em);
            const T ez = T(0.5) * (emm + em);
            const T xyw = sS[bo][3][ty][tx], xzw = sS[bo][4][ty][tx], yzw = sS[bo][5][ty][tx];
            if (flm & F_VCC) {
                const T g0 = a.grav_field ? gvx : a.grav[0];
                sto(uvx, cv + ovo, upd_V_ib(vxm, dif(Pw, sS[bo][0][ty][tx - 1], ix), dif(xxw, sS[bo][1][ty][tx - 1], ix),
                                        dif(sS[bo][3][ty + 1][tx], xyw, iy), dif(sS[bn][4][ty][tx], xzw, iz), g0, ex,
                                        mxm, a.dV[0]));
            } else if (wv) {
                sto(uvx, cv + ovo, T(0));
            }
            if (flm & F_CVC) {
                const T g1 = a.grav_field ? gvy : a.grav[1];
                sto(uvy, cv + ovo, upd_V_ib(vym, dif(Pw, sS[bo][0][ty - 1][tx], iy), dif(yyw, sS[bo][2][ty - 1][tx], iy),
                                        dif(sS[bo][3][ty][tx + 1], xyw, ix), dif(sS[bn][5][ty][tx], yzw, iz), g1, ey,
                                        mym, a.dV[1]));
            } else if (wv) {
                sto(uvy, cv + ovo, T(0));
            }
            if (flm & F_CCV) {
                const T g2 = a.grav_field ? gvz : a.grav[2];
                sto(uvz, cv + ovo, upd_V_ib(vzm, dif(Pw, Pwm, iz), dif(zzw, zzwm, iz), dif(sS[bo][4][ty][tx + 1], xzw, ix),
                                        dif(sS[bo][5][ty + 1][tx], yzw, iy), g2, ez, mzm, a.dV[2]));
            } else if (wv) {
                sto(uvz, cv + ovo, T(0));
            }
            if constexpr (VISC) {
                if (need_raw) {
                    const T t = tauII(xxr, yyr, zzr,
                                      lerp4(sS[bo][6][ty][tx], sS[bo][6][ty][tx + 1], sS[bo][6][ty + 1][tx], sS[bo][6][ty + 1][tx + 1]),
                                      lerp4(sS[bo][7][ty][tx], sS[bo][7][ty][tx + 1], sS[bn][7][ty][tx], sS[bn][7][ty][tx + 1]),
                                      lerp4(sS[bo][8][ty][tx], sS[bo][8][ty + 1][tx], sS[bn][8][ty][tx], sS[bn][8][ty + 1][tx]));
                    sto(a.eta_next, cv, rheology_o(a, cv, t, em));
                }
            }
        }

        // rotate the column registers
        Pwm = Pw;
        zzwm = zzw;
        Pw = Pwc;
        xxw = mc0 * xxc;
        yyw = mc0 * yyc;
        zzw = zzwc;
        xxr = xxc;
        yyr = yyc;
        zzr = zzc;
        vxm = vx0;
        vym = vy0;
        vzm = vz0;
        qxm = qx0;
        qym = qy0;
        mxm = mx0;
        mym = my0;
        mzm = mz0;
        flm = fl0;
        vx0 = nvx;
        vy0 = nvy;
        vz0 = vzp;
        mx0 = nmx;
        my0 = nmy;
        mz0 = mzp;
        qx0 = mx0 * vx0;
        qy0 = my0 * vy0;
        qz0 = qzp;
        vzp = nvz2;
        mzp = nmz2;
        qzp = mzp * vzp;
        mc0 = nmc;
        mxy0 = nmxy;
        mxz0 = nmxz;
        myz0 = nmyz;
        fl0 = nfl;
        emm = em;
        em = e0;
        e0 = ne;
        oP = nP;
        oxx = nxx;
        oyy = nyy;
        ozz = nzz;
        oxy = nxy;
        oxz = nxz;
        oyz = nyz;
        ofP = nfP;
        gvx = ngx;
        gvy = ngy;
        gvz = ngz;
    }
}

// The viscosity update of the interior cells a narrowed immersed-boundary
// sweep did not march: each tile of the launch (lanes tx, rows ty, z chunks
// fg.kch) recorded its march [ks, ke] in a.tmarch; a cell of its chunk
// outside it gets eta_next = rheology(tau_II = 0, eta), as the sweep would
// have given it (every stress its tau_II averages is at an invalid location).
// One thread per column and z chunk (blockIdx.z) walks the chunk's planes
// outside the march of its tile.
template <typename T>
__device__ __forceinline__ T eta_rest_cell(const StokesArgs<T>& a, const T* st0, uint32_t bo, T e)
{
    if (a.rh) return rheo_relax(a, rheo_scaled(a, rheo_target(a, bo, T(0))), e);
    // a constant-coefficient rheology has one target at tau_II = 0, and cells
    // that have only ever been relaxed towards it from the same start hold
    // the same viscosity: k_eta_target relaxed one such value (st0[1] ->
    // st0[2]) with the same operations, so a cell holding exactly its bits
    // takes the result and skips the log and exp
    return same_bits(e, st0[1]) ? st0[2] : rheo_relax(a, st0[0], e);
}

template <typename T>
__global__ __launch_bounds__(256) void k_eta_rest(StokesArgs<T> a, FusedGeom fg, int tx, int ty, const T* st0)
{
    const int i = fg.lo[0] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = fg.lo[1] + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int bz = (int)blockIdx.z;
    if (i > fg.hi[0] || j > fg.hi[1]) return;
    const int bx = (fg.nx0 && i < fg.xa) ? 0 : fg.nx0 + (i - fg.xa) / (tx - 2);
    const int by = (j - fg.lo[1]) / (ty - 2);
    const int nby = (fg.hi[1] - fg.lo[1] + 1 + ty - 3) / (ty - 2);
    const int2 r = a.tmarch[bx + fg.nbx * (by + nby * bz)];
    const int k0 = fg.lo[2] + bz * fg.kch, k1 = min(k0 + fg.kch - 1, fg.hi[2]);
    const uint32_t col = (uint32_t)((i + a.sy * j) * (long)sizeof(T)), bsz = (uint32_t)(a.sz * (long)sizeof(T));
    auto run = [&](int ka, int kb) {
        for (int k = ka; k <= kb; k += 4) {
            T e[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) e[u] = k + u <= kb ? ldo(a.eta, col + bsz * (uint32_t)(k + u)) : T(1);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (k + u <= kb) {
                    const uint32_t bo = col + bsz * (uint32_t)(k + u);
                    sto(a.eta_next, bo, eta_rest_cell(a, st0, bo, e[u]));
                }
        }
    };
    if (r.x > r.y) {
        run(k0, k1);
    } else {
        run(k0, min(k1, r.x - 1));
        run(max(k0, r.y + 1), k1);
    }
}

// st0[0]: the scaled target viscosity at tau_II = 0 of a constant-coefficient
// rheology; st0[1], st0[2]: the viscosity of the cell at byte offset probe
// and its relaxation towards that target
template <typename T>
__global__ void k_eta_target(StokesArgs<T> a, T* st0, uint32_t probe)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const T st = rheo_scaled(a, rheo_target(a, 0u, T(0)));
        const T e = ldo(a.eta, probe);
        st0[0] = st;
        st0[1] = e;
        st0[2] = rheo_relax(a, st, e);
    }
}

// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8), each with its own L2.  Tiles are numbered
// x fastest.  order 0: tile b (the XCDs interleave tile by tile); order 1:
// every XCD takes a contiguous run of 1/8 of the tiles, so the tiles running
// side by side on one XCD include x and y neighbours whose shared halo lines
// meet in that XCD's L2; order 2: every 64 workgroups, each XCD takes 8
// consecutive tiles (x neighbours share their halo lines in one L2 while the
// XCDs still interleave finely, which keeps uneven work balanced).  The grid
// is padded to a multiple of 64; padding workgroups exit at once.
constexpr int NXCD = 8;

__device__ __forceinline__ int xcd_run_tile(int b, int order, int ntiles)
{
    if (order == 1) {
        const int per = (ntiles + NXCD - 1) / NXCD;
        return (b % NXCD) * per + b / NXCD;
    }
    if (order == 2) {
        // groups of 64 workgroups: XCD x takes 8 consecutive tiles of the group
        const int g = b / (NXCD * 8), r = b % (NXCD * 8);
        return g * (NXCD * 8) + (r % NXCD) * 8 + r / NXCD;
    }
    return b;
}

__device__ __forceinline__ bool fused_tile_of(int b, int order, int nbx, int nby, int nbz, int& bx, int& by, int& bz)
{
    const int ntiles = nbx * nby * nbz;
    const int t = xcd_run_tile(b, order, ntiles);
    if (t >= ntiles) return false;
    bx = t % nbx;
    const int r = t / nbx;
    by = r % nby;
    bz = r / nby;
    return true;
}

// MINW: minimum waves per SIMD the register allocation must allow (1: none).
// Workgroups from nfused on update the stress of the boundary shell (the
// pointwise sigma_point3, one point per thread, over the boxes of shell
// whose starts count workgroups of TX * TY threads): the shell's stress
// reads only the previous iteration, like the fused tiles, and the two write
// disjoint points, so they share the launch and the shell fills its tail.
template <typename T, bool IB, int TX, int TY, int MINW, bool VISC>
__global__ __launch_bounds__(TX * TY, MINW) void k_fused3(StokesArgs<T> a, FusedGeom fg, BoxList shell, int nfused)
{
    if ((int)blockIdx.x >= nfused) {
        const int bid = (int)blockIdx.x - nfused;
        int q = 0;
        while (q + 1 < shell.nb && bid >= shell.start[q + 1]) ++q;
        const long t = (long)(bid - shell.start[q]) * (TX * TY) + threadIdx.x + TX * threadIdx.y;
        int i, j, k;
        if (!box_point(shell.b[q], t, i, j, k)) return;
        sigma_point3<T, IB>(a, (long)i + a.sy * j + a.sz * k, a.sy, a.sz);
        return;
    }
    const int nby = (fg.hi[1] - fg.lo[1] + 1 + TY - 3) / (TY - 2);
    const int nbz = (fg.hi[2] - fg.lo[2] + fg.kch) / fg.kch;
    int bx, by, bz;
    if (!fused_tile_of(blockIdx.x, fg.order, fg.nbx, nby, nbz, bx, by, bz)) return;
    if constexpr (IB) fused3_ib_tile<T, TX, TY, VISC, false>(a, fg, bx, by, bz);
    else fused3_tile<T, false, TX, TY, VISC>(a, fg, bx, by, bz);
}

// Packed fused sweep over many instances: a persistent grid walks the
// (instance, tile) work list; each item is one workgroup tile of one
// instance, swept exactly as the single-instance kernel sweeps it.
template <typename T, bool IB, int TX, int TY, int MINW, bool VISC>
__global__ __launch_bounds__(TX * TY, MINW) void k_fused3_packed(const PackedInst<T>* __restrict__ inst,
                                                                 const int4* __restrict__ items, int nitems)
{
    // items are in x-fastest tile order per instance, mapped to workgroups as
    // in fused_tile_of; gridDim.x is a multiple of 8 * NXCD
    const int order = inst[0].fg.order;
    const int per = (nitems + NXCD - 1) / NXCD;
    const int x = blockIdx.x % NXCD;
    const int first = order == 1 ? blockIdx.x / NXCD : blockIdx.x;
    const int stride = order == 1 ? gridDim.x / NXCD : gridDim.x;
    const int nslot = order == 1 ? per : order == 2 ? (nitems + 8 * NXCD - 1) / (8 * NXCD) * (8 * NXCD) : nitems;
    for (int slot = first; slot < nslot; slot += stride) {
        const int it = order == 1 ? x * per + slot : xcd_run_tile(slot, order, nitems);
        if (order == 2 && it >= nitems) continue;
        if (it >= nitems) break;
        const int4 w = items[it];
        const StokesArgs<T> a = inst[w.x].a;
        const FusedGeom fg = inst[w.x].fg;
        __syncthreads(); // the previous item's LDS reads are done
        if constexpr (IB) fused3_ib_tile<T, TX, TY, VISC, true>(a, fg, w.y, w.z, w.w);
        else fused3_tile<T, false, TX, TY, VISC>(a, fg, w.y, w.z, w.w);
    }
}

// ---------------------------------------------------------------------------
// residuals and compute_tau (3D pointwise, full work range)
// ---------------------------------------------------------------------------
#define IDX3                                                                      \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;                      \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;                      \
    const int k = blockIdx.z + 1;                                                 \
    if (i > a.nx + 1 || j > a.ny + 1) return;                                     \
    const long sy = a.sy, sz = a.sz;                                              \
    const long c = (long)i + sy * j + sz * k;

// compute_residuals!: FullStokes kernels_3d.jl:66-75, IB kernels_3d.jl:118-143
// positions of a column outside its plane range hold only invalid locations
__device__ __forceinline__ bool outside_range(const int2* crange, int i, int j, int k, long sy)
{
    const int2 rg = crange[(long)i + sy * j];
    return k < rg.x || k > rg.y;
}

template <typename T, bool IB>
__device__ __forceinline__ void residual3_at(const StokesArgs<T>& a, int i, int j, int k)
{
    const long sy = a.sy, sz = a.sz;
    const long c = (long)i + sy * j + sz * k;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *P = a.P, *txx = a.txx, *tyy = a.tyy, *tzz = a.tzz, *txy = a.txy, *txz = a.txz, *tyz = a.tyz;
    const T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz;
    T gx = a.grav_field ? a.gx[c] : a.grav[0];
    T gy = a.grav_field ? a.gy[c] : a.grav[1];
    T gz = a.grav_field ? a.gz[c] : a.grav[2];
    if (!IB) {
        a.rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy + (Vz[c + sz] - Vz[c]) * iz;
        if (a.fP) a.rP[c] = a.rP[c] + a.fP[c];
        a.rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy +
                   (txz[c + sz] - txz[c]) * iz - gx;
        a.rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix +
                   (tyz[c + sz] - tyz[c]) * iz - gy;
        a.rVz[c] = -((P[c] - P[c - sz]) * iz) + (tzz[c] - tzz[c - sz]) * iz + (txz[c + 1] - txz[c]) * ix +
                   (tyz[c + sy] - tyz[c]) * iy - gz;
    } else if (a.flags[c] & F_FULL) {
        // every weight of the stencils below is exactly 1: the masked
        // expressions without their (exact) multiplications by 1
        a.rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy + (Vz[c + sz] - Vz[c]) * iz +
                  (a.fP ? a.fP[c] : T(0));
        a.rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy +
                   (txz[c + sz] - txz[c]) * iz - gx;
        a.rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix +
                   (tyz[c + sz] - tyz[c]) * iz - gy;
        a.rVz[c] = -((P[c] - P[c - sz]) * iz) + (tzz[c] - tzz[c - sz]) * iz + (txz[c + 1] - txz[c]) * ix +
                   (tyz[c + sy] - tyz[c]) * iy - gz;
    } else {
        const auto* na = a.na;
        const auto* ns = a.ns;
        const auto wx = ns[OM_VCC], wy = ns[OM_CVC], wz = ns[OM_CCV];
        const auto wc = na[OM_CCC], wxy = na[OM_VVC], wxz = na[OM_VCV], wyz = na[OM_CVV];
        const unsigned fl = a.flags[c];
        if (fl & F_CCC)
            a.rP[c] = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix + (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy +
                      (wz[c + sz] * Vz[c + sz] - wz[c] * Vz[c]) * iz + (a.fP ? a.fP[c] : T(0));
        else a.rP[c] = T(0);
        if (fl & F_VCC)
            a.rVx[c] = -((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                       (wxy[c + sy] * txy[c + sy] - wxy[c] * txy[c]) * iy + (wxz[c + sz] * txz[c + sz] - wxz[c] * txz[c]) * iz - gx;
        else a.rVx[c] = T(0);
        if (fl & F_CVC)
            a.rVy[c] = -((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                       (wxy[c + 1] * txy[c + 1] - wxy[c] * txy[c]) * ix + (wyz[c + sz] * tyz[c + sz] - wyz[c] * tyz[c]) * iz - gy;
        else a.rVy[c] = T(0);
        if (fl & F_CCV)
            a.rVz[c] = -((wc[c] * P[c] - wc[c - sz] * P[c - sz]) * iz) + (wc[c] * tzz[c] - wc[c - sz] * tzz[c - sz]) * iz +
                       (wxz[c + 1] * txz[c + 1] - wxz[c] * txz[c]) * ix + (wyz[c + sy] * tyz[c + sy] - wyz[c] * tyz[c]) * iy - gz;
        else a.rVz[c] = T(0);
    }
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_residuals3(StokesArgs<T> a)
{
    IDX3
    (void)c;
    if (IB && a.skip_empty && outside_range(a.crange, i, j, k, sy)) return;
    residual3_at<T, IB>(a, i, j, k);
}

// compute_tau! store; LAG: also the FP32 difference old - new of the stress
// it replaces (written in place, each location by one thread)
template <bool LAG, typename T>
__device__ __forceinline__ void put_tau(T* dst, const T* old, float* lag, long c, T v)
{
    if (LAG) lag[c] = (float)(old[c] - v);
    dst[c] = v;
}
// the same at an invalid immersed-boundary location: the stress is 0, and so
// is the lag (the sweeps never read the old stress of an invalid location,
// and their first sweep overwrites it with 0)
template <bool LAG, typename T>
__device__ __forceinline__ void put_tau0(T* dst, float* lag, long c)
{
    if (LAG) lag[c] = 0.0f;
    dst[c] = T(0);
}

// compute_tau!: kernels_3d.jl:44-64 (IB variant: IB kernels_3d.jl:78-116 with
// its undefined `Pr` location read as the Center location); writes the
// current stress (txxn == txx)
template <typename T, bool IB, bool LAG>
__device__ __forceinline__ void tau3_at(const StokesArgs<T>& a, const TauLag& lg, int i, int j, int k)
{
    const long sy = a.sy, sz = a.sz;
    const long c = (long)i + sy * j + sz * k;
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz, *eta = a.eta;
    const auto* ns = a.ns;
    const unsigned fl = IB ? a.flags[c] : 0xffu;
    // masked derivatives only where a weight may differ from 1 (F_FULL: the
    // plain expressions are the masked ones bit for bit)
    const bool wtd = IB && !(fl & F_FULL);
    T ec = eta[c];
    if (fl & F_CCC) {
        T exx, eyy, ezz;
        if (wtd) {
            exx = (ns[OM_VCC][c + 1] * Vx[c + 1] - ns[OM_VCC][c] * Vx[c]) * ix;
            eyy = (ns[OM_CVC][c + sy] * Vy[c + sy] - ns[OM_CVC][c] * Vy[c]) * iy;
            ezz = (ns[OM_CCV][c + sz] * Vz[c + sz] - ns[OM_CCV][c] * Vz[c]) * iz;
        } else {
            exx = (Vx[c + 1] - Vx[c]) * ix;
            eyy = (Vy[c + sy] - Vy[c]) * iy;
            ezz = (Vz[c + sz] - Vz[c]) * iz;
        }
        T divV = exx + eyy + ezz;
        put_tau<LAG>(a.txxn, a.txx, lg.t[0], c, T(2) * ec * (exx - divV / T(3)));
        put_tau<LAG>(a.tyyn, a.tyy, lg.t[1], c, T(2) * ec * (eyy - divV / T(3)));
        put_tau<LAG>(a.tzzn, a.tzz, lg.t[2], c, T(2) * ec * (ezz - divV / T(3)));
    } else {
        put_tau0<LAG>(a.txxn, lg.t[0], c);
        put_tau0<LAG>(a.tyyn, lg.t[1], c);
        put_tau0<LAG>(a.tzzn, lg.t[2], c);
    }
    if (fl & F_VVC) {
        T exy;
        if (wtd) exy = T(0.5) * ((ns[OM_CVC][c] * Vy[c] - ns[OM_CVC][c - 1] * Vy[c - 1]) * ix +
                                (ns[OM_VCC][c] * Vx[c] - ns[OM_VCC][c - sy] * Vx[c - sy]) * iy);
        else exy = T(0.5) * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
        put_tau<LAG>(a.txyn, a.txy, lg.t[3], c, T(2) * lerp4(eta[c - 1 - sy], eta[c - sy], eta[c - 1], eta[c]) * exy);
    } else put_tau0<LAG>(a.txyn, lg.t[3], c);
    if (fl & F_VCV) {
        T exz;
        if (wtd) exz = T(0.5) * ((ns[OM_CCV][c] * Vz[c] - ns[OM_CCV][c - 1] * Vz[c - 1]) * ix +
                                (ns[OM_VCC][c] * Vx[c] - ns[OM_VCC][c - sz] * Vx[c - sz]) * iz);
        else exz = T(0.5) * ((Vz[c] - Vz[c - 1]) * ix + (Vx[c] - Vx[c - sz]) * iz);
        put_tau<LAG>(a.txzn, a.txz, lg.t[4], c, T(2) * lerp4(eta[c - 1 - sz], eta[c - sz], eta[c - 1], eta[c]) * exz);
    } else put_tau0<LAG>(a.txzn, lg.t[4], c);
    if (fl & F_CVV) {
        T eyz;
        if (wtd) eyz = T(0.5) * ((ns[OM_CCV][c] * Vz[c] - ns[OM_CCV][c - sy] * Vz[c - sy]) * iy +
                                (ns[OM_CVC][c] * Vy[c] - ns[OM_CVC][c - sz] * Vy[c - sz]) * iz);
        else eyz = T(0.5) * ((Vz[c] - Vz[c - sy]) * iy + (Vy[c] - Vy[c - sz]) * iz);
        put_tau<LAG>(a.tyzn, a.tyz, lg.t[5], c, T(2) * lerp4(eta[c - sy - sz], eta[c - sz], eta[c - sy], eta[c]) * eyz);
    } else put_tau0<LAG>(a.tyzn, lg.t[5], c);
}

template <typename T, bool IB, bool LAG>
__global__ __launch_bounds__(256) void k_tau3(StokesArgs<T> a, TauLag lg)
{
    IDX3
    (void)c;
    if (IB && a.skip_empty && outside_range(a.crange, i, j, k, sy)) return;
    tau3_at<T, IB, LAG>(a, lg, i, j, k);
}

static inline dim3 block2();
static inline dim3 grid3(const Geo& g);

// eta_next <- rheology(tau_II(tau), I) at the current stress (the viscosity half of
// update_V!, kernels_3d.jl:35-39 / kernels_2d.jl:27-28)
template <typename T>
__device__ __forceinline__ void eta_update_at(const StokesArgs<T>& a, long c, int nd)
{
    const long sy = a.sy, sz = a.sz;
    const T *txx = a.txx, *tyy = a.tyy, *txy = a.txy;
    T t;
    if (nd == 3) {
        const T *tzz = a.tzz, *txz = a.txz, *tyz = a.tyz;
        t = tauII(txx[c], tyy[c], tzz[c], lerp4(txy[c], txy[c + 1], txy[c + sy], txy[c + 1 + sy]),
                  lerp4(txz[c], txz[c + 1], txz[c + sz], txz[c + 1 + sz]), lerp4(tyz[c], tyz[c + sy], tyz[c + sz], tyz[c + sy + sz]));
    } else {
        T p = txx[c], q = tyy[c];
        T l = lerp4(txy[c], txy[c + 1], txy[c + sy], txy[c + 1 + sy]);
        t = sqrt(T(0.5) * (p * p + q * q) + l * l);
    }
    a.eta_next[c] = rheology(a, c, t);
}

template <typename T>
__global__ __launch_bounds__(256) void k_eta_update(StokesArgs<T> a, int nd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (nd == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > a.nx + 1 || j > a.ny + 1) return;
    const long sy = a.sy, sz = a.sz;
    const long c = (long)i + sy * j + sz * k;
    eta_update_at(a, c, nd);
}

// The same over [1, n+1]^3 of a 3D immersed-boundary model whose stress is
// 0 at every invalid location, one thread per column and chunk of ETA_KCH
// planes: on the planes of the column's erange as above; elsewhere every
// stress tau_II averages sits at an invalid location, so tau_II = +0 and
// only the old viscosity is read.  With a constant-coefficient rheology
// eta_next then depends on the old viscosity alone: a thread keeps its last
// input and result, and a cell holding the same bits takes the result
// without a log and an exp.
constexpr int ETA_KCH = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_eta_update_cols(StokesArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > a.nx + 1 || j > a.ny + 1) return;
    const long sy = a.sy, sz = a.sz;
    const int2 re = a.erange[(long)i + sy * j];
    const int k0 = 1 + (int)blockIdx.z * ETA_KCH, k1 = min(k0 + ETA_KCH - 1, a.nz + 1);
    const bool memo = a.rh == nullptr;
    T m_in = T(0), m_out = T(0);
    bool m_set = false;
    for (int k = k0; k <= k1; ++k) {
        const long c = (long)i + sy * j + sz * k;
        if (k >= re.x && k <= re.y) {
            eta_update_at(a, c, 3);
        } else {
            const T e = a.eta[c];
            if (!(memo && m_set && same_bits(e, m_in))) {
                m_in = e;
                m_out = rheology(a, c, T(0));
                m_set = true;
            }
            a.eta_next[c] = m_out;
        }
    }
}

template <typename T>
void launch_eta_update(const Geo& g, const StokesArgs<T>& a, hipStream_t s)
{
    if (g.nd == 3 && a.skip_empty && a.erange) {
        const dim3 grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, (g.n[2] + 1 + ETA_KCH - 1) / ETA_KCH);
        hipLaunchKernelGGL((k_eta_update_cols<T>), grd, block2(), 0, s, a);
        return;
    }
    dim3 grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.nd == 3 ? g.n[2] + 1 : 1);
    hipLaunchKernelGGL((k_eta_update<T>), grd, block2(), 0, s, a, g.nd);
}

// IB validity bits, once per geometry (null_spaces.jl:1-43)
__device__ __forceinline__ bool lt(const double* w, long c) { return w[c] < 1e-6; }

__device__ bool null3(const double* const* w, int loc, long c, long sy, long sz)
{
    switch (loc) {
    case L_CCC: return lt(w[OM_VCC], c) || lt(w[OM_VCC], c + 1) || lt(w[OM_CVC], c) || lt(w[OM_CVC], c + sy) ||
                       lt(w[OM_CCV], c) || lt(w[OM_CCV], c + sz);
    case L_VCC: return lt(w[OM_CCC], c - 1) || lt(w[OM_CCC], c) || lt(w[OM_VVC], c) || lt(w[OM_VVC], c + sy) ||
                       lt(w[OM_VCV], c) || lt(w[OM_VCV], c + sz);
    case L_CVC: return lt(w[OM_VVC], c) || lt(w[OM_VVC], c + 1) || lt(w[OM_CCC], c - sy) || lt(w[OM_CCC], c) ||
                       lt(w[OM_CVV], c) || lt(w[OM_CVV], c + sz);
    case L_CCV: return lt(w[OM_VCV], c) || lt(w[OM_VCV], c + 1) || lt(w[OM_CVV], c) || lt(w[OM_CVV], c + sy) ||
                       lt(w[OM_CCC], c - sz) || lt(w[OM_CCC], c);
    case L_VVC: return lt(w[OM_CVC], c - 1) || lt(w[OM_CVC], c) || lt(w[OM_VCC], c - sy) || lt(w[OM_VCC], c);
    case L_VCV: return lt(w[OM_CCV], c - 1) || lt(w[OM_CCV], c) || lt(w[OM_VCC], c - sz) || lt(w[OM_VCC], c);
    case L_CVV: return lt(w[OM_CCV], c - sy) || lt(w[OM_CCV], c) || lt(w[OM_CVC], c - sz) || lt(w[OM_CVC], c);
    }
    return true;
}

__device__ bool null2(const double* const* w, int loc, long c, long sy)
{
    switch (loc) {
    case L_CCC: return lt(w[OM2_VC], c) || lt(w[OM2_VC], c + 1) || lt(w[OM2_CV], c) || lt(w[OM2_CV], c + sy);
    case L_VCC: return lt(w[OM2_CC], c - 1) || lt(w[OM2_CC], c) || lt(w[OM2_VV], c) || lt(w[OM2_VV], c + sy);
    case L_CVC: return lt(w[OM2_VV], c) || lt(w[OM2_VV], c + 1) || lt(w[OM2_CC], c - sy) || lt(w[OM2_CC], c);
    case L_VVC: return lt(w[OM2_CV], c - 1) || lt(w[OM2_CV], c) || lt(w[OM2_VC], c - sy) || lt(w[OM2_VC], c);
    }
    return true;
}

struct MaskPtrs {
    const double* na[8];
    const double* ns[8];
};

__global__ __launch_bounds__(256) void k_flags(MaskPtrs m, uint8_t* flags, int nd, int nx, int ny, long sy, long sz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (nd == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > nx + 1 || j > ny + 1) return;
    const long c = (long)i + sy * j + sz * k;
    unsigned f = 0;
    if (nd == 3) {
        // (location, component, bit, sigma side?)
        const int locs[7] = {L_CCC, L_VVC, L_VCV, L_CVV, L_VCC, L_CVC, L_CCV};
        const int comp[7] = {OM_CCC, OM_VVC, OM_VCV, OM_CVV, OM_VCC, OM_CVC, OM_CCV};
        for (int q = 0; q < 7; ++q) {
            const double* const* w = (q < 4) ? m.na : m.ns;  // isvalid(w) ...
            const double* const* w2 = (q < 4) ? m.ns : m.na; // ... && !isempty(w2)
            bool v = !(null3(w, locs[q], c, sy, sz) || lt(w[comp[q]], c)) && !lt(w2[comp[q]], c);
            if (v) f |= 1u << q;
        }
        bool full = true;
        const long nb[7] = {0, 1, -1, sy, -sy, sz, -sz};
        for (int q = 0; q < 8 && full; ++q)
            for (int o = 0; o < 7; ++o) full = full && m.na[q][c + nb[o]] == 1.0 && m.ns[q][c + nb[o]] == 1.0;
        if (full) f |= F_FULL;
    } else {
        const int locs[4] = {L_CCC, L_VVC, L_VCC, L_CVC};
        const int comp[4] = {OM2_CC, OM2_VV, OM2_VC, OM2_CV};
        const unsigned bit[4] = {F_CCC, F_VVC, F_VCC, F_CVC};
        for (int q = 0; q < 4; ++q) {
            const double* const* w = (q < 2) ? m.na : m.ns;
            const double* const* w2 = (q < 2) ? m.ns : m.na;
            bool v = !(null2(w, locs[q], c, sy) || lt(w[comp[q]], c)) && !lt(w2[comp[q]], c);
            if (v) f |= bit[q];
        }
    }
    flags[c] = (uint8_t)f;
}

void launch_ib_flags(const Geo& g, const double* const na[8], const double* const ns[8], uint8_t* flags, hipStream_t s)
{
    MaskPtrs m;
    for (int q = 0; q < 8; ++q) {
        m.na[q] = na[q];
        m.ns[q] = ns[q];
    }
    dim3 blk(64, 4, 1), grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.nd == 3 ? g.n[2] + 1 : 1);
    hipLaunchKernelGGL(k_flags, grd, blk, 0, s, m, flags, g.nd, g.n[0], g.n[1], g.sy, g.nd == 3 ? g.sz : 0);
}

// Plane ranges per column (i, j), i in [0, nx+2], j in [0, ny+2], once per
// geometry.  own: the first and last plane k whose flags are non-zero (some
// location of the cell is valid); every plane outside it has no valid
// location.  eta: the own ranges of the 3 x 3 columns around (i, j), each
// widened by one plane on both sides: a valid location at (i', j', k') reads
// the viscosity of the cells (i'-1..i', j'-1..j', k'-1..k') only, so outside
// this range no valid location reads the viscosity of the column.
__global__ __launch_bounds__(256) void k_col_range(const uint8_t* __restrict__ flags, int2* __restrict__ own, int nx,
                                                   int ny, int nz, long sy, long sz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i > nx + 2 || j > ny + 2) return;
    const long col = i + sy * j;
    int lo = RANGE_NONE, hi = -RANGE_NONE;
    for (int k = 0; k <= nz + 2; ++k)
        if (flags[col + sz * k]) {
            lo = min(lo, k);
            hi = k;
        }
    own[col] = make_int2(lo, hi);
}

__global__ __launch_bounds__(256) void k_col_widen(const int2* __restrict__ own, int2* __restrict__ eta, int nx,
                                                   int ny, long sy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i > nx + 2 || j > ny + 2) return;
    int lo = RANGE_NONE, hi = -RANGE_NONE;
    for (int jj = max(j - 1, 0); jj <= min(j + 1, ny + 2); ++jj)
        for (int ii = max(i - 1, 0); ii <= min(i + 1, nx + 2); ++ii) {
            const int2 r = own[ii + sy * jj];
            if (r.x <= r.y) {
                lo = min(lo, r.x - 1);
                hi = max(hi, r.y + 1);
            }
        }
    eta[i + sy * j] = make_int2(lo, hi);
}

void launch_col_ranges(const Geo& g, const uint8_t* flags, int2* own, int2* eta, hipStream_t s)
{
    dim3 grd((g.n[0] + 3 + 255) / 256, g.n[1] + 3, 1);
    hipLaunchKernelGGL(k_col_range, grd, dim3(256), 0, s, flags, own, g.n[0], g.n[1], g.n[2], g.sy, g.sz);
    hipLaunchKernelGGL(k_col_widen, grd, dim3(256), 0, s, own, eta, g.n[0], g.n[1], g.sy);
}

// ---------------------------------------------------------------------------
// 2D: FullStokes kernels_2d.jl; the IB 2D path restates the 3D IB kernels
// ---------------------------------------------------------------------------
#define IDX2                                                                      \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;                      \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;                      \
    if (i > a.nx + 1 || j > a.ny + 1) return;                                     \
    const long sy = a.sy;                                                         \
    const long c = (long)i + sy * j;

template <typename T, bool IB>
__device__ __forceinline__ void sigma_point2(const StokesArgs<T>& a, long c, long sy)
{
    const T ix = a.ix, iy = a.iy;
    const T *Vx = a.Vx, *Vy = a.Vy, *eta = a.eta;
    const T ec = eta[c];
    const T exy_eta = lerp4(eta[c - 1 - sy], eta[c - sy], eta[c - 1], eta[c]);
    if (!IB) {
        // kernels_2d.jl:7-17
        T exx = dif(Vx[c + 1], Vx[c], ix);
        T eyy = dif(Vy[c + sy], Vy[c], iy);
        T exy = T(0.5) * (dif(Vy[c], Vy[c - 1], ix) + dif(Vx[c], Vx[c - sy], iy));
        T divV = exx + eyy;
        a.Pn[c] = upd_P(a.P[c], a.fP ? divV + a.fP[c] : divV, ec, a.dPr);
        a.txxn[c] = upd_td(a.txx[c], ec, exx, divV, a.dtau[0]);
        a.tyyn[c] = upd_td(a.tyy[c], ec, eyy, divV, a.dtau[1]);
        a.txyn[c] = upd_to(a.txy[c], exy_eta, exy, a.dtau[2]);
    } else {
        const auto wx = a.ns[OM2_VC], wy = a.ns[OM2_CV];
        const unsigned fl = a.flags[c];
        if (fl & F_CCC) {
            T exx = dif(wx[c + 1] * Vx[c + 1], wx[c] * Vx[c], ix);
            T eyy = dif(wy[c + sy] * Vy[c + sy], wy[c] * Vy[c], iy);
            T divV = exx + eyy;
            a.Pn[c] = upd_P(a.P[c], a.fP ? divV + a.fP[c] : divV, ec, a.dPr);
            a.txxn[c] = upd_td(a.txx[c], ec, exx, divV, a.dtau[0]);
            a.tyyn[c] = upd_td(a.tyy[c], ec, eyy, divV, a.dtau[1]);
        } else {
            a.Pn[c] = T(0); a.txxn[c] = T(0); a.tyyn[c] = T(0);
        }
        if (fl & F_VVC) {
            T exy = T(0.5) * (dif(wy[c] * Vy[c], wy[c - 1] * Vy[c - 1], ix) + dif(wx[c] * Vx[c], wx[c - sy] * Vx[c - sy], iy));
            a.txyn[c] = upd_to(a.txy[c], exy_eta, exy, a.dtau[2]);
        } else a.txyn[c] = T(0);
    }
}

template <typename T, bool IB>
__device__ __forceinline__ void velocity_point2(const StokesArgs<T>& a, long c, long sy)
{
    const T ix = a.ix, iy = a.iy;
    const T *P = a.Pn, *txx = a.txxn, *tyy = a.tyyn, *txy = a.txyn, *eta = a.eta;
    T gx = a.grav_field ? a.gx[c] : a.grav[0];
    T gy = a.grav_field ? a.gy[c] : a.grav[1];
    // each component interpolates eta to its own location (kernels_2d.jl:24-25
    // reuses location(V.x) for both)
    T ex = T(0.5) * (eta[c - 1] + eta[c]);
    T ey = T(0.5) * (eta[c - sy] + eta[c]);
    const T zero = T(0);
    if (!IB) {
        a.Vxn[c] = upd_V(a.Vx[c], dif(P[c], P[c - 1], ix), dif(txx[c], txx[c - 1], ix), dif(txy[c + sy], txy[c], iy), zero, gx, ex, a.dV[0]);
        a.Vyn[c] = upd_V(a.Vy[c], dif(P[c], P[c - sy], iy), dif(tyy[c], tyy[c - sy], iy), dif(txy[c + 1], txy[c], ix), zero, gy, ey, a.dV[1]);
    } else {
        const auto wc = a.na[OM2_CC], wv = a.na[OM2_VV];
        const unsigned fl = a.flags[c];
        a.Vxn[c] = (fl & F_VCC) ? upd_V_ib(a.Vx[c], dif(wc[c] * P[c], wc[c - 1] * P[c - 1], ix), dif(wc[c] * txx[c], wc[c - 1] * txx[c - 1], ix),
                                           dif(wv[c + sy] * txy[c + sy], wv[c] * txy[c], iy), zero, gx, ex, a.ns[OM2_VC][c], a.dV[0])
                                : T(0);
        a.Vyn[c] = (fl & F_CVC) ? upd_V_ib(a.Vy[c], dif(wc[c] * P[c], wc[c - sy] * P[c - sy], iy), dif(wc[c] * tyy[c], wc[c - sy] * tyy[c - sy], iy),
                                           dif(wv[c + 1] * txy[c + 1], wv[c] * txy[c], ix), zero, gy, ey, a.ns[OM2_CV][c], a.dV[1])
                                : T(0);
    }
    if (a.frozen_eta) return;
    // kernels_2d.jl:27-28
    T p = txx[c], q = tyy[c];
    T l = lerp4(txy[c], txy[c + 1], txy[c + sy], txy[c + 1 + sy]);
    a.eta_next[c] = rheology(a, c, sqrt(T(0.5) * (p * p + q * q) + l * l));
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_sigma2(StokesArgs<T> a)
{
    IDX2
    sigma_point2<T, IB>(a, c, sy);
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_velocity2(StokesArgs<T> a)
{
    IDX2
    velocity_point2<T, IB>(a, c, sy);
}

// Packed pointwise sweeps over the boxes of many instances (3D shells, whole
// 2D grids): blk2box maps a workgroup to its box, each box covers whole
// workgroups of 256 points (i fastest).
template <typename T, bool IB, int ND, bool SIGMA>
__global__ __launch_bounds__(256) void k_point_packed(const PackedInst<T>* __restrict__ inst,
                                                      const PackedBox* __restrict__ boxes,
                                                      const int* __restrict__ blk2box)
{
    const PackedBox b = boxes[blk2box[blockIdx.x]];
    const long t = (long)(blockIdx.x - b.start) * 256 + threadIdx.x;
    int ii, jj, kk;
    if (!box_point(Box{b.i0, b.i1, b.j0, b.j1, b.k0, b.k1}, t, ii, jj, kk)) return;
    const long i = ii, j = jj, k = kk;
    const StokesArgs<T>& a = inst[b.inst].a;
    if constexpr (ND == 3) {
        const long sy = a.sy, sz = a.sz;
        if constexpr (SIGMA) sigma_point3<T, IB>(a, i + sy * j + sz * k, sy, sz);
        else velocity_point3<T, IB>(a, i + sy * j + sz * k, sy, sz);
    } else {
        const long sy = a.sy;
        if constexpr (SIGMA) sigma_point2<T, IB>(a, i + sy * j, sy);
        else velocity_point2<T, IB>(a, i + sy * j, sy);
    }
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_residuals2(StokesArgs<T> a)
{
    IDX2
    const T ix = a.ix, iy = a.iy;
    const T *P = a.P, *txx = a.txx, *tyy = a.tyy, *txy = a.txy, *Vx = a.Vx, *Vy = a.Vy;
    T gx = a.grav_field ? a.gx[c] : a.grav[0];
    T gy = a.grav_field ? a.gy[c] : a.grav[1];
    if (!IB) {
        // kernels_2d.jl:50-53
        a.rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy;
        if (a.fP) a.rP[c] = a.rP[c] + a.fP[c];
        a.rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy - gx;
        a.rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix - gy;
    } else {
        const auto wx = a.ns[OM2_VC], wy = a.ns[OM2_CV], wc = a.na[OM2_CC], wv = a.na[OM2_VV];
        const unsigned fl = a.flags[c];
        if (fl & F_CCC)
            a.rP[c] = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix + (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy +
                      (a.fP ? a.fP[c] : T(0));
        else a.rP[c] = T(0);
        if (fl & F_VCC)
            a.rVx[c] = -((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                       (wv[c + sy] * txy[c + sy] - wv[c] * txy[c]) * iy - gx;
        else a.rVx[c] = T(0);
        if (fl & F_CVC)
            a.rVy[c] = -((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                       (wv[c + 1] * txy[c + 1] - wv[c] * txy[c]) * ix - gy;
        else a.rVy[c] = T(0);
    }
}

template <typename T, bool IB, bool LAG>
__global__ __launch_bounds__(256) void k_tau2(StokesArgs<T> a, TauLag lg)
{
    IDX2
    const T ix = a.ix, iy = a.iy;
    const T *Vx = a.Vx, *Vy = a.Vy, *eta = a.eta;
    const auto* ns = a.ns;
    const unsigned fl = IB ? a.flags[c] : 0xffu;
    T ec = eta[c];
    if (fl & F_CCC) {
        T exx, eyy;
        if (IB) {
            exx = (ns[OM2_VC][c + 1] * Vx[c + 1] - ns[OM2_VC][c] * Vx[c]) * ix;
            eyy = (ns[OM2_CV][c + sy] * Vy[c + sy] - ns[OM2_CV][c] * Vy[c]) * iy;
        } else {
            exx = (Vx[c + 1] - Vx[c]) * ix;
            eyy = (Vy[c + sy] - Vy[c]) * iy;
        }
        T divV = exx + eyy;
        put_tau<LAG>(a.txxn, a.txx, lg.t[0], c, T(2) * ec * (exx - divV / T(3)));
        put_tau<LAG>(a.tyyn, a.tyy, lg.t[1], c, T(2) * ec * (eyy - divV / T(3)));
    } else {
        put_tau0<LAG>(a.txxn, lg.t[0], c);
        put_tau0<LAG>(a.tyyn, lg.t[1], c);
    }
    if (fl & F_VVC) {
        T exy;
        if (IB) exy = T(0.5) * ((ns[OM2_CV][c] * Vy[c] - ns[OM2_CV][c - 1] * Vy[c - 1]) * ix +
                                (ns[OM2_VC][c] * Vx[c] - ns[OM2_VC][c - sy] * Vx[c - sy]) * iy);
        else exy = T(0.5) * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
        put_tau<LAG>(a.txyn, a.txy, lg.t[3], c, T(2) * lerp4(eta[c - 1 - sy], eta[c - sy], eta[c - 1], eta[c]) * exy);
    } else put_tau0<LAG>(a.txyn, lg.t[3], c);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 block2() { return dim3(64, 4, 1); }
static inline dim3 grid2(const Geo& g) { return dim3((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, 1); }
static inline dim3 grid3(const Geo& g) { return dim3((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[2] + 1); }

static void finish_boxes(BoxList& bl)
{
    bl.start[0] = 0;
    for (int q = 0; q < bl.nb; ++q) {
        const Box& b = bl.b[q];
        const long n = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
        bl.start[q + 1] = bl.start[q] + (int)((n + 255) / 256);
    }
}

BoxList full_box(const Geo& g)
{
    BoxList bl;
    bl.nb = 1;
    bl.b[0] = Box{1, g.n[0] + 1, 1, g.n[1] + 1, 1, g.n[2] + 1};
    finish_boxes(bl);
    return bl;
}

bool fused_split(const Geo& g, int ty, int kch, FusedGeom& fg, BoxList& shell)
{
    // interior [3, n-1] in every dimension; the last workgroup tile in x and y
    // may be partial (its footprint is clamped to hi+1)
    (void)ty;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = 3;
        hi[d] = g.n[d] - 1;
        if (hi[d] < lo[d]) {
            shell = full_box(g);
            return false;
        }
    }
    for (int d = 0; d < 3; ++d) {
        fg.lo[d] = lo[d];
        fg.hi[d] = hi[d];
    }
    fg.kch = kch;
    fg.order = 0;
    fused_xtiles(fg, 64, 8, g.off);
    const int N0 = g.n[0] + 1, N1 = g.n[1] + 1, N2 = g.n[2] + 1;
    BoxList& bl = shell;
    bl.nb = 0;
    auto add = [&](int i0, int i1, int j0, int j1, int k0, int k1) {
        if (i0 <= i1 && j0 <= j1 && k0 <= k1) bl.b[bl.nb++] = Box{i0, i1, j0, j1, k0, k1};
    };
    add(1, N0, 1, N1, 1, lo[2] - 1);                       // z low slab
    add(1, N0, 1, N1, hi[2] + 1, N2);                      // z high slab
    add(1, N0, 1, lo[1] - 1, lo[2], hi[2]);                // y low
    add(1, N0, hi[1] + 1, N1, lo[2], hi[2]);               // y high
    add(1, lo[0] - 1, lo[1], hi[1], lo[2], hi[2]);         // x low
    add(hi[0] + 1, N0, lo[1], hi[1], lo[2], hi[2]);        // x high
    finish_boxes(bl);
    return true;
}

unsigned face_layers_empty(const Geo& g, const int2* r)
{
    if (g.nd != 3) return 0u;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    unsigned full = 0u; // faces whose layer holds a valid location
    for (int j = 0; j <= n1 + 2; ++j)
        for (int i = 0; i <= n0 + 2; ++i) {
            const int2 c = r[(size_t)i + (size_t)g.sy * j];
            if (c.x > c.y) continue; // no valid location in the column
            if (i <= 2) full |= 1u << 0;
            if (i >= n0) full |= 1u << 1;
            if (j <= 2) full |= 1u << 2;
            if (j >= n1) full |= 1u << 3;
            if (c.x <= 2) full |= 1u << 4;
            if (c.y >= n2) full |= 1u << 5;
        }
    return ~full & 0x3fu;
}

BoxList drop_empty_boxes(const Geo& g, const BoxList& bl, unsigned empty)
{
    if (!empty || g.nd != 3) return bl;
    BoxList out;
    out.nb = 0;
    for (int q = 0; q < bl.nb; ++q) {
        const Box& b = bl.b[q];
        const int lo[3] = {b.i0, b.j0, b.k0}, hi[3] = {b.i1, b.j1, b.k1};
        bool drop = false;
        for (int d = 0; d < 3; ++d) {
            if ((empty & (1u << (2 * d))) && hi[d] <= 2) drop = true;
            if ((empty & (1u << (2 * d + 1))) && lo[d] >= g.n[d]) drop = true;
        }
        if (!drop) out.b[out.nb++] = b;
    }
    finish_boxes(out);
    return out;
}

template <typename T>
void launch_sigma_boxes(const Geo& g, const StokesArgs<T>& a, const BoxList& bl, bool ib, hipStream_t s)
{
    if (g.nd == 3) {
        if (bl.nb == 0 || bl.start[bl.nb] == 0) return;
        dim3 grd(bl.start[bl.nb]);
        if (ib) hipLaunchKernelGGL((k_sigma_box<T, true>), grd, dim3(256), 0, s, a, bl);
        else hipLaunchKernelGGL((k_sigma_box<T, false>), grd, dim3(256), 0, s, a, bl);
    } else {
        if (ib) hipLaunchKernelGGL((k_sigma2<T, true>), grid2(g), block2(), 0, s, a);
        else hipLaunchKernelGGL((k_sigma2<T, false>), grid2(g), block2(), 0, s, a);
    }
}

template <typename T>
void launch_velocity_boxes(const Geo& g, const StokesArgs<T>& a, const BoxList& bl, bool ib, hipStream_t s)
{
    if (g.nd == 3) {
        if (bl.nb == 0 || bl.start[bl.nb] == 0) return;
        dim3 grd(bl.start[bl.nb]);
        if (ib) hipLaunchKernelGGL((k_velocity_box<T, true>), grd, dim3(256), 0, s, a, bl);
        else hipLaunchKernelGGL((k_velocity_box<T, false>), grd, dim3(256), 0, s, a, bl);
    } else {
        if (ib) hipLaunchKernelGGL((k_velocity2<T, true>), grid2(g), block2(), 0, s, a);
        else hipLaunchKernelGGL((k_velocity2<T, false>), grid2(g), block2(), 0, s, a);
    }
}

// (tile lanes in x, tile rows, minimum waves per SIMD) variants of the fused
// sweep; the first entry is the default.  64-lane tiles sweep 62 columns at
// any alignment; 66-lane tiles sweep 64 line-aligned columns.
#define FI_FUSED_CFG_D(X) X(64, 8, 1) X(64, 8, 4) X(64, 4, 1) X(64, 4, 3) X(66, 7, 1) X(66, 8, 3)
#define FI_FUSED_CFG_F(X) X(64, 8, 1) X(64, 16, 1) X(66, 15, 1) X(66, 15, 4) X(66, 7, 1)

int fused_configs(bool fp32, int cfg[][3], int max)
{
    int n = 0;
#define FI_LIST_(TX_, TY_, W_) if (n < max) { cfg[n][0] = TX_; cfg[n][1] = TY_; cfg[n][2] = W_; ++n; }
    if (fp32) { FI_FUSED_CFG_F(FI_LIST_) }
    else { FI_FUSED_CFG_D(FI_LIST_) }
#undef FI_LIST_
    return n;
}

// x tiling of fg for tiles of tx lanes over fields of esize-byte elements
// whose Chmy index 0 sits `off` elements past a 128-byte line: sets xa, nx0,
// nbx (see fused_xrange)
void fused_xtiles(FusedGeom& fg, int tx, int esize, long off)
{
    const int lo = fg.lo[0], hi = fg.hi[0];
    if (tx == 64) {
        fg.xa = lo;
        fg.nx0 = 0;
    } else {
        const long line = 128 / esize;
        long c = lo + 1; // a leading tile needs at least the columns [lo, lo]
        while ((off + c) % line != 0) ++c;
        fg.xa = (int)(c <= hi ? c : hi + 1);
        fg.nx0 = 1;
    }
    fg.nbx = fg.nx0 + (hi - fg.xa + 1 + tx - 3) / (tx - 2);
}

// z planes per workgroup when not set: long chunks (fewer halo planes) as
// long as the launch keeps about 1024 workgroups for the 256 CUs
int fused_auto_kch(const FusedGeom& fg, int tiles_per_plane)
{
    const long nz = fg.hi[2] - fg.lo[2] + 1;
    long k = 64;
    while (k > 8 && nz * tiles_per_plane / k < 1024) k /= 2;
    return (int)k;
}

// the fused launch; fg receives the geometry it used (x tiles, z chunk)
template <typename T>
static void launch_fused_sweep(const Geo& g, const StokesArgs<T>& a, const FusedGeom& fg0, bool ib, int tx, int ty,
                               int minw, hipStream_t s, const BoxList* shell_sigma, FusedGeom& fg)
{
    fg = fg0;
    fused_xtiles(fg, tx, (int)sizeof(T), g.off);
    const int nby = (fg.hi[1] - fg.lo[1] + 1 + ty - 3) / (ty - 2);
    if (fg.kch <= 0) {
        fg.kch = fused_auto_kch(fg, fg.nbx * nby);
        // FP32 immersed-boundary sweeps skip empty space, so the work per
        // chunk is uneven: shorter chunks balance it better
        if (ib && sizeof(T) == 4) fg.kch = std::min(fg.kch, 32);
    }
    const int nbz = (fg.hi[2] - fg.lo[2] + fg.kch) / fg.kch;
    const int ntiles = fg.nbx * nby * nbz;
    const int nfused = (ntiles + 8 * NXCD - 1) / (8 * NXCD) * (8 * NXCD);
    BoxList sh;
    sh.nb = 0;
    sh.start[0] = 0;
    if (shell_sigma) {
        // box starts in workgroups of tx * ty threads
        sh = *shell_sigma;
        for (int q = 0; q < sh.nb; ++q) {
            const Box& b = sh.b[q];
            const long n = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
            sh.start[q + 1] = sh.start[q] + (int)((n + tx * ty - 1) / (tx * ty));
        }
    }
    dim3 grd(nfused + (sh.nb ? sh.start[sh.nb] : 0)), blk(tx, ty, 1);
    // FP32 sweeps with frozen viscosity run the variant compiled without the viscosity update
    const bool visc = sizeof(T) == 8 || !a.frozen_eta;
#define FI_LAUNCH_(TX_, TY_, W_)                                                                                    \
    if (tx == TX_ && ty == TY_ && minw == W_) {                                                                     \
        if constexpr (sizeof(T) == 4) {                                                                             \
            if (!visc) {                                                                                            \
                if (ib) hipLaunchKernelGGL((k_fused3<T, true, TX_, TY_, W_, false>), grd, blk, 0, s, a, fg, sh, nfused); \
                else hipLaunchKernelGGL((k_fused3<T, false, TX_, TY_, W_, false>), grd, blk, 0, s, a, fg, sh, nfused);   \
                return;                                                                                             \
            }                                                                                                       \
        }                                                                                                           \
        if (ib) hipLaunchKernelGGL((k_fused3<T, true, TX_, TY_, W_, true>), grd, blk, 0, s, a, fg, sh, nfused);    \
        else hipLaunchKernelGGL((k_fused3<T, false, TX_, TY_, W_, true>), grd, blk, 0, s, a, fg, sh, nfused);      \
        return;                                                                                                     \
    }
    if constexpr (sizeof(T) == 8) { FI_FUSED_CFG_D(FI_LAUNCH_) }
    else { FI_FUSED_CFG_F(FI_LAUNCH_) }
#undef FI_LAUNCH_
}

template <typename T>
void launch_fused(const Geo& g, const StokesArgs<T>& a, const FusedGeom& fg0, bool ib, int tx, int ty, int minw,
                  hipStream_t s, const BoxList* shell_sigma)
{
    FusedGeom fg;
    launch_fused_sweep(g, a, fg0, ib, tx, ty, minw, s, shell_sigma, fg);
    // a narrowed immersed-boundary sweep with a viscosity update leaves the
    // viscosity of the planes it did not march to k_eta_rest
    if (ib && !a.frozen_eta && a.z_wsig && a.z_wvel && a.tmarch) {
        // three scalars after the tile entries of tmarch: the scaled target
        // and the relaxation of the viscosity of a cell on the region's top
        // plane (in a glacier box, air: empty space)
        T* st0 = reinterpret_cast<T*>(a.tmarch + tile_march_count(g));
        const uint32_t probe = (uint32_t)((fg.lo[0] + g.sy * fg.lo[1] + g.sz * fg.hi[2]) * (long)sizeof(T));
        if (!a.rh) hipLaunchKernelGGL((k_eta_target<T>), dim3(1), dim3(64), 0, s, a, st0, probe);
        const int nbz = (fg.hi[2] - fg.lo[2] + fg.kch) / fg.kch;
        dim3 grd((fg.hi[0] - fg.lo[0] + 64) / 64, (fg.hi[1] - fg.lo[1] + 4) / 4, nbz);
        hipLaunchKernelGGL((k_eta_rest<T>), grd, dim3(64, 4, 1), 0, s, a, fg, tx, ty, st0);
    }
}

template <typename T>
void launch_residuals(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s)
{
    if (g.nd == 3) {
        if (ib) hipLaunchKernelGGL((k_residuals3<T, true>), grid3(g), block2(), 0, s, a);
        else hipLaunchKernelGGL((k_residuals3<T, false>), grid3(g), block2(), 0, s, a);
    } else {
        if (ib) hipLaunchKernelGGL((k_residuals2<T, true>), grid2(g), block2(), 0, s, a);
        else hipLaunchKernelGGL((k_residuals2<T, false>), grid2(g), block2(), 0, s, a);
    }
}

template <typename T>
void launch_tau(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s, const TauLag* lag)
{
    TauLag lg{};
    if (lag) lg = *lag;
    if (g.nd == 3) {
        if (lag) {
            if (ib) hipLaunchKernelGGL((k_tau3<T, true, true>), grid3(g), block2(), 0, s, a, lg);
            else hipLaunchKernelGGL((k_tau3<T, false, true>), grid3(g), block2(), 0, s, a, lg);
        } else {
            if (ib) hipLaunchKernelGGL((k_tau3<T, true, false>), grid3(g), block2(), 0, s, a, lg);
            else hipLaunchKernelGGL((k_tau3<T, false, false>), grid3(g), block2(), 0, s, a, lg);
        }
    } else {
        if (lag) {
            if (ib) hipLaunchKernelGGL((k_tau2<T, true, true>), grid2(g), block2(), 0, s, a, lg);
            else hipLaunchKernelGGL((k_tau2<T, false, true>), grid2(g), block2(), 0, s, a, lg);
        } else {
            if (ib) hipLaunchKernelGGL((k_tau2<T, true, false>), grid2(g), block2(), 0, s, a, lg);
            else hipLaunchKernelGGL((k_tau2<T, false, false>), grid2(g), block2(), 0, s, a, lg);
        }
    }
}

template <typename T>
void launch_fused_packed(const PackedInst<T>* inst, const int4* items, int nitems, bool ib, bool visc, int tx, int ty,
                         int minw, int grid, hipStream_t s)
{
    if (nitems <= 0) return;
    int gsz = nitems < grid ? nitems : grid;
    dim3 grd((gsz + 8 * NXCD - 1) / (8 * NXCD) * (8 * NXCD)), blk(tx, ty);
#define FI_LAUNCH_(TX_, TY_, W_)                                                                                        \
    if (tx == TX_ && ty == TY_ && minw == W_) {                                                                         \
        if constexpr (sizeof(T) == 4) {                                                                                 \
            if (!visc) {                                                                                                \
                if (ib) hipLaunchKernelGGL((k_fused3_packed<T, true, TX_, TY_, W_, false>), grd, blk, 0, s, inst, items, nitems); \
                else hipLaunchKernelGGL((k_fused3_packed<T, false, TX_, TY_, W_, false>), grd, blk, 0, s, inst, items, nitems);   \
                return;                                                                                                 \
            }                                                                                                           \
        }                                                                                                               \
        if (ib) hipLaunchKernelGGL((k_fused3_packed<T, true, TX_, TY_, W_, true>), grd, blk, 0, s, inst, items, nitems);  \
        else hipLaunchKernelGGL((k_fused3_packed<T, false, TX_, TY_, W_, true>), grd, blk, 0, s, inst, items, nitems);    \
        return;                                                                                                         \
    }
    if constexpr (sizeof(T) == 8) { FI_FUSED_CFG_D(FI_LAUNCH_) }
    else { FI_FUSED_CFG_F(FI_LAUNCH_) }
#undef FI_LAUNCH_
}

template <typename T>
void launch_point_packed(const PackedInst<T>* inst, const PackedBox* boxes, const int* blk2box, int nblocks, bool ib,
                         int nd, bool sigma, hipStream_t s)
{
    if (nblocks <= 0) return;
    dim3 grd(nblocks), blk(256);
#define FI_PP(IB_, ND_, S_) hipLaunchKernelGGL((k_point_packed<T, IB_, ND_, S_>), grd, blk, 0, s, inst, boxes, blk2box)
    if (nd == 3) {
        if (ib) { if (sigma) FI_PP(true, 3, true); else FI_PP(true, 3, false); }
        else { if (sigma) FI_PP(false, 3, true); else FI_PP(false, 3, false); }
    } else {
        if (ib) { if (sigma) FI_PP(true, 2, true); else FI_PP(true, 2, false); }
        else { if (sigma) FI_PP(false, 2, true); else FI_PP(false, 2, false); }
    }
#undef FI_PP
}

#define INST(T)                                                                                               \
    template void launch_sigma_boxes<T>(const Geo&, const StokesArgs<T>&, const BoxList&, bool, hipStream_t);    \
    template void launch_velocity_boxes<T>(const Geo&, const StokesArgs<T>&, const BoxList&, bool, hipStream_t); \
    template void launch_fused<T>(const Geo&, const StokesArgs<T>&, const FusedGeom&, bool, int, int, int,       \
                                  hipStream_t, const BoxList*);                                                 \
    template void launch_residuals<T>(const Geo&, const StokesArgs<T>&, bool, hipStream_t);                      \
    template void launch_tau<T>(const Geo&, const StokesArgs<T>&, bool, hipStream_t, const TauLag*);                      \
    template void launch_eta_update<T>(const Geo&, const StokesArgs<T>&, hipStream_t);                     \
    template void launch_fused_packed<T>(const PackedInst<T>*, const int4*, int, bool, bool, int, int, int, int,   \
                                         hipStream_t);                                                          \
    template void launch_point_packed<T>(const PackedInst<T>*, const PackedBox*, const int*, int, bool, int, bool, \
                                         hipStream_t);
INST(double)
INST(float)

} // namespace fi
