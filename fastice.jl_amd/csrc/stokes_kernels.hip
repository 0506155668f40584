// stokes_kernels.hip -- PT-Stokes sweep kernels for gfx950.
//
// One thread per staggered index I in [1, n+1]^d, x fastest across the
// wavefront (coalesced 64-wide rows).  Every expression keeps the operation
// order of the reference source line it cites (Julia evaluates left to right)
// and the library is compiled with -ffp-contract=off, so the FP64 kernels
// round exactly like the CPU restatement of the same formulas.
//
// Discrete operators (Chmy semantics, fixed by this library):
//   d/dx of a Vertex-in-x field at a Center index: (f[i+1]-f[i]) * (1/dx)
//   d/dx of a Center-in-x field at a Vertex index: (f[i]-f[i-1]) * (1/dx)
//   lerp Center->Vertex 0.5*(f[i-1]+f[i]); Vertex->Center 0.5*(f[i]+f[i+1]),
//   nested x innermost, then y, then z.
//   masked d/dx(f, w): (w[r]*f[r] - w[l]*f[l]) * (1/dx), w at f's location.
#include "fi_internal.h"

namespace fi {

#define TOL_EMPTY 1e-6 // null_spaces.jl:16,41

template <typename T>
__device__ __forceinline__ T ipow_(T x, T p)
{
    if (p == T(0)) return T(1);
    if (p == T(1)) return x;
    if (p == T(2)) return x * x;
    if (p == T(3)) return x * x * x;
    return pow(x, p);
}

// Physics.jl:11,18,26,46-47
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

// null_spaces.jl:1-43 in three dimensions
template <typename T>
__device__ __forceinline__ bool lt(const T* w, long c)
{
    return w[c] < T(TOL_EMPTY);
}

template <typename T>
__device__ __forceinline__ bool null3(const T* const* w, int loc, long c, long sy, long sz)
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

__device__ __forceinline__ int comp3(int loc)
{
    switch (loc) {
    case L_CCC: return OM_CCC;
    case L_VCC: return OM_VCC;
    case L_CVC: return OM_CVC;
    case L_CCV: return OM_CCV;
    case L_VVC: return OM_VVC;
    case L_VCV: return OM_VCV;
    case L_CVV: return OM_CVV;
    default: return OM_VVV;
    }
}

// isvalid(w, loc, I) && !isempty(w2, loc, I)
template <typename T>
__device__ __forceinline__ bool valid3(const T* const* w, const T* const* w2, int loc, long c, long sy, long sz)
{
    int k = comp3(loc);
    return !(null3(w, loc, c, sy, sz) || lt(w[k], c)) && !lt(w2[k], c);
}

template <typename T>
__device__ __forceinline__ bool null2(const T* const* w, int loc, long c, long sy)
{
    switch (loc) {
    case L_CCC: return lt(w[OM2_VC], c) || lt(w[OM2_VC], c + 1) || lt(w[OM2_CV], c) || lt(w[OM2_CV], c + sy);
    case L_VCC: return lt(w[OM2_CC], c - 1) || lt(w[OM2_CC], c) || lt(w[OM2_VV], c) || lt(w[OM2_VV], c + sy);
    case L_CVC: return lt(w[OM2_VV], c) || lt(w[OM2_VV], c + 1) || lt(w[OM2_CC], c - sy) || lt(w[OM2_CC], c);
    case L_VVC: return lt(w[OM2_CV], c - 1) || lt(w[OM2_CV], c) || lt(w[OM2_VC], c - sy) || lt(w[OM2_VC], c);
    }
    return true;
}

__device__ __forceinline__ int comp2(int loc)
{
    switch (loc) {
    case L_CCC: return OM2_CC;
    case L_VCC: return OM2_VC;
    case L_CVC: return OM2_CV;
    default: return OM2_VV;
    }
}

template <typename T>
__device__ __forceinline__ bool valid2(const T* const* w, const T* const* w2, int loc, long c, long sy)
{
    int k = comp2(loc);
    return !(null2(w, loc, c, sy) || lt(w[k], c)) && !lt(w2[k], c);
}

// ---------------------------------------------------------------------------
// 3D
// ---------------------------------------------------------------------------
#define IDX3                                                                      \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;                      \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;                      \
    const int k = blockIdx.z + 1;                                                 \
    if (i > a.nx + 1 || j > a.ny + 1) return;                                     \
    const long sy = a.sy, sz = a.sz;                                              \
    const long c = (long)i + sy * j + sz * k;

// update_sigma!: FullStokes kernels_3d.jl:3-25, IB kernels_3d.jl:141-182
template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_sigma3(StokesArgs<T> a)
{
    IDX3
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    T *P = a.P, *txx = a.txx, *tyy = a.tyy, *tzz = a.tzz, *txy = a.txy, *txz = a.txz, *tyz = a.tyz;
    const T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz, *eta = a.eta;
    if (!IB) {
        T exx = (Vx[c + 1] - Vx[c]) * ix;
        T eyy = (Vy[c + sy] - Vy[c]) * iy;
        T ezz = (Vz[c + sz] - Vz[c]) * iz;
        T exy = T(0.5) * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
        T exz = T(0.5) * ((Vz[c] - Vz[c - 1]) * ix + (Vx[c] - Vx[c - sz]) * iz);
        T eyz = T(0.5) * ((Vz[c] - Vz[c - sy]) * iy + (Vy[c] - Vy[c - sz]) * iz);
        T divV = exx + eyy + ezz;
        T ec = eta[c];
        T exy_eta = T(0.5) * (T(0.5) * (eta[c - 1 - sy] + eta[c - sy]) + T(0.5) * (eta[c - 1] + eta[c]));
        T exz_eta = T(0.5) * (T(0.5) * (eta[c - 1 - sz] + eta[c - sz]) + T(0.5) * (eta[c - 1] + eta[c]));
        T eyz_eta = T(0.5) * (T(0.5) * (eta[c - sy - sz] + eta[c - sz]) + T(0.5) * (eta[c - sy] + eta[c]));
        P[c] = P[c] - divV * ec * a.dPr;
        txx[c] = txx[c] - (txx[c] - T(2) * ec * (exx - divV / T(3))) * a.dtau[0];
        tyy[c] = tyy[c] - (tyy[c] - T(2) * ec * (eyy - divV / T(3))) * a.dtau[1];
        tzz[c] = tzz[c] - (tzz[c] - T(2) * ec * (ezz - divV / T(3))) * a.dtau[2];
        txy[c] = txy[c] - (txy[c] - T(2) * exy_eta * exy) * a.dtau[3];
        txz[c] = txz[c] - (txz[c] - T(2) * exz_eta * exz) * a.dtau[4];
        tyz[c] = tyz[c] - (tyz[c] - T(2) * eyz_eta * eyz) * a.dtau[5];
    } else {
        const T* const* na = a.na;
        const T* const* ns = a.ns;
        const T *wx = ns[OM_VCC], *wy = ns[OM_CVC], *wz = ns[OM_CCV];
        if (valid3(na, ns, L_CCC, c, sy, sz)) {
            T exx = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix;
            T eyy = (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy;
            T ezz = (wz[c + sz] * Vz[c + sz] - wz[c] * Vz[c]) * iz;
            T divV = exx + eyy + ezz;
            T ec = eta[c];
            P[c] = P[c] - divV * ec * a.dPr;
            txx[c] = txx[c] - (txx[c] - T(2) * ec * (exx - divV / T(3))) * a.dtau[0];
            tyy[c] = tyy[c] - (tyy[c] - T(2) * ec * (eyy - divV / T(3))) * a.dtau[1];
            tzz[c] = tzz[c] - (tzz[c] - T(2) * ec * (ezz - divV / T(3))) * a.dtau[2];
        } else {
            P[c] = T(0); txx[c] = T(0); tyy[c] = T(0); tzz[c] = T(0);
        }
        if (valid3(na, ns, L_VVC, c, sy, sz)) {
            T exy = T(0.5) * ((wy[c] * Vy[c] - wy[c - 1] * Vy[c - 1]) * ix + (wx[c] * Vx[c] - wx[c - sy] * Vx[c - sy]) * iy);
            T e = T(0.5) * (T(0.5) * (eta[c - 1 - sy] + eta[c - sy]) + T(0.5) * (eta[c - 1] + eta[c]));
            txy[c] = txy[c] - (txy[c] - T(2) * e * exy) * a.dtau[3];
        } else txy[c] = T(0);
        if (valid3(na, ns, L_VCV, c, sy, sz)) {
            T exz = T(0.5) * ((wz[c] * Vz[c] - wz[c - 1] * Vz[c - 1]) * ix + (wx[c] * Vx[c] - wx[c - sz] * Vx[c - sz]) * iz);
            T e = T(0.5) * (T(0.5) * (eta[c - 1 - sz] + eta[c - sz]) + T(0.5) * (eta[c - 1] + eta[c]));
            txz[c] = txz[c] - (txz[c] - T(2) * e * exz) * a.dtau[4];
        } else txz[c] = T(0);
        if (valid3(na, ns, L_CVV, c, sy, sz)) {
            T eyz = T(0.5) * ((wz[c] * Vz[c] - wz[c - sy] * Vz[c - sy]) * iy + (wy[c] * Vy[c] - wy[c - sz] * Vy[c - sz]) * iz);
            T e = T(0.5) * (T(0.5) * (eta[c - sy - sz] + eta[c - sz]) + T(0.5) * (eta[c - sy] + eta[c]));
            tyz[c] = tyz[c] - (tyz[c] - T(2) * e * eyz) * a.dtau[5];
        } else tyz[c] = T(0);
    }
}

// tau_II at a cell centre: kernels_3d.jl:162-165
template <typename T>
__device__ __forceinline__ T tauII3(const StokesArgs<T>& a, long c, long sy, long sz)
{
    const T *txy = a.txy, *txz = a.txz, *tyz = a.tyz;
    T p = a.txx[c], q = a.tyy[c], r = a.tzz[c];
    T lxy = T(0.5) * (T(0.5) * (txy[c] + txy[c + 1]) + T(0.5) * (txy[c + sy] + txy[c + 1 + sy]));
    T lxz = T(0.5) * (T(0.5) * (txz[c] + txz[c + 1]) + T(0.5) * (txz[c + sz] + txz[c + 1 + sz]));
    T lyz = T(0.5) * (T(0.5) * (tyz[c] + tyz[c + sy]) + T(0.5) * (tyz[c + sz] + tyz[c + sy + sz]));
    return sqrt(T(0.5) * (p * p + q * q + r * r) + lxy * lxy + lxz * lxz + lyz * lyz);
}

// update_V!: FullStokes kernels_3d.jl:154-167, IB kernels_3d.jl:184-212
template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_velocity3(StokesArgs<T> a)
{
    IDX3
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *P = a.P, *txx = a.txx, *tyy = a.tyy, *tzz = a.tzz, *txy = a.txy, *txz = a.txz, *tyz = a.tyz;
    T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz;
    const T* eta = a.eta;
    T gx = a.grav_field ? a.gx[c] : a.grav[0];
    T gy = a.grav_field ? a.gy[c] : a.grav[1];
    T gz = a.grav_field ? a.gz[c] : a.grav[2];
    T ex = T(0.5) * (eta[c - 1] + eta[c]);
    T ey = T(0.5) * (eta[c - sy] + eta[c]);
    T ez = T(0.5) * (eta[c - sz] + eta[c]);
    if (!IB) {
        Vx[c] = Vx[c] + (-((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy +
                         (txz[c + sz] - txz[c]) * iz - gx) / ex * a.dV[0];
        Vy[c] = Vy[c] + (-((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix +
                         (tyz[c + sz] - tyz[c]) * iz - gy) / ey * a.dV[1];
        Vz[c] = Vz[c] + (-((P[c] - P[c - sz]) * iz) + (tzz[c] - tzz[c - sz]) * iz + (txz[c + 1] - txz[c]) * ix +
                         (tyz[c + sy] - tyz[c]) * iy - gz) / ez * a.dV[2];
    } else {
        const T* const* na = a.na;
        const T* const* ns = a.ns;
        const T *wc = na[OM_CCC], *wxy = na[OM_VVC], *wxz = na[OM_VCV], *wyz = na[OM_CVV];
        if (valid3(ns, na, L_VCC, c, sy, sz))
            Vx[c] = Vx[c] + (-((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                             (wxy[c + sy] * txy[c + sy] - wxy[c] * txy[c]) * iy +
                             (wxz[c + sz] * txz[c + sz] - wxz[c] * txz[c]) * iz - gx) / ex / ns[OM_VCC][c] * a.dV[0];
        else Vx[c] = T(0);
        if (valid3(ns, na, L_CVC, c, sy, sz))
            Vy[c] = Vy[c] + (-((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                             (wxy[c + 1] * txy[c + 1] - wxy[c] * txy[c]) * ix +
                             (wyz[c + sz] * tyz[c + sz] - wyz[c] * tyz[c]) * iz - gy) / ey / ns[OM_CVC][c] * a.dV[1];
        else Vy[c] = T(0);
        if (valid3(ns, na, L_CCV, c, sy, sz))
            Vz[c] = Vz[c] + (-((wc[c] * P[c] - wc[c - sz] * P[c - sz]) * iz) + (wc[c] * tzz[c] - wc[c - sz] * tzz[c - sz]) * iz +
                             (wxz[c + 1] * txz[c + 1] - wxz[c] * txz[c]) * ix +
                             (wyz[c + sy] * tyz[c + sy] - wyz[c] * tyz[c]) * iy - gz) / ez / ns[OM_CCV][c] * a.dV[2];
        else Vz[c] = T(0);
    }
    a.eta_next[c] = rheology(a, c, tauII3(a, c, sy, sz));
}

// compute_residuals!: FullStokes kernels_3d.jl:193-202, IB kernels_3d.jl:256-281
template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_residuals3(StokesArgs<T> a)
{
    IDX3
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *P = a.P, *txx = a.txx, *tyy = a.tyy, *tzz = a.tzz, *txy = a.txy, *txz = a.txz, *tyz = a.tyz;
    const T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz;
    T gx = a.grav_field ? a.gx[c] : a.grav[0];
    T gy = a.grav_field ? a.gy[c] : a.grav[1];
    T gz = a.grav_field ? a.gz[c] : a.grav[2];
    if (!IB) {
        a.rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy + (Vz[c + sz] - Vz[c]) * iz;
        a.rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy +
                   (txz[c + sz] - txz[c]) * iz - gx;
        a.rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix +
                   (tyz[c + sz] - tyz[c]) * iz - gy;
        a.rVz[c] = -((P[c] - P[c - sz]) * iz) + (tzz[c] - tzz[c - sz]) * iz + (txz[c + 1] - txz[c]) * ix +
                   (tyz[c + sy] - tyz[c]) * iy - gz;
    } else {
        const T* const* na = a.na;
        const T* const* ns = a.ns;
        const T *wx = ns[OM_VCC], *wy = ns[OM_CVC], *wz = ns[OM_CCV];
        const T *wc = na[OM_CCC], *wxy = na[OM_VVC], *wxz = na[OM_VCV], *wyz = na[OM_CVV];
        if (valid3(na, ns, L_CCC, c, sy, sz))
            a.rP[c] = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix + (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy +
                      (wz[c + sz] * Vz[c + sz] - wz[c] * Vz[c]) * iz;
        else a.rP[c] = T(0);
        if (valid3(ns, na, L_VCC, c, sy, sz))
            a.rVx[c] = -((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                       (wxy[c + sy] * txy[c + sy] - wxy[c] * txy[c]) * iy + (wxz[c + sz] * txz[c + sz] - wxz[c] * txz[c]) * iz - gx;
        else a.rVx[c] = T(0);
        if (valid3(ns, na, L_CVC, c, sy, sz))
            a.rVy[c] = -((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                       (wxy[c + 1] * txy[c + 1] - wxy[c] * txy[c]) * ix + (wyz[c + sz] * tyz[c + sz] - wyz[c] * tyz[c]) * iz - gy;
        else a.rVy[c] = T(0);
        if (valid3(ns, na, L_CCV, c, sy, sz))
            a.rVz[c] = -((wc[c] * P[c] - wc[c - sz] * P[c - sz]) * iz) + (wc[c] * tzz[c] - wc[c - sz] * tzz[c - sz]) * iz +
                       (wxz[c + 1] * txz[c + 1] - wxz[c] * txz[c]) * ix + (wyz[c + sy] * tyz[c + sy] - wyz[c] * tyz[c]) * iy - gz;
        else a.rVz[c] = T(0);
    }
}

// compute_tau!: kernels_3d.jl:171-191 (IB variant: kernels_3d.jl:216-254 with
// its undefined `Pr` location read as the Center location)
template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_tau3(StokesArgs<T> a)
{
    IDX3
    const T ix = a.ix, iy = a.iy, iz = a.iz;
    const T *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz, *eta = a.eta;
    const T* const* na = a.na;
    const T* const* ns = a.ns;
    T ec = eta[c];
    if (!IB || valid3(na, ns, L_CCC, c, sy, sz)) {
        T exx, eyy, ezz;
        if (IB) {
            exx = (ns[OM_VCC][c + 1] * Vx[c + 1] - ns[OM_VCC][c] * Vx[c]) * ix;
            eyy = (ns[OM_CVC][c + sy] * Vy[c + sy] - ns[OM_CVC][c] * Vy[c]) * iy;
            ezz = (ns[OM_CCV][c + sz] * Vz[c + sz] - ns[OM_CCV][c] * Vz[c]) * iz;
        } else {
            exx = (Vx[c + 1] - Vx[c]) * ix;
            eyy = (Vy[c + sy] - Vy[c]) * iy;
            ezz = (Vz[c + sz] - Vz[c]) * iz;
        }
        T divV = exx + eyy + ezz;
        a.txx[c] = T(2) * ec * (exx - divV / T(3));
        a.tyy[c] = T(2) * ec * (eyy - divV / T(3));
        a.tzz[c] = T(2) * ec * (ezz - divV / T(3));
    } else {
        a.txx[c] = T(0); a.tyy[c] = T(0); a.tzz[c] = T(0);
    }
    if (!IB || valid3(na, ns, L_VVC, c, sy, sz)) {
        T exy;
        if (IB) exy = T(0.5) * ((ns[OM_CVC][c] * Vy[c] - ns[OM_CVC][c - 1] * Vy[c - 1]) * ix +
                                (ns[OM_VCC][c] * Vx[c] - ns[OM_VCC][c - sy] * Vx[c - sy]) * iy);
        else exy = T(0.5) * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
        T e = T(0.5) * (T(0.5) * (eta[c - 1 - sy] + eta[c - sy]) + T(0.5) * (eta[c - 1] + eta[c]));
        a.txy[c] = T(2) * e * exy;
    } else a.txy[c] = T(0);
    if (!IB || valid3(na, ns, L_VCV, c, sy, sz)) {
        T exz;
        if (IB) exz = T(0.5) * ((ns[OM_CCV][c] * Vz[c] - ns[OM_CCV][c - 1] * Vz[c - 1]) * ix +
                                (ns[OM_VCC][c] * Vx[c] - ns[OM_VCC][c - sz] * Vx[c - sz]) * iz);
        else exz = T(0.5) * ((Vz[c] - Vz[c - 1]) * ix + (Vx[c] - Vx[c - sz]) * iz);
        T e = T(0.5) * (T(0.5) * (eta[c - 1 - sz] + eta[c - sz]) + T(0.5) * (eta[c - 1] + eta[c]));
        a.txz[c] = T(2) * e * exz;
    } else a.txz[c] = T(0);
    if (!IB || valid3(na, ns, L_CVV, c, sy, sz)) {
        T eyz;
        if (IB) eyz = T(0.5) * ((ns[OM_CCV][c] * Vz[c] - ns[OM_CCV][c - sy] * Vz[c - sy]) * iy +
                                (ns[OM_CVC][c] * Vy[c] - ns[OM_CVC][c - sz] * Vy[c - sz]) * iz);
        else eyz = T(0.5) * ((Vz[c] - Vz[c - sy]) * iy + (Vy[c] - Vy[c - sz]) * iz);
        T e = T(0.5) * (T(0.5) * (eta[c - sy - sz] + eta[c - sz]) + T(0.5) * (eta[c - sy] + eta[c]));
        a.tyz[c] = T(2) * e * eyz;
    } else a.tyz[c] = T(0);
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
__global__ __launch_bounds__(256) void k_sigma2(StokesArgs<T> a)
{
    IDX2
    const T ix = a.ix, iy = a.iy;
    T *P = a.P, *txx = a.txx, *tyy = a.tyy, *txy = a.txy;
    const T *Vx = a.Vx, *Vy = a.Vy, *eta = a.eta;
    T ec = eta[c];
    T exy_eta = T(0.5) * (T(0.5) * (eta[c - 1 - sy] + eta[c - sy]) + T(0.5) * (eta[c - 1] + eta[c]));
    if (!IB) {
        // kernels_2d.jl:205-220
        T exx = (Vx[c + 1] - Vx[c]) * ix;
        T eyy = (Vy[c + sy] - Vy[c]) * iy;
        T exy = T(0.5) * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
        T divV = exx + eyy;
        P[c] = P[c] - divV * ec * a.dPr;
        txx[c] = txx[c] - (txx[c] - T(2) * ec * (exx - divV / T(3))) * a.dtau[0];
        tyy[c] = tyy[c] - (tyy[c] - T(2) * ec * (eyy - divV / T(3))) * a.dtau[1];
        txy[c] = txy[c] - (txy[c] - T(2) * exy_eta * exy) * a.dtau[2];
    } else {
        const T* const* na = a.na;
        const T* const* ns = a.ns;
        const T *wx = ns[OM2_VC], *wy = ns[OM2_CV];
        if (valid2(na, ns, L_CCC, c, sy)) {
            T exx = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix;
            T eyy = (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy;
            T divV = exx + eyy;
            P[c] = P[c] - divV * ec * a.dPr;
            txx[c] = txx[c] - (txx[c] - T(2) * ec * (exx - divV / T(3))) * a.dtau[0];
            tyy[c] = tyy[c] - (tyy[c] - T(2) * ec * (eyy - divV / T(3))) * a.dtau[1];
        } else {
            P[c] = T(0); txx[c] = T(0); tyy[c] = T(0);
        }
        if (valid2(na, ns, L_VVC, c, sy)) {
            T exy = T(0.5) * ((wy[c] * Vy[c] - wy[c - 1] * Vy[c - 1]) * ix + (wx[c] * Vx[c] - wx[c - sy] * Vx[c - sy]) * iy);
            txy[c] = txy[c] - (txy[c] - T(2) * exy_eta * exy) * a.dtau[2];
        } else txy[c] = T(0);
    }
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_velocity2(StokesArgs<T> a)
{
    IDX2
    const T ix = a.ix, iy = a.iy;
    const T *P = a.P, *txx = a.txx, *tyy = a.tyy, *txy = a.txy, *eta = a.eta;
    T *Vx = a.Vx, *Vy = a.Vy;
    T gx = a.grav_field ? a.gx[c] : a.grav[0];
    T gy = a.grav_field ? a.gy[c] : a.grav[1];
    // each component interpolates eta to its own location (kernels_2d.jl:226-227
    // reuses location(V.x) for both)
    T ex = T(0.5) * (eta[c - 1] + eta[c]);
    T ey = T(0.5) * (eta[c - sy] + eta[c]);
    if (!IB) {
        Vx[c] = Vx[c] + (-((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy - gx) / ex * a.dV[0];
        Vy[c] = Vy[c] + (-((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix - gy) / ey * a.dV[1];
    } else {
        const T* const* na = a.na;
        const T* const* ns = a.ns;
        const T *wc = na[OM2_CC], *wv = na[OM2_VV];
        if (valid2(ns, na, L_VCC, c, sy))
            Vx[c] = Vx[c] + (-((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                             (wv[c + sy] * txy[c + sy] - wv[c] * txy[c]) * iy - gx) / ex / ns[OM2_VC][c] * a.dV[0];
        else Vx[c] = T(0);
        if (valid2(ns, na, L_CVC, c, sy))
            Vy[c] = Vy[c] + (-((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                             (wv[c + 1] * txy[c + 1] - wv[c] * txy[c]) * ix - gy) / ey / ns[OM2_CV][c] * a.dV[1];
        else Vy[c] = T(0);
    }
    // kernels_2d.jl:229-230
    T p = txx[c], q = tyy[c];
    T l = T(0.5) * (T(0.5) * (txy[c] + txy[c + 1]) + T(0.5) * (txy[c + sy] + txy[c + 1 + sy]));
    a.eta_next[c] = rheology(a, c, sqrt(T(0.5) * (p * p + q * q) + l * l));
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
        // kernels_2d.jl:248-256
        a.rP[c] = (Vx[c + 1] - Vx[c]) * ix + (Vy[c + sy] - Vy[c]) * iy;
        a.rVx[c] = -((P[c] - P[c - 1]) * ix) + (txx[c] - txx[c - 1]) * ix + (txy[c + sy] - txy[c]) * iy - gx;
        a.rVy[c] = -((P[c] - P[c - sy]) * iy) + (tyy[c] - tyy[c - sy]) * iy + (txy[c + 1] - txy[c]) * ix - gy;
    } else {
        const T* const* na = a.na;
        const T* const* ns = a.ns;
        const T *wx = ns[OM2_VC], *wy = ns[OM2_CV], *wc = na[OM2_CC], *wv = na[OM2_VV];
        if (valid2(na, ns, L_CCC, c, sy))
            a.rP[c] = (wx[c + 1] * Vx[c + 1] - wx[c] * Vx[c]) * ix + (wy[c + sy] * Vy[c + sy] - wy[c] * Vy[c]) * iy;
        else a.rP[c] = T(0);
        if (valid2(ns, na, L_VCC, c, sy))
            a.rVx[c] = -((wc[c] * P[c] - wc[c - 1] * P[c - 1]) * ix) + (wc[c] * txx[c] - wc[c - 1] * txx[c - 1]) * ix +
                       (wv[c + sy] * txy[c + sy] - wv[c] * txy[c]) * iy - gx;
        else a.rVx[c] = T(0);
        if (valid2(ns, na, L_CVC, c, sy))
            a.rVy[c] = -((wc[c] * P[c] - wc[c - sy] * P[c - sy]) * iy) + (wc[c] * tyy[c] - wc[c - sy] * tyy[c - sy]) * iy +
                       (wv[c + 1] * txy[c + 1] - wv[c] * txy[c]) * ix - gy;
        else a.rVy[c] = T(0);
    }
}

template <typename T, bool IB>
__global__ __launch_bounds__(256) void k_tau2(StokesArgs<T> a)
{
    IDX2
    const T ix = a.ix, iy = a.iy;
    const T *Vx = a.Vx, *Vy = a.Vy, *eta = a.eta;
    const T* const* na = a.na;
    const T* const* ns = a.ns;
    T ec = eta[c];
    T exy_eta = T(0.5) * (T(0.5) * (eta[c - 1 - sy] + eta[c - sy]) + T(0.5) * (eta[c - 1] + eta[c]));
    if (!IB || valid2(na, ns, L_CCC, c, sy)) {
        T exx, eyy;
        if (IB) {
            exx = (ns[OM2_VC][c + 1] * Vx[c + 1] - ns[OM2_VC][c] * Vx[c]) * ix;
            eyy = (ns[OM2_CV][c + sy] * Vy[c + sy] - ns[OM2_CV][c] * Vy[c]) * iy;
        } else {
            exx = (Vx[c + 1] - Vx[c]) * ix;
            eyy = (Vy[c + sy] - Vy[c]) * iy;
        }
        T divV = exx + eyy;
        a.txx[c] = T(2) * ec * (exx - divV / T(3));
        a.tyy[c] = T(2) * ec * (eyy - divV / T(3));
    } else {
        a.txx[c] = T(0); a.tyy[c] = T(0);
    }
    if (!IB || valid2(na, ns, L_VVC, c, sy)) {
        T exy;
        if (IB) exy = T(0.5) * ((ns[OM2_CV][c] * Vy[c] - ns[OM2_CV][c - 1] * Vy[c - 1]) * ix +
                                (ns[OM2_VC][c] * Vx[c] - ns[OM2_VC][c - sy] * Vx[c - sy]) * iy);
        else exy = T(0.5) * ((Vy[c] - Vy[c - 1]) * ix + (Vx[c] - Vx[c - sy]) * iy);
        a.txy[c] = T(2) * exy_eta * exy;
    } else a.txy[c] = T(0);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 block3() { return dim3(64, 4, 1); }
static inline dim3 grid3(const Geo& g) { return dim3((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[2] + 1); }
static inline dim3 block2() { return dim3(64, 4, 1); }
static inline dim3 grid2(const Geo& g) { return dim3((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, 1); }

#define DISPATCH(K3, K2)                                                                          \
    if (g.nd == 3) {                                                                              \
        if (ib) hipLaunchKernelGGL((K3<T, true>), grid3(g), block3(), 0, s, a);                   \
        else hipLaunchKernelGGL((K3<T, false>), grid3(g), block3(), 0, s, a);                     \
    } else {                                                                                      \
        if (ib) hipLaunchKernelGGL((K2<T, true>), grid2(g), block2(), 0, s, a);                   \
        else hipLaunchKernelGGL((K2<T, false>), grid2(g), block2(), 0, s, a);                     \
    }

template <typename T>
void launch_sigma(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s)
{
    DISPATCH(k_sigma3, k_sigma2)
}
template <typename T>
void launch_velocity(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s)
{
    DISPATCH(k_velocity3, k_velocity2)
}
template <typename T>
void launch_residuals(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s)
{
    DISPATCH(k_residuals3, k_residuals2)
}
template <typename T>
void launch_tau(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s)
{
    DISPATCH(k_tau3, k_tau2)
}

template void launch_sigma<double>(const Geo&, const StokesArgs<double>&, bool, hipStream_t);
template void launch_velocity<double>(const Geo&, const StokesArgs<double>&, bool, hipStream_t);
template void launch_residuals<double>(const Geo&, const StokesArgs<double>&, bool, hipStream_t);
template void launch_tau<double>(const Geo&, const StokesArgs<double>&, bool, hipStream_t);

} // namespace fi
