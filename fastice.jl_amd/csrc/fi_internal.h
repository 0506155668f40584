// fi_internal.h -- shared definitions of libfasticehip.so (gfx950).
//
// Device layout of a field: one allocation per field, x-fastest, with the
// uniform staggered box of (nx+3) x (ny+3) x (nz+3) Chmy indices (index 0 =
// low halo).  Rows are padded to a multiple of one 128-byte line and the base
// is shifted so that Chmy index i = 1 of every row starts a line: a wavefront
// reading i = 1..64 of an FP64 field touches exactly four lines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "fastice_hip.h"

namespace fi {

// location bit d set = Vertex in dimension d
enum Loc { L_CCC = 0, L_VCC = 1, L_CVC = 2, L_VVC = 3, L_CCV = 4, L_VCV = 5, L_CVV = 6, L_VVV = 7 };
enum { OM_CCC = 0, OM_VVV, OM_VCC, OM_CVC, OM_CCV, OM_VVC, OM_VCV, OM_CVV };
enum { OM2_CC = 0, OM2_VV, OM2_VC, OM2_CV };

struct Geo {
    int nd;
    int n[3];
    double dx[3];
    double idx[3];
    long sy, sz;  // element strides of the device layout
    long off;     // elements from the allocation base to Chmy index (0,0,0)
    size_t elems; // elements per allocation
};

// Arguments of the PT sweep / residual kernels (passed by value).
template <typename T>
struct StokesArgs {
    T *P, *txx, *tyy, *tzz, *txy, *txz, *tyz;
    T *Vx, *Vy, *Vz;
    const T* eta;
    T* eta_next;
    const T *gx, *gy, *gz;
    const T* rh;
    T *rP, *rVx, *rVy, *rVz;
    const T* na[8];
    const T* ns[8];
    int nx, ny, nz;
    long sy, sz;
    T ix, iy, iz;
    T dPr, dtau[6], dV[3];
    T grav[3];
    int grav_field;
    int rh_kind, rh_has_reg, rh_relax;
    T rh_eta, rh_A, rh_n, rh_inv_eta_reg, rh_rate;
};

struct BcOp {
    double* ptr;
    int L;    // location bits
    int kind; // 0 Dirichlet, 1 Neumann
    double v;
};
struct BcBatch {
    BcOp op[8];
    int nop;
    int d;
};

// kernel launchers (stokes_kernels.hip)
template <typename T>
void launch_sigma(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s);
template <typename T>
void launch_velocity(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s);
template <typename T>
void launch_residuals(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s);
template <typename T>
void launch_tau(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s);

// boundary conditions, norms, fills (bc_norm_kernels.hip)
void launch_bc_dim(const Geo& g, const BcBatch& b, hipStream_t s);
// Inf-norm of up to 4 fields over their interiors; out = 4 uint64 (|x| bit patterns, NaN largest)
void launch_norms(const Geo& g, const double* const f[4], const int L[4], int nf, unsigned long long* out,
                  hipStream_t s);
void launch_fill(double* p, size_t n, double v, hipStream_t s);

// level sets (levelset_kernels.hip)
void launch_omega_from_psi(const Geo& g, const double* psi, double* const om[8], hipStream_t s);
struct DemArgs {
    const double* dem; // device, parent layout (dnx+3)*(dny+3)
    int nv[2];
    long sy;
    double o[2], d[2], ext[2];
    double R[9];
    int has_R;
};
void launch_levelset_from_dem(const Geo& g, const double origin[3], double* psi, const DemArgs& dem, hipStream_t s);
void launch_negate_vertex(const Geo& g, double* psi, hipStream_t s);

} // namespace fi
