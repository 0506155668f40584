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

// Elements from a double-buffered field's first buffer to its second, both
// in one allocation (a multiple of 64, so both buffers keep the alignment of
// their rows)
inline size_t pair_stride(const Geo& g) { return (g.elems + 63) / 64 * 64; }

// The FP64 fused sweep addresses fields with 32-bit byte offsets (the
// immersed-boundary sweep both buffers of a pair from the first)
inline bool fused_fits(const Geo& g) { return g.nd == 3 && 2 * pair_stride(g) * sizeof(double) < (1ull << 32); }

// Volume fractions as the sweeps read them: FP64 sweeps read the FP64 masks;
// FP32 sweeps (the corrections of the refined solver, whose FP64 residuals
// use the FP64 masks) read them in IEEE half precision, scaled by 2^15 so
// that every fraction down to 2^-29 is a normal half: 0 and 1 exactly, every
// fraction the reference can call non-empty (>= 1e-6) to within 2^-11 of
// itself.  The PT step at a cut location is the velocity residual divided by
// the ns fraction there, and near the glacier bed small fractions meet PT
// steps that are close to the stability limit: an 8-bit code (q/255, or an
// 8-bit float) perturbs such weights by several per cent, enough to make the
// FP32 correction sweeps unstable at those cells on 383x383x191 and larger
// glaciers although the FP64 iteration is stable there.
enum : uint32_t { H16_SCALE_EXP = 15 };
__device__ __forceinline__ float h16_decode(uint32_t h)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h) * (1.0f / (float)(1u << H16_SCALE_EXP));
}
struct H16Mask {
    const uint16_t* p;
    int s; // halves from one element to the next: 1, or 8 in the packed words below
    __device__ __forceinline__ float operator[](long i) const { return h16_decode(p[i * s]); }
};
// 3D FP32 sweeps with an immersed boundary keep the seven half-precision
// masks they read and the validity byte of a cell in one 16-byte word (one
// load per cell and plane): halves 0-3 na at CCC VVC VCV CVV, halves 4-6 ns
// at VCC CVC CCV, byte 14 the validity bits, byte 15 zero
enum { MW_NA_CCC = 0, MW_NA_VVC, MW_NA_VCV, MW_NA_CVV, MW_NS_VCC, MW_NS_CVC, MW_NS_CCV };
__device__ __forceinline__ float mw_mask(uint4 w, int half)
{
    const uint32_t d = half < 2 ? w.x : half < 4 ? w.y : half < 6 ? w.z : w.w;
    return h16_decode((d >> (16 * (half & 1))) & 0xffffu);
}
__device__ __forceinline__ unsigned mw_flags(uint4 w) { return (w.w >> 16) & 0xffu; }
template <typename T>
struct MaskOf {
    using ptr = const T*;
};
template <>
struct MaskOf<float> {
    using ptr = H16Mask;
};

// Arguments of the PT sweep / residual kernels (passed by value).  The
// stress and velocity state is double buffered: kernels read the previous
// iteration (P, txx, ..., Vx, ...) and write this iteration (Pn, ..., Vxn,
// ...).  Equal pointers give the in-place update.
template <typename T>
struct StokesArgs {
    const T *P, *txx, *tyy, *tzz, *txy, *txz, *tyz;
    T *Pn, *txxn, *tyyn, *tzzn, *txyn, *txzn, *tyzn;
    const T *Vx, *Vy, *Vz;
    T *Vxn, *Vyn, *Vzn;
    const T* eta;
    T* eta_next;
    const T *gx, *gy, *gz;
    const T* rh;
    T *rP, *rVx, *rVy, *rVz;
    typename MaskOf<T>::ptr na[8];
    typename MaskOf<T>::ptr ns[8];
    const uint8_t* flags; // immersed-boundary validity bits (F_* below)
    const uint4* mw;      // 3D FP32 immersed boundary: the packed mask words (MW_*), else null
    // 3D immersed boundary, per column i + sy j: the planes k holding a valid
    // location (crange) and the planes whose viscosity a valid location may
    // read (erange); empty columns have x > y
    const int2* crange;
    const int2* erange;
    // FP64 immersed-boundary sweeps: one entry per tile of the launch, the
    // z range each tile marched (see k_eta_rest); null where not used
    int2* tmarch;
    // 1: the current velocity is +0 at every invalid location of the fused
    // sweep's footprint, so the sweep may take it as 0 without reading it
    int z_read;
    // 1: the output stress / velocity buffers already hold +0 at every
    // invalid location of the fused interior, so the sweep need not store it
    int z_wsig, z_wvel;
    const T* fP;          // forcing of the continuity equation (defect correction), or null
    int frozen_eta;       // 1: eta is fixed, eta_next is not written
    // 3D immersed boundary: 1 when the outputs of compute_residuals! /
    // compute_tau! (and the stress lag) already hold 0 at every position
    // outside their column's crange; those kernels then skip such positions
    // (every location there is invalid, so they would store 0 again)
    int skip_empty;
    int nx, ny, nz;
    long sy, sz;
    T ix, iy, iz;
    T dPr, dtau[6], dV[3];
    T grav[3];
    int grav_field;
    int rh_kind, rh_has_reg, rh_relax;
    T rh_eta, rh_A, rh_n, rh_inv_eta_reg, rh_rate;
};

// validity bits: sigma locations use isvalid(na) && !isempty(ns), velocity
// locations isvalid(ns) && !isempty(na) (IB kernels_3d.jl:6,26,32,38,50,56,62)
enum { F_CCC = 1, F_VVC = 2, F_VCV = 4, F_CVV = 8, F_VCC = 16, F_CVC = 32, F_CCV = 64 };
// 3D: every FP64 volume fraction (both phases, all eight locations) of the
// cell and of its six face neighbours is exactly 1.  The masked stencils of
// compute_tau! and compute_residuals! then equal the plain ones bit for bit
// (1 * x is x), so those kernels skip the mask loads there.  Not a validity
// bit: fi_ib_validity clears it.
enum { F_FULL = 128 };

// inclusive index boxes over Chmy indices
struct Box {
    int i0, i1, j0, j1, k0, k1;
};
struct BoxList {
    Box b[6];
    int nb;
    int start[7]; // block prefix sums
};

// interior region of the fused sweep
struct FusedGeom {
    int lo[3], hi[3];
    int kch;   // planes per workgroup
    int order; // workgroup-to-tile order (0: x fastest, 1: contiguous runs per XCD)
    int xa;    // first column of the line-aligned x tiles (see fused_xrange)
    int nx0;   // 1: a leading tile covers [lo, xa-1]
    int nbx;   // tiles in x
};

// one instance of a packed (batched) sweep: its kernel arguments at one
// state parity and its fused interior region
template <typename T>
struct PackedInst {
    StokesArgs<T> a;
    FusedGeom fg;
};

// a box of pointwise work of one instance; start = its first workgroup
struct PackedBox {
    int inst;
    int i0, i1, j0, j1, k0, k1;
    int start;
};

struct BcOp {
    void* ptr;
    int L;    // location bits
    int kind; // bit 0: Neumann; bits 1-2: side restriction (0 both, 1 low, 2 high)
    int f32;  // field stored as float
    double v;
};
struct BcBatch {
    BcOp op[8];
    int nop;
    int d;
};

// kernel launchers (stokes_kernels.hip)
// pointwise sweeps over a box list (3D) or the full work range (2D)
template <typename T>
void launch_sigma_boxes(const Geo& g, const StokesArgs<T>& a, const BoxList& bl, bool ib, hipStream_t s);
template <typename T>
void launch_velocity_boxes(const Geo& g, const StokesArgs<T>& a, const BoxList& bl, bool ib, hipStream_t s);
// fused stress + velocity sweep over the interior region (3D)
// (shell_sigma: also the pointwise stress update of those shell boxes, in the same launch)
template <typename T>
void launch_fused(const Geo& g, const StokesArgs<T>& a, const FusedGeom& fg, bool ib, int tx, int ty, int minw,
                  hipStream_t s, const BoxList* shell_sigma = nullptr);
// compiled (tile lanes in x, tile rows, min waves per SIMD) variants of the fused sweep
int fused_configs(bool fp32, int cfg[][3], int max);
// x tiling of a fused region for tiles of tx lanes (sets xa, nx0, nbx)
void fused_xtiles(FusedGeom& fg, int tx, int esize, long off);
// z planes per workgroup of the fused sweep when none is set (kch <= 0)
int fused_auto_kch(const FusedGeom& fg, int tiles_per_plane);
// interior region and the shell boxes around it; false when the interior is empty
bool fused_split(const Geo& g, int ty, int kch, FusedGeom& fg, BoxList& shell);
BoxList full_box(const Geo& g);
void launch_ib_flags(const Geo& g, const double* const na[8], const double* const ns[8], uint8_t* flags,
                     hipStream_t s);
// 3D: the per-column plane ranges of the flags (StokesArgs::crange / erange),
// arrays of sy * (n[1] + 3) entries; own and eta must not overlap
void launch_col_ranges(const Geo& g, const uint8_t* flags, int2* own, int2* eta, hipStream_t s);
inline size_t col_ranges_count(const Geo& g) { return (size_t)g.sy * (size_t)(g.n[1] + 3); }
// tiles of any fused launch over g (64- or 66-lane tiles, at least 2 rows
// and 1 plane per tile): the capacity of StokesArgs::tmarch (whose
// allocation holds three more 8-byte slots, for k_eta_target's scalars)
inline size_t tile_march_count(const Geo& g)
{
    return (size_t)((g.n[0] + 61) / 62 + 2) * (size_t)((g.n[1] + 1) / 2 + 1) * (size_t)(g.n[2] + 1);
}
template <typename T>
void launch_residuals(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s);
// FP32 outputs of the stress lag (old - new) of a compute_tau (xx yy zz xy xz yz)
struct TauLag {
    float* t[6];
};
template <typename T>
void launch_tau(const Geo& g, const StokesArgs<T>& a, bool ib, hipStream_t s, const TauLag* lag = nullptr);
template <typename T>
void launch_eta_update(const Geo& g, const StokesArgs<T>& a, hipStream_t s);

// boundary conditions, norms, fills (bc_norm_kernels.hip)
void launch_bc_dim(const Geo& g, const BcBatch& b, hipStream_t s);
// the batches of every dimension (b[0..nd-1]) in one launch, same result;
// skip_ranges (3D, a stress group): the per-column plane ranges of the
// validity bits (StokesArgs::crange); the ghosts whose point and chain
// start lie outside them already hold their value and are skipped; with
// them, the faces in empty_faces (face_layers_empty) are left out of the
// launch, every ghost of theirs being one of those
void launch_bc_dims(const Geo& g, const BcBatch* b, hipStream_t s, const int2* skip_ranges = nullptr,
                    unsigned empty_faces = 0);
// the k_bc_all launch of a group: its fields, with per pass and side the
// face rule, and one slice of workgroups per written face
struct BcField {
    void* ptr;
    int L, f32;
    int kind[3][2]; // -1: pass d does not write side s; else 0 Dirichlet, 1 Neumann
    double v[3][2];
};
// one face of one field: pass d, side s, ext1 x ext2 points, workgroups from start
struct BcSlice {
    int fi, d, s;
    int ext1, ext2;
    int start;
};
struct BcAll {
    BcField f[8];
    int nf, nd;
    const int2* skip_ranges; // see launch_bc_dims; null: nothing is skipped
    BcSlice sl[48];
    int nsl;
};
bool bc_all_build(const Geo& g, const BcBatch* b, const int2* skip_ranges, unsigned empty_faces, BcAll& a,
                  int& nblocks);
// Faces of a 3D grid whose layer holds no valid location (bit 2d+s): the
// planes 0..2 (s = 0) or n..n+2 (s = 1) of dimension d, from the column
// ranges (host copy, indexed i + sy * j).  The shell box and every boundary
// condition point and source of such a face lie in its layer.
unsigned face_layers_empty(const Geo& g, const int2* host_ranges);
// bl without the boxes that lie in the layer of a face of empty_faces
BoxList drop_empty_boxes(const Geo& g, const BoxList& bl, unsigned empty_faces);
// one model's k_bc_all launch inside a packed launch over many models
struct PackedBcAll {
    BcAll a;
    int n[3];
    long sy, sz;
    int start; // its first workgroup in the packed launch
};
void launch_bc_all_packed(const PackedBcAll* all, const int* blk2inst, int nblocks, hipStream_t s);

// packed sweeps over many instances (batch.hip builds the tables)
// visc false: every listed instance sweeps with frozen viscosity (FP32 only)
template <typename T>
void launch_fused_packed(const PackedInst<T>* inst, const int4* items, int nitems, bool ib, bool visc, int tx, int ty,
                         int minw, int grid, hipStream_t s);
template <typename T>
void launch_point_packed(const PackedInst<T>* inst, const PackedBox* boxes, const int* blk2box, int nblocks, bool ib,
                         int nd, bool sigma, hipStream_t s);

// one boundary-condition face operation of one instance in one dimension,
// with the geometry it needs; start = its first workgroup
struct PackedBcOp {
    BcOp op;
    long st[3];  // element strides of x, y, z
    int d;       // normal dimension
    int nn;      // cells in the normal dimension
    int e1, e2;  // tangential dimensions (e2 = -1 in 2D)
    int ext1, ext2;
    int start;
};
void launch_bc_packed(const PackedBcOp* ops, const int* blk2op, int nblocks, hipStream_t s);
// Inf-norm of up to 4 fields over their interiors; out = 4 uint64 (|x| bit patterns, NaN largest);
// crange (3D, or null): per-column plane ranges outside which every field is known to be 0
void launch_norms(const Geo& g, const double* const f[4], const int L[4], int nf, unsigned long long* out,
                  hipStream_t s, const int2* crange = nullptr);
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
void launch_levelset_from_polygon(const Geo& g, const double origin[3], double* psi, const double* xy, int npts,
                                  hipStream_t s);
void launch_negate_vertex(const Geo& g, double* psi, hipStream_t s);

} // namespace fi
