// levelset_kernels.hip -- FP64 level-set and volume-fraction set-up kernels.
//
// omega_from_psi: volume fraction of {psi < 0} in the dual cell around each of
// the eight staggered locations, by splitting the cell into five tetrahedra
// and cutting each with the linear interpolant of psi
// (src/LevelSets/volume_fractions.jl:1-127, volume_fractions_kernels.jl:31-97;
// 2D: two triangles, volume_fractions.jl:8-45, volume_fractions_kernels.jl:2-28).
// levelset_from_dem: signed distance of each grid vertex to a triangulated
// DEM, searched over the DEM cells within 4*max(spacing)
// (src/LevelSets/signed_distances.jl:22-92, compute_levelsets.jl:1-25).
#include "fi_internal.h"

namespace fi {

struct v3 {
    double x, y, z;
};
struct v2 {
    double x, y;
};

__device__ __forceinline__ double perturb(double p) { return fabs(p) > 1e-20 ? p : (p > 0.0 ? 1e-20 : -1e-20); }
__device__ __forceinline__ v3 vsub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// determinant of the 3x3 matrix with columns a, b, c (cofactors of the first row)
__device__ __forceinline__ double det3(v3 a, v3 b, v3 c)
{
    return a.x * (b.y * c.z - c.y * b.z) - b.x * (a.y * c.z - c.y * a.z) + c.x * (a.y * b.z - b.y * a.z);
}

__device__ __forceinline__ double tetvol(v3 a, v3 b, v3 c, v3 d)
{
    return (1.0 / 6.0) * fabs(det3(vsub(b, a), vsub(c, a), vsub(d, a)));
}

__device__ __forceinline__ v3 vij3(const double* p, const v3* v, int i, int j)
{
    double wi = p[i] / (p[i] - p[j]);
    double wj = p[j] / (p[i] - p[j]);
    return v3{v[j].x * wi - v[i].x * wj, v[j].y * wi - v[i].y * wj, v[j].z * wi - v[i].z * wj};
}

__device__ double tetcut(const double* p, const v3* v)
{
    int nneg = (p[0] < 0) + (p[1] < 0) + (p[2] < 0) + (p[3] < 0);
#define VIJ(a, b) vij3(p, v, a, b)
    if (nneg == 0) return 0.0;
    if (nneg == 1) {
        if (p[0] < 0) return tetvol(v[0], VIJ(0, 1), VIJ(0, 2), VIJ(0, 3));
        if (p[1] < 0) return tetvol(v[1], VIJ(1, 0), VIJ(1, 2), VIJ(1, 3));
        if (p[2] < 0) return tetvol(v[2], VIJ(2, 0), VIJ(2, 1), VIJ(2, 3));
        return tetvol(v[3], VIJ(3, 0), VIJ(3, 1), VIJ(3, 2));
    }
    if (nneg == 2) {
        if (p[0] < 0 && p[1] < 0)
            return tetvol(v[0], v[1], VIJ(0, 2), VIJ(1, 3)) + tetvol(VIJ(1, 2), v[1], VIJ(0, 2), VIJ(1, 3)) +
                   tetvol(v[0], VIJ(0, 3), VIJ(0, 2), VIJ(1, 3));
        if (p[0] < 0 && p[2] < 0)
            return tetvol(v[0], v[2], VIJ(0, 3), VIJ(2, 1)) + tetvol(VIJ(2, 3), v[2], VIJ(0, 3), VIJ(2, 1)) +
                   tetvol(v[0], VIJ(0, 1), VIJ(0, 3), VIJ(2, 1));
        if (p[0] < 0 && p[3] < 0)
            return tetvol(v[0], v[3], VIJ(0, 1), VIJ(3, 2)) + tetvol(VIJ(3, 1), v[3], VIJ(0, 1), VIJ(3, 2)) +
                   tetvol(v[0], VIJ(0, 2), VIJ(0, 1), VIJ(3, 2));
        if (p[1] < 0 && p[2] < 0)
            return tetvol(v[2], v[1], VIJ(2, 0), VIJ(1, 3)) + tetvol(VIJ(1, 0), v[1], VIJ(2, 0), VIJ(1, 3)) +
                   tetvol(v[2], VIJ(2, 3), VIJ(2, 0), VIJ(1, 3));
        if (p[1] < 0 && p[3] < 0)
            return tetvol(v[3], v[1], VIJ(3, 0), VIJ(1, 2)) + tetvol(VIJ(1, 0), v[1], VIJ(3, 0), VIJ(1, 2)) +
                   tetvol(v[3], VIJ(3, 2), VIJ(3, 0), VIJ(1, 2));
        return tetvol(v[2], v[3], VIJ(2, 0), VIJ(3, 1)) + tetvol(VIJ(3, 0), v[3], VIJ(2, 0), VIJ(3, 1)) +
               tetvol(v[2], VIJ(2, 1), VIJ(2, 0), VIJ(3, 1));
    }
    if (nneg == 3) {
        double tot = tetvol(v[0], v[1], v[2], v[3]);
        if (p[0] >= 0) return tot - tetvol(v[0], VIJ(0, 1), VIJ(0, 2), VIJ(0, 3));
        if (p[1] >= 0) return tot - tetvol(v[1], VIJ(1, 0), VIJ(1, 2), VIJ(1, 3));
        if (p[2] >= 0) return tot - tetvol(v[2], VIJ(2, 0), VIJ(2, 1), VIJ(2, 3));
        return tot - tetvol(v[3], VIJ(3, 0), VIJ(3, 1), VIJ(3, 2));
    }
    return tetvol(v[0], v[1], v[2], v[3]);
#undef VIJ
}

__device__ double psi2omega3(const double* q, double dx, double dy, double dz)
{
    // q = (000,100,010,110,001,101,011,111) in (x,y,z); volume_fractions.jl:106-115
    const v3 v000{0, 0, 0}, v001{dx, 0, 0}, v010{0, dy, 0}, v011{dx, dy, 0};
    const v3 v100{0, 0, dz}, v101{dx, 0, dz}, v110{0, dy, dz}, v111{dx, dy, dz};
    double p[8];
    for (int t = 0; t < 8; ++t) p[t] = perturb(q[t]);
    double t1[4] = {p[0], p[4], p[2], p[1]};
    v3 w1[4] = {v000, v100, v010, v001};
    double t2[4] = {p[6], p[4], p[2], p[7]};
    v3 w2[4] = {v110, v100, v010, v111};
    double t3[4] = {p[5], p[4], p[7], p[1]};
    v3 w3[4] = {v101, v100, v111, v001};
    double t4[4] = {p[3], p[7], p[2], p[1]};
    v3 w4[4] = {v011, v111, v010, v001};
    double t5[4] = {p[7], p[4], p[2], p[1]};
    v3 w5[4] = {v111, v100, v010, v001};
    double vol = tetcut(t1, w1) + tetcut(t2, w2) + tetcut(t3, w3) + tetcut(t4, w4) + tetcut(t5, w5);
    return vol / (dx * dy * dz);
}

__device__ __forceinline__ double trivol(v2 a, v2 b, v2 c)
{
    double ux = c.x - a.x, uy = c.y - a.y, wx = b.x - a.x, wy = b.y - a.y;
    return 0.5 * fabs(ux * wy - wx * uy);
}

__device__ __forceinline__ v2 vij2(const double* p, const v2* v, int i, int j)
{
    double wi = p[i] / (p[i] - p[j]);
    double wj = p[j] / (p[i] - p[j]);
    return v2{v[j].x * wi - v[i].x * wj, v[j].y * wi - v[i].y * wj};
}

__device__ double tricut(const double* p, const v2* v)
{
    if (p[0] < 0 && p[1] < 0 && p[2] < 0) return trivol(v[0], v[1], v[2]);
    if (p[0] > 0 && p[1] > 0 && p[2] > 0) return 0.0;
    v2 v12 = vij2(p, v, 0, 1), v13 = vij2(p, v, 0, 2), v23 = vij2(p, v, 1, 2);
    if (p[0] < 0) {
        if (p[1] < 0) return trivol(v[0], v23, v13) + trivol(v[0], v[1], v23);
        if (p[2] < 0) return trivol(v[2], v12, v23) + trivol(v[2], v[0], v12);
        return trivol(v[0], v12, v13);
    }
    if (p[1] < 0) {
        if (p[2] < 0) return trivol(v[1], v13, v12) + trivol(v[1], v[2], v13);
        return trivol(v12, v[1], v23);
    }
    return trivol(v13, v23, v[2]);
}

__device__ double psi2omega2(const double* q, double dx, double dy)
{
    const v2 v00{0, 0}, v01{dx, 0}, v10{0, dy}, v11{dx, dy};
    double p00 = perturb(q[0]), p01 = perturb(q[1]), p10 = perturb(q[2]), p11 = perturb(q[3]);
    double a[3] = {p00, p01, p11};
    v2 va[3] = {v00, v01, v11};
    double b[3] = {p00, p11, p10};
    v2 vb[3] = {v00, v11, v10};
    return (tricut(a, va) + tricut(b, vb)) / (dx * dy);
}

// psi averaged to a dual location; mask bit d = average along dim d (x innermost)
__device__ __forceinline__ double psi_lerp(const double* psi, long c, int mask, long sy, long sz)
{
    const int nxp = (mask & 1) ? 2 : 1, nyp = (mask & 2) ? 2 : 1, nzp = (mask & 4) ? 2 : 1;
    double zs[2] = {0.0, 0.0};
    for (int kk = 0; kk < nzp; ++kk) {
        double ys[2] = {0.0, 0.0};
        for (int jj = 0; jj < nyp; ++jj) {
            long b = c + jj * sy + kk * sz;
            ys[jj] = (nxp == 2) ? 0.5 * (psi[b] + psi[b + 1]) : psi[b];
        }
        zs[kk] = (nyp == 2) ? 0.5 * (ys[0] + ys[1]) : ys[0];
    }
    return (nzp == 2) ? 0.5 * (zs[0] + zs[1]) : zs[0];
}

struct OmegaPtrs {
    double* om[8];
};

__global__ __launch_bounds__(256) void k_omega3(const double* psi, OmegaPtrs o, int nx, int ny, long sy, long sz,
                                                double dx, double dy, double dz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > nx + 1 || j > ny + 1) return;
    const long c = (long)i + sy * j + sz * k;
    // component order ccc vvv vcc cvc ccv vvc vcv cvv; lerp masks and corner offsets
    const int lmask[8] = {0, 7, 1, 2, 4, 3, 5, 6};
    const int off[8][3] = {{0, 0, 0}, {-1, -1, -1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 0, -1}, {0, -1, -1}};
    for (int q = 0; q < 8; ++q) {
        double p8[8];
        for (int cz = 0; cz < 2; ++cz)
            for (int cy = 0; cy < 2; ++cy)
                for (int cx = 0; cx < 2; ++cx) {
                    long b = c + (off[q][0] + cx) + sy * (off[q][1] + cy) + sz * (off[q][2] + cz);
                    p8[cx + 2 * cy + 4 * cz] = psi_lerp(psi, b, lmask[q], sy, sz);
                }
        o.om[q][c] = psi2omega3(p8, dx, dy, dz);
    }
}

__global__ __launch_bounds__(256) void k_omega2(const double* psi, OmegaPtrs o, int nx, int ny, long sy, double dx,
                                                double dy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > nx + 1 || j > ny + 1) return;
    const long c = (long)i + sy * j;
    const int lmask[4] = {0, 3, 1, 2};
    const int off[4][2] = {{0, 0}, {-1, -1}, {-1, 0}, {0, -1}};
    for (int q = 0; q < 4; ++q) {
        double p4[4];
        for (int cy = 0; cy < 2; ++cy)
            for (int cx = 0; cx < 2; ++cx) {
                long b = c + (off[q][0] + cx) + sy * (off[q][1] + cy);
                p4[cx + 2 * cy] = psi_lerp(psi, b, lmask[q], sy, 0);
            }
        o.om[q][c] = psi2omega2(p4, dx, dy);
    }
}

void launch_omega_from_psi(const Geo& g, const double* psi, double* const om[8], hipStream_t s)
{
    OmegaPtrs o;
    for (int q = 0; q < 8; ++q) o.om[q] = om[q];
    if (g.nd == 3) {
        dim3 blk(64, 4, 1), grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[2] + 1);
        hipLaunchKernelGGL(k_omega3, grd, blk, 0, s, psi, o, g.n[0], g.n[1], g.sy, g.sz, g.dx[0], g.dx[1], g.dx[2]);
    } else {
        dim3 blk(64, 4, 1), grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, 1);
        hipLaunchKernelGGL(k_omega2, grd, blk, 0, s, psi, o, g.n[0], g.n[1], g.sy, g.dx[0], g.dx[1]);
    }
}

// ---------------------------------------------------------------------------
// signed distance to a DEM
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross3(v3 a, v3 b)
{
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ v3 vscale(v3 a, double s) { return v3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
__device__ __forceinline__ double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

__device__ double ud_triangle(v3 p, v3 a, v3 b, v3 c)
{
    v3 ba = vsub(b, a), pa = vsub(p, a), cb = vsub(c, b), pb = vsub(p, b), ac = vsub(a, c), pc = vsub(p, c);
    v3 nor = cross3(ba, ac);
    double s = sgn(dot3(cross3(ba, nor), pa)) + sgn(dot3(cross3(cb, nor), pb)) + sgn(dot3(cross3(ac, nor), pc));
    if (s < 2.0) {
        v3 e1 = vsub(vscale(ba, clamp01(dot3(ba, pa) / dot3(ba, ba))), pa);
        v3 e2 = vsub(vscale(cb, clamp01(dot3(cb, pb) / dot3(cb, cb))), pb);
        v3 e3 = vsub(vscale(ac, clamp01(dot3(ac, pc) / dot3(ac, ac))), pc);
        double d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(e3, e3);
        double mn = d1 < d2 ? d1 : d2;
        mn = mn < d3 ? mn : d3;
        return sqrt(mn);
    }
    double np = dot3(nor, pa);
    return sqrt(np * np / dot3(nor, nor));
}

__device__ __forceinline__ double sign_triangle(v3 p, v3 a, v3 b, v3 c)
{
    double x = dot3(vsub(p, a), cross3(vsub(b, a), vsub(c, a)));
    return x == 0.0 ? 1.0 : sgn(x);
}

__device__ __forceinline__ v3 sample_dem(const DemArgs& g, int i, int j)
{
    return v3{g.o[0] + (double)(i - 1) * g.d[0], g.o[1] + (double)(j - 1) * g.d[1], g.dem[i + g.sy * j]};
}

__device__ __forceinline__ int closest_vertex(const DemArgs& g, double p, int d)
{
    int I = (int)floor((p - g.o[d]) / g.d[d]) + 1;
    I = I < 1 ? 1 : I;
    I = I > g.nv[d] ? g.nv[d] : I;
    return I;
}

__device__ __forceinline__ double pair_ud(const DemArgs& g, v3 p, int i, int j, double* sg)
{
    v3 a = sample_dem(g, i, j), b = sample_dem(g, i + 1, j), c = sample_dem(g, i, j + 1), e = sample_dem(g, i + 1, j + 1);
    double u1 = ud_triangle(p, a, b, c), u2 = ud_triangle(p, c, b, e);
    if (sg) *sg = sign_triangle(p, a, b, c);
    return u1 < u2 ? u1 : u2;
}

__global__ __launch_bounds__(256) void k_levelset(DemArgs g, double* psi, int nx, int ny, long sy, long sz, double ox,
                                                  double oy, double oz, double dx, double dy, double dz, double cutoff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > nx + 1 || j > ny + 1) return;
    double x = ox + (double)(i - 1) * dx, y = oy + (double)(j - 1) * dy, z = oz + (double)(k - 1) * dz;
    v3 p{x, y, z};
    if (g.has_R) {
        p.x = g.R[0] * x + g.R[1] * y + g.R[2] * z;
        p.y = g.R[3] * x + g.R[4] * y + g.R[5] * z;
        p.z = g.R[6] * x + g.R[7] * y + g.R[8] * z;
    }
    double px = p.x, py = p.y;
    double lo0 = g.o[0], hi0 = g.o[0] + g.ext[0], lo1 = g.o[1], hi1 = g.o[1] + g.ext[1];
    px = px < lo0 ? lo0 : (px > hi0 ? hi0 : px);
    py = py < lo1 ? lo1 : (py > hi1 ? hi1 : py);
    int bl0 = closest_vertex(g, px - cutoff, 0), bl1 = closest_vertex(g, py - cutoff, 1);
    int tr0 = closest_vertex(g, px + cutoff, 0), tr1 = closest_vertex(g, py + cutoff, 1);
    int ic0 = closest_vertex(g, px, 0), ic1 = closest_vertex(g, py, 1);
    double sg;
    double ud = pair_ud(g, p, ic0, ic1, &sg);
    for (int jv = bl1; jv <= tr1; ++jv)
        for (int iv = bl0; iv <= tr0; ++iv) {
            if (iv == ic0 && jv == ic1) continue;
            double u = pair_ud(g, p, iv, jv, nullptr);
            ud = u < ud ? u : ud;
        }
    psi[(long)i + sy * j + sz * k] = ud * sg;
}

void launch_levelset_from_dem(const Geo& g, const double origin[3], double* psi, const DemArgs& dem, hipStream_t s)
{
    double dmax = g.dx[0];
    for (int d = 1; d < 3; ++d) dmax = g.dx[d] > dmax ? g.dx[d] : dmax;
    dim3 blk(64, 4, 1), grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[2] + 1);
    hipLaunchKernelGGL(k_levelset, grd, blk, 0, s, dem, psi, g.n[0], g.n[1], g.sy, g.sz, origin[0], origin[1], origin[2],
                       g.dx[0], g.dx[1], g.dx[2], 4.0 * dmax);
}

// Signed distance of (px, py) to a closed polygon (xy pairs), negative
// inside by the even-odd rule: sd_poly, signed_distances.jl:3-19.
__device__ double sd_poly(double px, double py, const double* __restrict__ xy, int npts)
{
    const double dx0 = px - xy[0], dy0 = py - xy[1];
    double d = dx0 * dx0 + dy0 * dy0;
    double s = 1.0;
    int j = npts - 1;
    for (int i = 0; i < npts; ++i) {
        const double ex = xy[2 * j] - xy[2 * i], ey = xy[2 * j + 1] - xy[2 * i + 1];
        const double wx = px - xy[2 * i], wy = py - xy[2 * i + 1];
        double t = (wx * ex + wy * ey) / (ex * ex + ey * ey);
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double bx = wx - ex * t, by = wy - ey * t;
        const double bb = bx * bx + by * by;
        d = bb < d ? bb : d;
        const bool c1 = py >= xy[2 * i + 1], c2 = py < xy[2 * j + 1], c3 = ex * wy > ey * wx;
        if ((c1 && c2 && c3) || (!c1 && !c2 && !c3)) s = -s;
        j = i;
    }
    return s * sqrt(d);
}

__global__ __launch_bounds__(256) void k_levelset_poly(const double* __restrict__ xy, int npts, double* psi, int nx,
                                                       int ny, long sy, double ox, double oy, double dx, double dy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > nx + 1 || j > ny + 1) return;
    const double x = ox + (double)(i - 1) * dx;
    const double y = oy + (double)(j - 1) * dy;
    psi[(long)i + sy * j] = sd_poly(x, y, xy, npts);
}

void launch_levelset_from_polygon(const Geo& g, const double origin[3], double* psi, const double* xy, int npts,
                                  hipStream_t s)
{
    dim3 blk(64, 4, 1), grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, 1);
    hipLaunchKernelGGL(k_levelset_poly, grd, blk, 0, s, xy, npts, psi, g.n[0], g.n[1], g.sy, origin[0], origin[1],
                       g.dx[0], g.dx[1]);
}

__global__ __launch_bounds__(256) void k_negate(double* psi, int nx, int ny, long sy, long sz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + (sz ? 1 : 0);
    if (i > nx + 1 || j > ny + 1) return;
    long c = (long)i + sy * j + sz * k;
    psi[c] = -psi[c];
}

void launch_negate_vertex(const Geo& g, double* psi, hipStream_t s)
{
    dim3 blk(64, 4, 1), grd((g.n[0] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.nd == 3 ? g.n[2] + 1 : 1);
    hipLaunchKernelGGL(k_negate, grd, blk, 0, s, psi, g.n[0], g.n[1], g.sy, g.nd == 3 ? g.sz : 0);
}

} // namespace fi
