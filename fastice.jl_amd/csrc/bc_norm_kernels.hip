// bc_norm_kernels.hip -- boundary-condition face kernels, Inf-norm reductions
// and fills.
//
// Boundary conditions follow Chmy's bc! as this library fixes it: one pass
// per dimension (x, then y, then z), covering both sides and every field of
// the batch, over the field's full padded extent in the other dimensions
// (so edge and corner ghosts end up as functions of interior values only).
// k_bc_dim runs one pass per launch; k_bc_all gives the result of all of
// them in one launch.
//   Dirichlet(v), Vertex in the normal dim: boundary node = v
//   Dirichlet(v), Center in the normal dim: ghost = 2v - first interior
//   Neumann(),   Center: ghost = first interior
//   Neumann(),   Vertex: ghost = mirror of the node next to the boundary node
#include "fi_internal.h"

#include <algorithm>

namespace fi {

template <typename T>
__device__ __forceinline__ void apply_face(T* f, int kind, int vert, int s, int nn, long base, long st, double v)
{
    if (kind == 0) {
        if (vert) f[base + (long)((s == 0) ? 1 : nn + 1) * st] = T(v);
        else {
            long t = (s == 0) ? 0 : nn + 1, src = (s == 0) ? 1 : nn;
            f[base + t * st] = T(2) * T(v) - f[base + src * st];
        }
    } else {
        long t, src;
        if (vert) { t = (s == 0) ? 0 : nn + 2; src = (s == 0) ? 2 : nn; }
        else { t = (s == 0) ? 0 : nn + 1; src = (s == 0) ? 1 : nn; }
        f[base + t * st] = f[base + src * st];
    }
}

__global__ __launch_bounds__(256) void k_bc_dim(BcBatch b, int nd, int n0, int n1, int n2, long sy, long sz)
{
    const int s = blockIdx.z; // side
    const int q1 = blockIdx.x * blockDim.x + threadIdx.x;
    const int q2 = blockIdx.y * blockDim.y + threadIdx.y;
    const int n[3] = {n0, n1, n2};
    const long st[3] = {1, sy, sz};
    const int d = b.d;
    int e1 = -1, e2 = -1;
    for (int e = 0; e < nd; ++e) {
        if (e == d) continue;
        if (e1 < 0) e1 = e; else e2 = e;
    }
    for (int o = 0; o < b.nop; ++o) {
        const BcOp op = b.op[o];
        // op.kind: bit 0 = Neumann, bits 1-2 = side restriction (0 both, 1 low, 2 high)
        const int sides = op.kind >> 1;
        if ((sides == 1 && s != 0) || (sides == 2 && s != 1)) continue;
        const int kind = op.kind & 1;
        const int ext1 = n[e1] + ((op.L >> e1) & 1) + 2;
        const int ext2 = (e2 >= 0) ? n[e2] + ((op.L >> e2) & 1) + 2 : 1;
        if (q1 >= ext1 || q2 >= ext2) continue;
        const long base = q1 * st[e1] + ((e2 >= 0) ? q2 * st[e2] : 0);
        const int vert = (op.L >> d) & 1;
        const int nn = n[d];
        if (op.f32) apply_face(static_cast<float*>(op.ptr), kind, vert, s, nn, base, st[d], op.v);
        else apply_face(static_cast<double*>(op.ptr), kind, vert, s, nn, base, st[d], op.v);
    }
}

void launch_bc_dim(const Geo& g, const BcBatch& b, hipStream_t s)
{
    if (b.nop == 0) return;
    int e1 = -1, e2 = -1;
    for (int e = 0; e < g.nd; ++e) {
        if (e == b.d) continue;
        if (e1 < 0) e1 = e; else e2 = e;
    }
    const int m1 = g.n[e1] + 3, m2 = (e2 >= 0) ? g.n[e2] + 3 : 1;
    dim3 blk(64, 4, 1);
    dim3 grd((m1 + 63) / 64, (m2 + 3) / 4, 2);
    hipLaunchKernelGGL(k_bc_dim, grd, blk, 0, s, b, g.nd, g.n[0], g.n[1], g.n[2], g.sy, g.sz);
}

// Every dimension of a boundary-condition group in one launch, with the
// result of the launches x, then y, then z above, bit for bit.
//
// Per field, pass d writes the points whose index in d is a target of one of
// its faces and reads each one's source (an index in d that pass d never
// writes).  The value a point ends with is therefore: the last pass e that
// writes it applies its face rule to the value its source held after passes
// 0..e-1, and so on down.  The thread of a face point of pass d writes the
// point only if no later pass writes it, and evaluates that chain itself:
// it follows the sources from pass d down to pass 0, loads the one value the
// chain starts from (a point that no pass writes, so no thread of this
// launch changes it), or starts from v at a Dirichlet vertex, and applies
// the face rules of the chain in pass order, in the same arithmetic.
struct BcField {
    void* ptr;
    int L, f32;
    int kind[3][2]; // -1: pass d does not write side s; else 0 Dirichlet, 1 Neumann
    double v[3][2];
};
struct BcAll {
    BcField f[8];
    int nf, nd;
    const int2* skip_ranges; // see launch_bc_dims; null: nothing is skipped
};

__device__ __forceinline__ int bc_target(int kind, int vert, int s, int nn)
{
    if (kind == 0 && vert) return s == 0 ? 1 : nn + 1;
    if (kind == 1 && vert) return s == 0 ? 0 : nn + 2;
    return s == 0 ? 0 : nn + 1;
}
__device__ __forceinline__ int bc_source(int vert, int s, int nn)
{
    if (vert) return s == 0 ? 2 : nn; // Neumann at a vertex (a Dirichlet vertex reads nothing)
    return s == 0 ? 1 : nn;
}

template <typename T>
__device__ __forceinline__ void bc_chain(const BcField& F, int nd, int d, const int n[3], const long st[3], int p[3])
{
    // the face rules met along the chain, outermost pass first
    int cd[3], cs[3], nc = 0;
    bool from_v = false;
    double v0 = 0.0;
    int q[3] = {p[0], p[1], p[2]};
    for (int e = d; e >= 0 && !from_v; --e) {
        const int vert = (F.L >> e) & 1;
        for (int s = 0; s < 2; ++s) {
            const int k = F.kind[e][s];
            if (k < 0 || q[e] != bc_target(k, vert, s, n[e])) continue;
            if (k == 0 && vert) {
                from_v = true;
                v0 = F.v[e][s];
            } else {
                cd[nc] = e;
                cs[nc] = s;
                ++nc;
                q[e] = bc_source(vert, s, n[e]);
            }
            break;
        }
    }
    T* f = static_cast<T*>(F.ptr);
    T x;
    if (from_v) x = T(v0);
    else x = f[q[0] * st[0] + q[1] * st[1] + (nd == 3 ? q[2] * st[2] : 0)];
    for (int c = nc - 1; c >= 0; --c)
        if (F.kind[cd[c]][cs[c]] == 0) x = T(2) * T(F.v[cd[c]][cs[c]]) - x;
    f[p[0] * st[0] + p[1] * st[1] + (nd == 3 ? p[2] * st[2] : 0)] = x;
}

__global__ __launch_bounds__(256) void k_bc_all(BcAll b, int n0, int n1, int n2, long sy, long sz)
{
    const int z = blockIdx.z;
    const int fi = z / (2 * b.nd), d = (z >> 1) % b.nd, s = z & 1;
    const BcField& F = b.f[fi];
    const int k = F.kind[d][s];
    if (k < 0) return;
    const int n[3] = {n0, n1, n2};
    const long st[3] = {1, sy, sz};
    int e1 = -1, e2 = -1;
    for (int e = 0; e < b.nd; ++e) {
        if (e == d) continue;
        if (e1 < 0) e1 = e; else e2 = e;
    }
    const int q1 = blockIdx.x * blockDim.x + threadIdx.x;
    const int q2 = blockIdx.y * blockDim.y + threadIdx.y;
    const int ext1 = n[e1] + ((F.L >> e1) & 1) + 2;
    const int ext2 = (e2 >= 0) ? n[e2] + ((F.L >> e2) & 1) + 2 : 1;
    if (q1 >= ext1 || q2 >= ext2) return;
    int p[3] = {0, 0, 0};
    p[d] = bc_target(k, (F.L >> d) & 1, s, n[d]);
    p[e1] = q1;
    if (e2 >= 0) p[e2] = q2;
    // a later pass that writes p owns it
    for (int e = d + 1; e < b.nd; ++e)
        for (int t = 0; t < 2; ++t) {
            const int ke = F.kind[e][t];
            if (ke >= 0 && p[e] == bc_target(ke, (F.L >> e) & 1, t, n[e])) return;
        }
    // an x-face ghost whose point and source lie on planes of their columns
    // without a valid location: the sweep stored nothing at either since the
    // last application, which wrote the value a +0 source gives (a
    // Dirichlet vertex has no source)
    if (b.skip_ranges && d == 0) {
        const int2 rt = b.skip_ranges[p[0] + sy * p[1]];
        bool inv = p[2] < rt.x || p[2] > rt.y;
        if (inv && !(k == 0 && (F.L & 1))) {
            const int2 rs = b.skip_ranges[bc_source(F.L & 1, s, n0) + sy * p[1]];
            inv = p[2] < rs.x || p[2] > rs.y;
        }
        if (inv) return;
    }
    if (F.f32) bc_chain<float>(F, b.nd, d, n, st, p);
    else bc_chain<double>(F, b.nd, d, n, st, p);
}

void launch_bc_dims(const Geo& g, const BcBatch* b, hipStream_t s, const int2* skip_ranges)
{
    BcAll a;
    a.nf = 0;
    a.nd = g.nd;
    a.skip_ranges = g.nd == 3 ? skip_ranges : nullptr;
    bool ok = true;
    for (int d = 0; d < g.nd && ok; ++d)
        for (int o = 0; o < b[d].nop && ok; ++o) {
            const BcOp& op = b[d].op[o];
            int fi = 0;
            while (fi < a.nf && a.f[fi].ptr != op.ptr) ++fi;
            if (fi == a.nf) {
                if (a.nf == 8) { ok = false; break; }
                BcField& F = a.f[a.nf++];
                F.ptr = op.ptr;
                F.L = op.L;
                F.f32 = op.f32;
                for (int e = 0; e < 3; ++e)
                    for (int t = 0; t < 2; ++t) {
                        F.kind[e][t] = -1;
                        F.v[e][t] = 0.0;
                    }
            }
            BcField& F = a.f[fi];
            if (F.L != op.L || F.f32 != op.f32) { ok = false; break; }
            const int sides = op.kind >> 1;
            for (int t = 0; t < 2; ++t) {
                if ((sides == 1 && t != 0) || (sides == 2 && t != 1)) continue;
                if (F.kind[d][t] >= 0) { ok = false; break; } // two rules on one face: keep the launch order
                F.kind[d][t] = op.kind & 1;
                F.v[d][t] = op.v;
            }
        }
    if (!ok) {
        for (int d = 0; d < g.nd; ++d) launch_bc_dim(g, b[d], s);
        return;
    }
    if (a.nf == 0) return;
    int m1 = 1, m2 = 1;
    for (int d = 0; d < g.nd; ++d) {
        int e1 = -1, e2 = -1;
        for (int e = 0; e < g.nd; ++e) {
            if (e == d) continue;
            if (e1 < 0) e1 = e; else e2 = e;
        }
        m1 = std::max(m1, g.n[e1] + 3);
        m2 = std::max(m2, (e2 >= 0) ? g.n[e2] + 3 : 1);
    }
    dim3 blk(64, 4, 1);
    dim3 grd((m1 + 63) / 64, (m2 + 3) / 4, 2 * g.nd * a.nf);
    hipLaunchKernelGGL(k_bc_all, grd, blk, 0, s, a, g.n[0], g.n[1], g.n[2], g.sy, g.sz);
}

// Packed boundary conditions of many instances in one dimension: each op
// covers both sides of its face plane (full padded tangential extent).
__global__ __launch_bounds__(256) void k_bc_packed(const PackedBcOp* __restrict__ ops, const int* __restrict__ blk2op)
{
    const PackedBcOp& o = ops[blk2op[blockIdx.x]];
    const long plane = (long)o.ext1 * o.ext2;
    const long t = (long)(blockIdx.x - o.start) * 256 + threadIdx.x;
    if (t >= 2 * plane) return;
    const int s = (int)(t / plane);
    const long rem = t - s * plane;
    const int sides = o.op.kind >> 1;
    if ((sides == 1 && s != 0) || (sides == 2 && s != 1)) return;
    const long q1 = rem % o.ext1, q2 = rem / o.ext1;
    const long base = q1 * o.st[o.e1] + (o.e2 >= 0 ? q2 * o.st[o.e2] : 0);
    const int vert = (o.op.L >> o.d) & 1;
    const int kind = o.op.kind & 1;
    if (o.op.f32) apply_face(static_cast<float*>(o.op.ptr), kind, vert, s, o.nn, base, o.st[o.d], o.op.v);
    else apply_face(static_cast<double*>(o.op.ptr), kind, vert, s, o.nn, base, o.st[o.d], o.op.v);
}

void launch_bc_packed(const PackedBcOp* ops, const int* blk2op, int nblocks, hipStream_t s)
{
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_bc_packed, dim3(nblocks), dim3(256), 0, s, ops, blk2op);
}

// Inf-norm over the interior.  |x| is compared through its bit pattern:
// for non-negative doubles the unsigned order equals the numeric order and
// every NaN sorts above +Inf, so a NaN anywhere propagates to the result and
// the reduction is exact and order independent.
struct NormArgs {
    const double* f[4];
    int hi[4][3];
};

__global__ __launch_bounds__(256) void k_norms(NormArgs a, long sy, long sz, int nd, unsigned long long* out)
{
    const int q = blockIdx.y;
    const double* f = a.f[q];
    const int hx = a.hi[q][0], hy = a.hi[q][1], hz = a.hi[q][2];
    const long total = (long)hx * hy * hz;
    unsigned long long m = 0ull;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long i = t % hx + 1;
        long r = t / hx;
        long j = r % hy + 1;
        long k = (nd == 3) ? r / hy + 1 : 0;
        unsigned long long b = (unsigned long long)__double_as_longlong(f[i + sy * j + sz * k]) & 0x7fffffffffffffffull;
        m = b > m ? b : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long v = __shfl_xor(m, o);
        m = v > m ? v : m;
    }
    __shared__ unsigned long long red[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = red[w] > r ? red[w] : r;
        atomicMax(out + q, r);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

void launch_norms(const Geo& g, const double* const f[4], const int L[4], int nf, unsigned long long* out,
                  hipStream_t s)
{
    NormArgs a;
    for (int q = 0; q < 4; ++q) {
        a.f[q] = f[q < nf ? q : 0];
        for (int d = 0; d < 3; ++d) a.hi[q][d] = (d < g.nd) ? g.n[d] + ((L[q < nf ? q : 0] >> d) & 1) : 1;
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, out, 4);
    long total = (long)a.hi[0][0] * a.hi[0][1] * a.hi[0][2];
    int blocks = (int)((total + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_norms, dim3(blocks, nf, 1), dim3(256), 0, s, a, g.sy, g.sz, g.nd, out);
}

__global__ __launch_bounds__(256) void k_fill(double* p, size_t n, double v)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}

void launch_fill(double* p, size_t n, double v, hipStream_t s)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

} // namespace fi
