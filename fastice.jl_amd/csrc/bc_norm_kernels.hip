// bc_norm_kernels.hip -- boundary-condition face kernels, Inf-norm reductions
// and fills.
//
// Boundary conditions follow Chmy's bc! as this library fixes it: one launch
// per dimension (x, then y, then z), covering both sides and every field of
// the batch, over the field's full padded extent in the other dimensions
// (so edge and corner ghosts end up as functions of interior values only).
//   Dirichlet(v), Vertex in the normal dim: boundary node = v
//   Dirichlet(v), Center in the normal dim: ghost = 2v - first interior
//   Neumann(),   Center: ghost = first interior
//   Neumann(),   Vertex: ghost = mirror of the node next to the boundary node
#include "fi_internal.h"

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
