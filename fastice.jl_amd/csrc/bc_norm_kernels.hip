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

// Points as three scalars selected by dimension (no run-time indexed arrays,
// which would live in scratch memory).
struct BcPt {
    int x, y, z;
    __device__ __forceinline__ int at(int e) const { return e == 0 ? x : e == 1 ? y : z; }
    __device__ __forceinline__ void set(int e, int v)
    {
        if (e == 0) x = v;
        else if (e == 1) y = v;
        else z = v;
    }
};

// The chain of the point p of pass d (see above): the face rules met (bit
// 2e: pass e applies a rule, bit 2e+1: its side), the point the value starts
// from, or v at a Dirichlet vertex.
struct BcChain {
    unsigned rules;
    bool from_v;
    double v0;
    long src, tgt;
    BcPt sp; // the point the value starts from (unless from_v)
};

__device__ __forceinline__ int bc_n(int e, int n0, int n1, int n2) { return e == 0 ? n0 : e == 1 ? n1 : n2; }

__device__ __forceinline__ BcChain bc_chain(const BcField& F, int nd, int d, int n0, int n1, int n2, long sy, long sz,
                                            BcPt p)
{
    BcChain c;
    c.rules = 0u;
    c.from_v = false;
    c.v0 = 0.0;
    BcPt q = p;
    for (int e = d; e >= 0 && !c.from_v; --e) {
        const int vert = (F.L >> e) & 1, ne = bc_n(e, n0, n1, n2);
        for (int s = 0; s < 2; ++s) {
            const int k = F.kind[e][s];
            if (k < 0 || q.at(e) != bc_target(k, vert, s, ne)) continue;
            if (k == 0 && vert) {
                c.from_v = true;
                c.v0 = F.v[e][s];
            } else {
                c.rules |= (1u | (unsigned)s << 1) << (2 * e);
                q.set(e, bc_source(vert, s, ne));
            }
            break;
        }
    }
    c.sp = q;
    c.src = q.x + sy * q.y + (nd == 3 ? sz * q.z : 0);
    c.tgt = p.x + sy * p.y + (nd == 3 ? sz * p.z : 0);
    return c;
}

// p lies on a plane of its column (x, y) without a valid location
__device__ __forceinline__ bool bc_outside(const int2* ranges, long sy, const BcPt& p)
{
    const int2 r = ranges[p.x + sy * p.y];
    return p.z < r.x || p.z > r.y;
}

// the points of one face a thread handles: they are set up and their
// sources loaded first, then their values stored (no thread of the launch
// writes a source, so the loads may all go ahead of the stores); a face
// holds fewer than 2^32 points (32-bit index arithmetic)
constexpr int BC_PTS = 4;

template <typename T>
__device__ __forceinline__ void bc_points(const BcAll& b, const BcSlice& S, const BcField& F, unsigned w0, int n0, int n1,
                                          int n2, long sy, long sz)
{
    const int d = S.d, s = S.s, k = F.kind[d][s];
    const int e1 = d == 0 ? 1 : 0, e2 = b.nd == 3 ? (d == 2 ? 1 : 2) : -1;
    T* f = static_cast<T*>(F.ptr);
    const unsigned np = (unsigned)S.ext1 * (unsigned)S.ext2, ext1 = (unsigned)S.ext1;
    BcChain c[BC_PTS];
    bool live[BC_PTS];
    T x[BC_PTS];
#pragma unroll
    for (int u = 0; u < BC_PTS; ++u) {
        const unsigned w = w0 + (unsigned)u * blockDim.x;
        live[u] = w < np;
        x[u] = T(0);
        c[u].rules = 0u;
        c[u].tgt = 0;
        if (!live[u]) continue;
        BcPt p{0, 0, 0};
        p.set(d, bc_target(k, (F.L >> d) & 1, s, bc_n(d, n0, n1, n2)));
        p.set(e1, (int)(w % ext1));
        if (e2 >= 0) p.set(e2, (int)(w / ext1));
        // a later pass that writes p owns it
        for (int e = d + 1; e < b.nd; ++e)
            for (int t = 0; t < 2; ++t) {
                const int ke = F.kind[e][t];
                if (ke >= 0 && p.at(e) == bc_target(ke, (F.L >> e) & 1, t, bc_n(e, n0, n1, n2))) live[u] = false;
            }
        if (!live[u]) continue;
        c[u] = bc_chain(F, b.nd, d, n0, n1, n2, sy, sz, p);
        // a ghost whose point and the point its chain starts from lie on
        // planes of their columns without a valid location: the sweep stored
        // nothing at either since the last application, which wrote the
        // value the chain gives a +0 start (a chain from a Dirichlet vertex
        // starts from v and reads nothing)
        if (b.skip_ranges && bc_outside(b.skip_ranges, sy, p) &&
            (c[u].from_v || bc_outside(b.skip_ranges, sy, c[u].sp))) {
            live[u] = false;
            continue;
        }
        x[u] = c[u].from_v ? T(c[u].v0) : f[c[u].src];
    }
#pragma unroll
    for (int u = 0; u < BC_PTS; ++u) {
        if (!live[u]) continue;
        T v = x[u];
        for (int e = 0; e <= d; ++e) {
            const unsigned r = c[u].rules >> (2 * e);
            if ((r & 1u) && F.kind[e][(r >> 1) & 1u] == 0) v = T(2) * T(F.v[e][(r >> 1) & 1u]) - v;
        }
        f[c[u].tgt] = v;
    }
}

__global__ __launch_bounds__(256) void k_bc_all(BcAll b, int n0, int n1, int n2, long sy, long sz)
{
    // the slice of this workgroup (slices in launch order, starts ascending)
    const int bid = blockIdx.x;
    int q = 0;
    while (q + 1 < b.nsl && bid >= b.sl[q + 1].start) ++q;
    const BcSlice S = b.sl[q];
    const BcField& F = b.f[S.fi];
    const unsigned w0 = (unsigned)(bid - S.start) * blockDim.x * BC_PTS + threadIdx.x;
    if (F.f32) bc_points<float>(b, S, F, w0, n0, n1, n2, sy, sz);
    else bc_points<double>(b, S, F, w0, n0, n1, n2, sy, sz);
}

// The k_bc_all launch of one model's group (b: one batch per dimension):
// fills a and its slice starts, nblocks = workgroups.  false: a field has
// two rules on one face or there are more than 8 fields, and the passes
// must run one launch per dimension (launch_bc_dim) to keep their order.
bool bc_all_build(const Geo& g, const BcBatch* b, const int2* skip_ranges, unsigned empty_faces, BcAll& a,
                  int& nblocks)
{
    if (!skip_ranges || g.nd != 3) empty_faces = 0;
    a.nf = 0;
    a.nd = g.nd;
    a.nsl = 0;
    a.skip_ranges = g.nd == 3 ? skip_ranges : nullptr;
    nblocks = 0;
    for (int d = 0; d < g.nd; ++d)
        for (int o = 0; o < b[d].nop; ++o) {
            const BcOp& op = b[d].op[o];
            int fi = 0;
            while (fi < a.nf && a.f[fi].ptr != op.ptr) ++fi;
            if (fi == a.nf) {
                if (a.nf == 8) return false;
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
            if (F.L != op.L || F.f32 != op.f32) return false;
            const int sides = op.kind >> 1;
            for (int t = 0; t < 2; ++t) {
                if ((sides == 1 && t != 0) || (sides == 2 && t != 1)) continue;
                if (F.kind[d][t] >= 0) return false; // two rules on one face: keep the launch order
                F.kind[d][t] = op.kind & 1;
                F.v[d][t] = op.v;
            }
        }
    // one slice per face with a rule, each sized to its field's extents
    for (int fi = 0; fi < a.nf; ++fi)
        for (int d = 0; d < g.nd; ++d) {
            int e1 = -1, e2 = -1;
            for (int e = 0; e < g.nd; ++e) {
                if (e == d) continue;
                if (e1 < 0) e1 = e; else e2 = e;
            }
            const int L = a.f[fi].L;
            const int ext1 = g.n[e1] + ((L >> e1) & 1) + 2;
            const int ext2 = (e2 >= 0) ? g.n[e2] + ((L >> e2) & 1) + 2 : 1;
            for (int t = 0; t < 2; ++t) {
                if (a.f[fi].kind[d][t] < 0) continue;
                // every point and chain start of the face lies in its layer:
                // all of them would be skipped
                if (empty_faces & (1u << (2 * d + t))) continue;
                BcSlice& S = a.sl[a.nsl++];
                S.fi = fi;
                S.d = d;
                S.s = t;
                S.ext1 = ext1;
                S.ext2 = ext2;
                S.start = nblocks;
                nblocks += (int)(((long)ext1 * ext2 + 256 * BC_PTS - 1) / (256 * BC_PTS));
            }
        }
    return true;
}

void launch_bc_dims(const Geo& g, const BcBatch* b, hipStream_t s, const int2* skip_ranges, unsigned empty_faces)
{
    BcAll a;
    int nblocks = 0;
    if (!bc_all_build(g, b, skip_ranges, empty_faces, a, nblocks)) {
        for (int d = 0; d < g.nd; ++d) launch_bc_dim(g, b[d], s);
        return;
    }
    if (a.nsl == 0) return;
    hipLaunchKernelGGL(k_bc_all, dim3(nblocks), dim3(256), 0, s, a, g.n[0], g.n[1], g.n[2], g.sy, g.sz);
}

// k_bc_all of many models in one launch: workgroup w works for model
// blk2inst[w], as workgroup w - start of that model's own launch would
__global__ __launch_bounds__(256) void k_bc_all_packed(const PackedBcAll* __restrict__ all,
                                                       const int* __restrict__ blk2inst)
{
    const PackedBcAll& P = all[blk2inst[blockIdx.x]];
    const BcAll& b = P.a;
    const int bid = (int)blockIdx.x - P.start;
    int q = 0;
    while (q + 1 < b.nsl && bid >= b.sl[q + 1].start) ++q;
    const BcSlice S = b.sl[q];
    const BcField& F = b.f[S.fi];
    const unsigned w0 = (unsigned)(bid - S.start) * blockDim.x * BC_PTS + threadIdx.x;
    if (F.f32) bc_points<float>(b, S, F, w0, P.n[0], P.n[1], P.n[2], P.sy, P.sz);
    else bc_points<double>(b, S, F, w0, P.n[0], P.n[1], P.n[2], P.sy, P.sz);
}

void launch_bc_all_packed(const PackedBcAll* all, const int* blk2inst, int nblocks, hipStream_t s)
{
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_bc_all_packed, dim3(nblocks), dim3(256), 0, s, all, blk2inst);
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
    // 3D immersed boundary, or null: per column, the planes holding a valid
    // location; the fields are known to be 0 at every other position, which
    // cannot raise the maximum, so those are not read
    const int2* crange;
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
        if (a.crange) {
            const int2 rg = a.crange[i + sy * j];
            if (k < rg.x || k > rg.y) continue;
        }
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

// The same maximum over the column ranges: one thread per column (i, j) of
// a field's interior and chunk of NORM_KCH planes (blockIdx.z = chunk * nf +
// field) reads the planes of its chunk inside the column's range only.
constexpr int NORM_KCH = 64;

__global__ __launch_bounds__(256) void k_norms_cols(NormArgs a, long sy, long sz, int nf, unsigned long long* out)
{
    const int q = (int)blockIdx.z % nf, chunk = (int)blockIdx.z / nf;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
    unsigned long long m = 0ull;
    if (i <= a.hi[q][0] && j <= a.hi[q][1]) {
        const int2 rg = a.crange[i + sy * j];
        const int k0 = max(1 + chunk * NORM_KCH, rg.x), k1 = min(min(chunk * NORM_KCH + NORM_KCH, a.hi[q][2]), rg.y);
        const double* f = a.f[q] + i + sy * j;
        for (int k = k0; k <= k1; ++k) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(f[sz * k]) & 0x7fffffffffffffffull;
            m = b > m ? b : m;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long v = __shfl_xor(m, o);
        m = v > m ? v : m;
    }
    __shared__ unsigned long long red[4];
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = t & 63, wv = t >> 6;
    if (lane == 0) red[wv] = m;
    __syncthreads();
    if (t == 0) {
        unsigned long long r = red[0];
        for (int w = 1; w < 4; ++w) r = red[w] > r ? red[w] : r;
        if (r) atomicMax(out + q, r);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

void launch_norms(const Geo& g, const double* const f[4], const int L[4], int nf, unsigned long long* out,
                  hipStream_t s, const int2* crange)
{
    NormArgs a;
    a.crange = g.nd == 3 ? crange : nullptr;
    for (int q = 0; q < 4; ++q) {
        a.f[q] = f[q < nf ? q : 0];
        for (int d = 0; d < 3; ++d) a.hi[q][d] = (d < g.nd) ? g.n[d] + ((L[q < nf ? q : 0] >> d) & 1) : 1;
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, out, 4);
    if (a.crange) {
        int hx = 1, hy = 1, hz = 1;
        for (int q = 0; q < nf; ++q) {
            hx = std::max(hx, a.hi[q][0]);
            hy = std::max(hy, a.hi[q][1]);
            hz = std::max(hz, a.hi[q][2]);
        }
        const dim3 grd((unsigned)((hx + 63) / 64), (unsigned)((hy + 3) / 4), (unsigned)(((hz + NORM_KCH - 1) / NORM_KCH) * nf));
        hipLaunchKernelGGL(k_norms_cols, grd, dim3(64, 4), 0, s, a, g.sy, g.sz, nf, out);
        return;
    }
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
