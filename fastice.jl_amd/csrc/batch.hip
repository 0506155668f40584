// batch.hip -- batched PT solves: many FP64 models advanced together by
// packed launches over (instance, tile) work lists, replayed as HIP graphs.
//
// Per iteration of the running set the batch launches
//   fused interior tiles of every 3D instance       (k_fused3_packed, per IB kind)
//   pointwise stress update of shells / 2D grids    (k_point_packed, per kind)
//   stress boundary conditions, x then y then z     (k_bc_packed)
//   pointwise velocity update of shells / 2D grids  (k_point_packed)
//   velocity + eta_next boundary conditions         (k_bc_packed)
// which is the single-model iteration (api.hip: iterate) with every launch
// widened to all instances, so each instance computes exactly what it
// computes alone.  The kernel arguments of an instance depend on its buffer
// parity; the tables hold both parities and the iterations alternate them.
#include "refine.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace fi {

struct BatchInst {
    fi_model_s* m = nullptr;
    bool refined = false; // solved by FP32 sweeps in FP64 defect correction
    RefineJob job;
    int left = 0;         // sweeps of the job's pending chunk still to run
    hipStream_t home = nullptr; // the model's own stream, restored on removal
    int max_iter = 0, ncheck = 1;
    double tol = 0.0;
    int iters = 0;
    int state = FI_BATCH_RUNNING;
    double norms[4] = {};
};

enum { K_3D_PLAIN = 0, K_3D_IB = 1, K_2D_PLAIN = 2, K_2D_IB = 3 };

// device tables of the running set for one tile configuration
struct BatchTables {
    char* dev = nullptr;
    size_t cap = 0;
    void* inst[2] = {}; // PackedInst<double> (FP64 batches) or PackedInst<float> (refined batches)
    int4* items[4] = {}; // plain, IB; then the same for frozen-viscosity (FP32 correction) instances
    int nitems[4] = {};
    PackedBox* box[4] = {};
    int* blk2box[4] = {};
    int nblk_box[4] = {};
    PackedBcOp* bc[2][2][3] = {}; // parity, group (0 stress, 1 velocity), dim
    int* blk2bc[2][2][3] = {};
    int nblk_bc[2][2][3] = {};
    // the same passes as one k_bc_all launch per parity and group, where
    // every instance's group has the form it takes (bc_all_build)
    bool bc_all_ok[2][2] = {};
    PackedBcAll* bc_all[2][2] = {};
    int* blk2bcall[2][2] = {};
    int nblk_bcall[2][2] = {};
    unsigned gen = 0; // bumped by every build: graphs record the generation they were captured against
};

} // namespace fi

using namespace fi;

struct fi_batch_s {
    hipStream_t stream = nullptr;
    std::vector<BatchInst> inst;
    std::vector<int> run; // indices of running instances in table order
    bool dirty = true;
    int phase = 0; // iterations since the tables were built, mod 2
    BatchTables t;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // what each graph was captured against: the table generation and the
    // running models (their device buffers are baked into the graph's launches)
    unsigned graph_gen[2] = {0, 0};
    std::vector<fi_model_s*> graph_models;
    int tx = 64, ty = 8, minw = 1, kch = 32; // fused variant and z chunk
    int refined = -1; // -1: no model yet, 0: FP64 solves, 1: refined solves
    int order = 2;                  // workgroup-to-tile order of the packed fused sweep
    int autotune = 1;
    int graph_iters = 8;
    long tuned_cells = 0; // size of the working set the tile choice was made for
    int tune_runs = 0;
    int rebuilds = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // refined batches: the FP64 outer steps of the models due after a chunk
    // are independent of one another and run spread over these streams
    static constexpr int NAUX = 4;
    hipStream_t aux[NAUX] = {};
    hipEvent_t ev_fork = nullptr;
    std::vector<long> sig; // refined batches: what the current tables were built from
    std::vector<unsigned> zbuilt; // FP64 batches: the running models' clean-buffer bits the tables hold
};

namespace fi {

static void destroy_graphs(fi_batch_s* b)
{
    // a replay still in flight must end before its executable goes
    if (b->graph[0] || b->graph[1]) (void)hipStreamSynchronize(b->stream);
    for (int p = 0; p < 2; ++p)
        if (b->graph[p]) {
            (void)hipGraphExecDestroy(b->graph[p]);
            b->graph[p] = nullptr;
        }
}

void batch_touch(fi_model_s* m)
{
    if (m->batch) m->batch->dirty = true;
}

static int find_inst(fi_batch_s* b, fi_model_s* m)
{
    for (size_t q = 0; q < b->inst.size(); ++q)
        if (b->inst[q].m == m) return (int)q;
    return -1;
}

static int detach(fi_batch_s* b, int q)
{
    HIPCHK(hipStreamSynchronize(b->stream));
    fi_model_s* m = b->inst[q].m;
    m->stream = b->inst[q].home;
    m->batch = nullptr;
    b->inst.erase(b->inst.begin() + q);
    b->dirty = true;
    return FI_OK;
}

void batch_detach(fi_model_s* m)
{
    fi_batch_s* b = m->batch;
    int q = find_inst(b, m);
    if (q >= 0) (void)detach(b, q);
}


static bool fused_ok(const fi_model_s* m)
{
    return fused_fits(m->g);
}

static int kind_of(const fi_model_s* m)
{
    const bool ib = m->desc.immersed_boundary != 0;
    return m->g.nd == 3 ? (ib ? K_3D_IB : K_3D_PLAIN) : (ib ? K_2D_IB : K_2D_PLAIN);
}

template <typename X>
static size_t put(std::vector<char>& h, const std::vector<X>& v)
{
    size_t off = (h.size() + 255) / 256 * 256;
    h.resize(off + v.size() * sizeof(X));
    if (!v.empty()) std::memcpy(h.data() + off, v.data(), v.size() * sizeof(X));
    return off;
}

static void add_bc_ops(std::vector<PackedBcOp>& ops, std::vector<int>& blk, const fi_model_s* m, const BcBatch& bb)
{
    const Geo& g = m->g;
    int e1 = -1, e2 = -1;
    for (int e = 0; e < g.nd; ++e) {
        if (e == bb.d) continue;
        if (e1 < 0) e1 = e;
        else e2 = e;
    }
    for (int o = 0; o < bb.nop; ++o) {
        PackedBcOp p;
        std::memset(&p, 0, sizeof(p));
        p.op = bb.op[o];
        p.st[0] = 1;
        p.st[1] = g.sy;
        p.st[2] = g.sz;
        p.d = bb.d;
        p.nn = g.n[bb.d];
        p.e1 = e1;
        p.e2 = e2;
        p.ext1 = g.n[e1] + ((p.op.L >> e1) & 1) + 2;
        p.ext2 = (e2 >= 0) ? g.n[e2] + ((p.op.L >> e2) & 1) + 2 : 1;
        p.start = (int)blk.size();
        const long work = 2L * p.ext1 * p.ext2;
        const int nb = (int)((work + 255) / 256);
        const int idx = (int)ops.size();
        ops.push_back(p);
        for (int k = 0; k < nb; ++k) blk.push_back(idx);
    }
}

// one instance's group as a k_bc_all launch inside the packed one; ok turns
// false when the group needs its passes launched one dimension at a time
static void add_bc_all(std::vector<PackedBcAll>& all, std::vector<int>& blk, bool& ok, const fi_model_s* m,
                       const BcBatch* bb)
{
    PackedBcAll p;
    std::memset(&p, 0, sizeof(p));
    int nb = 0;
    if (!bc_all_build(m->g, bb, nullptr, 0u, p.a, nb)) {
        ok = false;
        return;
    }
    if (p.a.nsl == 0) return;
    for (int d = 0; d < 3; ++d) p.n[d] = m->g.n[d];
    p.sy = m->g.sy;
    p.sz = m->g.sz;
    p.start = (int)blk.size();
    const int idx = (int)all.size();
    all.push_back(p);
    for (int k = 0; k < nb; ++k) blk.push_back(idx);
}

// the FP64 sweep of a model (its PT iteration) and the FP32 sweep of a
// refined model (phase 1 or correction), as the table builder sees them
template <typename T>
struct SweepOf;

template <>
struct SweepOf<double> {
    using Parity = RefineParity;
    static Parity get(fi_model_s* m) { return Parity{m->sig_cur, m->v_cur, m->eta_cur}; }
    static void set(fi_model_s* m, Parity p)
    {
        m->sig_cur = p.sig;
        m->v_cur = p.v;
        m->eta_cur = p.eta;
    }
    static Parity flipped(fi_model_s* m, Parity p, int n)
    {
        (void)m;
        return Parity{p.sig ^ (n & 1), p.v ^ (n & 1), p.eta ^ (n & 1)};
    }
    static StokesArgs<double> args(fi_model_s* m) { return make_args(m, false); }
    static void zflags(fi_model_s* m, Parity p, StokesArgs<double>& a)
    {
        a.z_read = m->zv[p.v];
        a.z_wsig = m->zs[1 - p.sig];
        a.z_wvel = m->zv[1 - p.v];
    }
    static unsigned zbits(fi_model_s* m) { return m->zs[0] | m->zs[1] << 1 | m->zv[0] << 2 | m->zv[1] << 3; }
    static void swept(fi_model_s* m, Parity p0, int n)
    {
        if (n >= 1) m->zs[1 - p0.sig] = m->zv[1 - p0.v] = true;
        if (n >= 2) m->zs[p0.sig] = m->zv[p0.v] = true;
        m->gs[0] = m->gs[1] = false;
    }
    static BcBatch bc(fi_model_s* m, int group, int d)
    {
        return group == 0 ? stress_batch(m, d) : velocity_batch(m, d, true);
    }
    static bool fused(fi_model_s* m, int ty, int kch, FusedGeom& fg, BoxList& shell)
    {
        return fused_ok(m) && fused_split(m->g, ty, kch, fg, shell);
    }
};

template <>
struct SweepOf<float> {
    using Parity = RefineParity;
    static Parity get(fi_model_s* m) { return refine_parity(m); }
    static void set(fi_model_s* m, Parity p) { refine_set_parity(m, p); }
    static Parity flipped(fi_model_s* m, Parity p, int n)
    {
        return Parity{p.sig ^ (n & 1), p.v ^ (n & 1), p.eta ^ (refine_full(m) ? (n & 1) : 0)};
    }
    static StokesArgs<float> args(fi_model_s* m) { return refine_args(m, false); }
    static void zflags(fi_model_s* m, Parity p, StokesArgs<float>& a)
    {
        a.z_read = m->zv32[p.v];
        a.z_wsig = m->zs32[1 - p.sig];
        a.z_wvel = m->zv32[1 - p.v];
    }
    static unsigned zbits(fi_model_s* m) { return m->zs32[0] | m->zs32[1] << 1 | m->zv32[0] << 2 | m->zv32[1] << 3; }
    static void swept(fi_model_s* m, Parity p0, int n)
    {
        if (n >= 1) m->zs32[1 - p0.sig] = m->zv32[1 - p0.v] = true;
        if (n >= 2) m->zs32[p0.sig] = m->zv32[p0.v] = true;
        m->gs32[0] = m->gs32[1] = false;
    }
    static BcBatch bc(fi_model_s* m, int group, int d)
    {
        return refine_bc_batch(m, group == 0 ? FI_BCGROUP_STRESS : FI_BCGROUP_VELOCITY, d);
    }
    static bool fused(fi_model_s* m, int ty, int kch, FusedGeom& fg, BoxList& shell)
    {
        (void)ty;
        const bool f = refine_fused(m, fg, shell) && m->g.elems * sizeof(float) < (1ull << 32);
        fg.kch = kch;
        return f;
    }
};

// build the tables of the running set for tile rows ty and z chunk kch
template <typename T>
static int build_tables_t(fi_batch_s* b, int ty, int kch)
{
    using S = SweepOf<T>;
    const int nrun = (int)b->run.size();
    std::vector<PackedInst<T>> pinst[2];
    std::vector<int4> items[4];
    std::vector<PackedBox> boxes[4];
    std::vector<int> blk2box[4];
    std::vector<PackedBcOp> bcops[2][2][3];
    std::vector<int> bcblk[2][2][3];
    std::vector<PackedBcAll> bcall[2][2];
    std::vector<int> bcallblk[2][2];
    bool bcall_ok[2][2] = {{true, true}, {true, true}};
    for (int r = 0; r < nrun; ++r) {
        fi_model_s* m = b->inst[b->run[r]].m;
        const int kind = kind_of(m);
        FusedGeom fg;
        BoxList shell;
        bool interior = S::fused(m, ty, kch, fg, shell);
        if (!interior) {
            shell = full_box(m->g);
            std::memset(&fg, 0, sizeof(fg));
        }
        fg.order = b->order;
        if (interior) fused_xtiles(fg, b->tx, (int)sizeof(T), m->g.off);
        const typename S::Parity p0 = S::get(m);
        for (int p = 0; p < 2; ++p) {
            typename S::Parity pp = S::flipped(m, p0, p);
            S::set(m, pp);
            PackedInst<T> pi;
            pi.a = S::args(m);
            pi.fg = fg;
            // which buffers of this parity hold +0 at every invalid location,
            // as the model tracks it when the tables are built (the packed
            // sweeps keep a clean buffer clean, so the flags stay true while
            // the tables live; a change of them rebuilds the tables); without
            // a march table the packed sweep updates the viscosity of the
            // planes it does not march itself
            pi.a.tmarch = nullptr;
            S::zflags(m, pp, pi.a);
            pinst[p].push_back(pi);
            // stress BCs see the new stress buffers, velocity BCs the new velocity
            for (int gr = 0; gr < 2; ++gr) {
                if (gr == 0) pp.sig ^= 1;
                else pp.v ^= 1;
                S::set(m, pp);
                BcBatch bb[3];
                for (int d = 0; d < m->g.nd; ++d) {
                    bb[d] = S::bc(m, gr, d);
                    add_bc_ops(bcops[p][gr][d], bcblk[p][gr][d], m, bb[d]);
                }
                if (bcall_ok[p][gr]) add_bc_all(bcall[p][gr], bcallblk[p][gr], bcall_ok[p][gr], m, bb);
            }
            S::set(m, p0);
        }
        if (interior) {
            // frozen-viscosity sweeps go to the packed kernel compiled without the update
            const int ib = (m->desc.immersed_boundary ? 1 : 0) | (pinst[0].back().a.frozen_eta ? 2 : 0);
            const int nbx = fg.nbx;
            const int nby = (fg.hi[1] - fg.lo[1] + 1 + ty - 3) / (ty - 2);
            const int nbz = (fg.hi[2] - fg.lo[2] + fg.kch) / fg.kch;
            for (int bz = 0; bz < nbz; ++bz)
                for (int by = 0; by < nby; ++by)
                    for (int bx = 0; bx < nbx; ++bx) items[ib].push_back(make_int4(r, bx, by, bz));
        }
        for (int q = 0; q < shell.nb; ++q) {
            const Box& x = shell.b[q];
            PackedBox pb;
            pb.inst = r;
            pb.i0 = x.i0;
            pb.i1 = x.i1;
            pb.j0 = x.j0;
            pb.j1 = x.j1;
            pb.k0 = m->g.nd == 3 ? x.k0 : 0;
            pb.k1 = m->g.nd == 3 ? x.k1 : 0;
            pb.start = (int)blk2box[kind].size();
            const long n = (long)(pb.i1 - pb.i0 + 1) * (pb.j1 - pb.j0 + 1) * (pb.k1 - pb.k0 + 1);
            const int nb = (int)((n + 255) / 256);
            const int idx = (int)boxes[kind].size();
            boxes[kind].push_back(pb);
            for (int k = 0; k < nb; ++k) blk2box[kind].push_back(idx);
        }
    }
    std::vector<char> h;
    size_t o_inst[2], o_items[4], o_box[4], o_blk[4], o_bc[2][2][3], o_bcblk[2][2][3];
    for (int p = 0; p < 2; ++p) o_inst[p] = put(h, pinst[p]);
    for (int q = 0; q < 4; ++q) o_items[q] = put(h, items[q]);
    for (int k = 0; k < 4; ++k) {
        o_box[k] = put(h, boxes[k]);
        o_blk[k] = put(h, blk2box[k]);
    }
    for (int p = 0; p < 2; ++p)
        for (int gr = 0; gr < 2; ++gr)
            for (int d = 0; d < 3; ++d) {
                o_bc[p][gr][d] = put(h, bcops[p][gr][d]);
                o_bcblk[p][gr][d] = put(h, bcblk[p][gr][d]);
            }
    size_t o_bcall[2][2], o_bcallblk[2][2];
    for (int p = 0; p < 2; ++p)
        for (int gr = 0; gr < 2; ++gr) {
            if (!bcall_ok[p][gr]) {
                bcall[p][gr].clear();
                bcallblk[p][gr].clear();
            }
            o_bcall[p][gr] = put(h, bcall[p][gr]);
            o_bcallblk[p][gr] = put(h, bcallblk[p][gr]);
        }
    BatchTables& t = b->t;
    if (h.size() > t.cap) {
        HIPCHK(hipStreamSynchronize(b->stream)); // graphs in flight may still read the old tables
        destroy_graphs(b);
        if (t.dev) HIPCHK(hipFree(t.dev));
        t.dev = nullptr;
        t.cap = 0;
        size_t cap = h.size() + h.size() / 2 + 4096;
        if (hipMalloc(&t.dev, cap) != hipSuccess) return fail(FI_ERR_OOM, "batch tables (%zu bytes)", cap);
        t.cap = cap;
    } else {
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    if (!h.empty()) HIPCHK(hipMemcpy(t.dev, h.data(), h.size(), hipMemcpyHostToDevice));
    for (int p = 0; p < 2; ++p) t.inst[p] = t.dev + o_inst[p];
    for (int q = 0; q < 4; ++q) {
        t.items[q] = reinterpret_cast<int4*>(t.dev + o_items[q]);
        t.nitems[q] = (int)items[q].size();
    }
    for (int k = 0; k < 4; ++k) {
        t.box[k] = reinterpret_cast<PackedBox*>(t.dev + o_box[k]);
        t.blk2box[k] = reinterpret_cast<int*>(t.dev + o_blk[k]);
        t.nblk_box[k] = (int)blk2box[k].size();
    }
    for (int p = 0; p < 2; ++p)
        for (int gr = 0; gr < 2; ++gr)
            for (int d = 0; d < 3; ++d) {
                t.bc[p][gr][d] = reinterpret_cast<PackedBcOp*>(t.dev + o_bc[p][gr][d]);
                t.blk2bc[p][gr][d] = reinterpret_cast<int*>(t.dev + o_bcblk[p][gr][d]);
                t.nblk_bc[p][gr][d] = (int)bcblk[p][gr][d].size();
            }
    for (int p = 0; p < 2; ++p)
        for (int gr = 0; gr < 2; ++gr) {
            t.bc_all_ok[p][gr] = bcall_ok[p][gr];
            t.bc_all[p][gr] = reinterpret_cast<PackedBcAll*>(t.dev + o_bcall[p][gr]);
            t.blk2bcall[p][gr] = reinterpret_cast<int*>(t.dev + o_bcallblk[p][gr]);
            t.nblk_bcall[p][gr] = (int)bcallblk[p][gr].size();
        }
    b->ty = ty;
    b->kch = kch;
    b->phase = 0;
    b->rebuilds++;
    ++t.gen;
    return FI_OK;
}

static int build_tables(fi_batch_s* b, int ty, int kch)
{
    return b->refined == 1 ? build_tables_t<float>(b, ty, kch) : build_tables_t<double>(b, ty, kch);
}

// persistent grid of the fused sweep: enough workgroups for every CU
static int fused_grid(int nitems) { return nitems < 8192 ? nitems : 8192; }

template <typename T>
static void launch_fused_all_t(fi_batch_s* b, int p)
{
    const BatchTables& t = b->t;
    for (int q = 0; q < 4; ++q)
        launch_fused_packed<T>(static_cast<const PackedInst<T>*>(t.inst[p]), t.items[q], t.nitems[q], (q & 1) != 0,
                               (q & 2) == 0, b->tx, b->ty, b->minw, fused_grid(t.nitems[q]), b->stream);
}

template <typename T>
static void launch_points_t(fi_batch_s* b, int p, bool sigma)
{
    const BatchTables& t = b->t;
    const int kib[4] = {0, 1, 0, 1}, knd[4] = {3, 3, 2, 2};
    for (int k = 0; k < 4; ++k)
        launch_point_packed<T>(static_cast<const PackedInst<T>*>(t.inst[p]), t.box[k], t.blk2box[k], t.nblk_box[k],
                               kib[k] != 0, knd[k], sigma, b->stream);
}

static void launch_fused_all(fi_batch_s* b, int p)
{
    if (b->refined == 1) launch_fused_all_t<float>(b, p);
    else launch_fused_all_t<double>(b, p);
}

static void launch_points(fi_batch_s* b, int p, bool sigma)
{
    if (b->refined == 1) launch_points_t<float>(b, p, sigma);
    else launch_points_t<double>(b, p, sigma);
}

static void launch_bcs(fi_batch_s* b, int p, int group)
{
    const BatchTables& t = b->t;
    if (t.bc_all_ok[p][group]) {
        launch_bc_all_packed(t.bc_all[p][group], t.blk2bcall[p][group], t.nblk_bcall[p][group], b->stream);
        return;
    }
    for (int d = 0; d < 3; ++d)
        launch_bc_packed(t.bc[p][group][d], t.blk2bc[p][group][d], t.nblk_bc[p][group][d], b->stream);
}

// one PT iteration of every running instance at table parity p
static void launch_iteration(fi_batch_s* b, int p)
{
    launch_fused_all(b, p);
    launch_points(b, p, true);
    launch_bcs(b, p, 0);
    launch_points(b, p, false);
    launch_bcs(b, p, 1);
}

static std::vector<fi_model_s*> running_models(const fi_batch_s* b)
{
    std::vector<fi_model_s*> v;
    for (int r : b->run) v.push_back(b->inst[r].m);
    return v;
}

static int capture_graphs(fi_batch_s* b)
{
    destroy_graphs(b);
    b->graph_models.clear();
    if (b->graph_iters <= 0) return FI_OK;
    for (int p = 0; p < 2; ++p) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < b->graph_iters; ++it) launch_iteration(b, (p + it) & 1);
        HIPCHK(hipStreamEndCapture(b->stream, &g));
        hipError_t e = hipGraphInstantiate(&b->graph[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(FI_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        b->graph_gen[p] = b->t.gen;
    }
    b->graph_models = running_models(b);
    return FI_OK;
}

// Replay the graph of parity p, capturing both graphs first when there are
// none for the current tables (capture is lazy: a running set that changes
// before its graphs would be replayed never pays for them).  A graph's
// launches point into the table buffer and the member models' fields: one
// captured against other tables or another running set is never replayed
// (the tables are rewritten in place, and removed models may be destroyed).
static int graph_launch(fi_batch_s* b, int p)
{
    if (!b->graph[0] || !b->graph[1]) {
        int rc = capture_graphs(b);
        if (rc) return rc;
    }
    if (!b->graph[p] || b->graph_gen[p] != b->t.gen || b->graph_models != running_models(b))
        return fail(FI_ERR_STATE, "batch graph %d is stale (captured at table generation %u, tables at %u)", p,
                    b->graph_gen[p], b->t.gen);
    HIPCHK(hipGraphLaunch(b->graph[p], b->stream));
    return FI_OK;
}

static long running_cells(fi_batch_s* b)
{
    long c = 0;
    for (int r : b->run) {
        const Geo& g = b->inst[r].m->g;
        c += (long)g.n[0] * g.n[1] * (g.nd == 3 ? g.n[2] : 1);
    }
    return c;
}

// Time the fused sweep of every tile candidate on genuine iterations of the
// running set (the tile shape does not change any result: every compiled
// variant equals the pointwise sweep bit for bit) and keep the fastest.
// FP64 batches try every FP64 variant with z chunks of 32 and 64 and tile
// orders 0 and 2; refined batches every FP32 variant (their sweeps run in
// chunks of about 32 between outer steps).  Each candidate runs one warm-up
// iteration and three timed ones.  Uses at most `budget` iterations;
// returns how many it ran.
static std::vector<unsigned> fp64_zbits(fi_batch_s* b);
static std::vector<long> refined_signature(fi_batch_s* b);

static int autotune(fi_batch_s* b, int budget, int* used)
{
    *used = 0;
    const bool f32 = b->refined == 1;
    struct Cand {
        int tx, ty, minw, kch, order;
    };
    std::vector<Cand> cands;
    int cfg[16][3];
    const int nv = fused_configs(f32, cfg, 16);
    for (int v = 0; v < nv; ++v) {
        if (f32) {
            cands.push_back(Cand{cfg[v][0], cfg[v][1], cfg[v][2], 32, 2});
            continue;
        }
        for (int kch : {32, 64})
            for (int order : {0, 2}) cands.push_back(Cand{cfg[v][0], cfg[v][1], cfg[v][2], kch, order});
    }
    const int nc = (int)cands.size();
    const int reps = 4;
    if (budget < nc * reps) return FI_OK;
    int best = -1;
    float best_ms = 0.0f;
    for (int c = 0; c < nc; ++c) {
        b->tx = cands[c].tx;
        b->minw = cands[c].minw;
        b->order = cands[c].order;
        int rc = build_tables(b, cands[c].ty, cands[c].kch);
        if (rc) return rc;
        float ms = 0.0f;
        for (int k = 0; k < reps; ++k) {
            // full iterations; iterations 1 to reps-1 are timed up to the
            // last one's fused launches (the rest costs every candidate the same)
            const int p = b->phase;
            if (k == 1) HIPCHK(hipEventRecord(b->ev[0], b->stream));
            launch_fused_all(b, p);
            if (k == reps - 1) HIPCHK(hipEventRecord(b->ev[1], b->stream));
            launch_points(b, p, true);
            launch_bcs(b, p, 0);
            launch_points(b, p, false);
            launch_bcs(b, p, 1);
            // the tables were rebuilt for this candidate: keep the models' parity
            // in step with every iteration run here
            for (int r : b->run) {
                fi_model_s* m = b->inst[r].m;
                if (f32) {
                    using S = SweepOf<float>;
                    S::swept(m, S::get(m), 1);
                    S::set(m, S::flipped(m, S::get(m), 1));
                    b->inst[r].left--;
                } else {
                    SweepOf<double>::swept(m, SweepOf<double>::get(m), 1);
                    m->sig_cur ^= 1;
                    m->v_cur ^= 1;
                    m->eta_cur ^= 1;
                    b->inst[r].iters++;
                }
            }
            b->phase ^= 1;
            (*used)++;
        }
        HIPCHK(hipEventSynchronize(b->ev[1]));
        HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        if (best < 0 || ms < best_ms) {
            best = c;
            best_ms = ms;
        }
        // the next candidate's tables start from the models' current parity
        b->phase = 0;
    }
    b->tune_runs++;
    b->tuned_cells = running_cells(b);
    b->tx = cands[best].tx;
    b->minw = cands[best].minw;
    b->order = cands[best].order;
    int rc = build_tables(b, cands[best].ty, cands[best].kch);
    if (rc) return rc;
    destroy_graphs(b); // captured when first replayed
    // the tuning iterations changed the models' parities and clean-buffer
    // bits, and the tables just built hold the new ones
    if (f32) b->sig = refined_signature(b);
    else b->zbuilt = fp64_zbits(b);
    b->dirty = false;
    return FI_OK;
}

static std::vector<unsigned> fp64_zbits(fi_batch_s* b)
{
    std::vector<unsigned> z;
    for (int r : b->run) z.push_back(SweepOf<double>::zbits(b->inst[r].m));
    return z;
}

static int rebuild(fi_batch_s* b)
{
    b->run.clear();
    for (size_t q = 0; q < b->inst.size(); ++q)
        if (b->inst[q].state == FI_BATCH_RUNNING) b->run.push_back((int)q);
    for (int r : b->run) {
        int rc = prepare(b->inst[r].m);
        if (rc) return rc;
    }
    int rc = build_tables(b, b->ty, b->kch);
    if (rc) return rc;
    destroy_graphs(b); // captured when first replayed
    b->zbuilt = fp64_zbits(b);
    b->dirty = false;
    return FI_OK;
}

static int finish_check(BatchInst& bi)
{
    residual_norms_read(bi.m, bi.norms);
    bool conv = true;
    for (int q = 0; q < 4; ++q) {
        if (!std::isfinite(bi.norms[q])) {
            bi.state = FI_BATCH_FAILED;
            return 1;
        }
        if (!(bi.norms[q] < bi.tol)) conv = false;
    }
    if (conv) {
        bi.state = FI_BATCH_CONVERGED;
        return 1;
    }
    return 0;
}

// A refined batch advances every running job's FP32 sweeps together up to
// the nearest job's check, then runs the checks (and any FP64 outer steps)
// of the jobs whose chunk is complete.  Tables follow the jobs' buffers and
// modes, so they are rebuilt every step.
// what the refined tables depend on: tile options, the running set and each
// model's buffer parity and phase (full nonlinear or correction)
static std::vector<long> refined_signature(fi_batch_s* b)
{
    std::vector<long> sg = {b->tx, b->ty, b->minw, b->kch, b->order};
    for (int r : b->run) {
        fi_model_s* m = b->inst[r].m;
        const RefineParity p = refine_parity(m);
        sg.push_back(r);
        sg.push_back((long)(intptr_t)m);
        sg.push_back(p.sig | (p.v << 1) | (p.eta << 2) | (refine_full(m) ? 8 : 0) | SweepOf<float>::zbits(m) << 4);
    }
    return sg;
}

// Refined batches: every running job has a pending chunk of FP32 sweeps; run
// the smallest of them for all jobs (packed, replayed from graphs while the
// tables stay valid), then give every job whose chunk ended its check: the
// checks of all due jobs are enqueued around one synchronisation.
static int step_refined(fi_batch_s* b, int* n_finished)
{
    b->run.clear();
    for (size_t q = 0; q < b->inst.size(); ++q)
        if (b->inst[q].state == FI_BATCH_RUNNING) b->run.push_back((int)q);
    if (b->run.empty()) return FI_OK;
    int delta = 1 << 30;
    for (int r : b->run) delta = std::min(delta, b->inst[r].left);
    if (delta < 1) return fail(FI_ERR_STATE, "refined batch job without pending sweeps");
    for (int r : b->run) {
        int rc = prepare(b->inst[r].m, false);
        if (rc) return rc;
    }
    std::vector<long> sg = refined_signature(b);
    if (b->dirty || sg != b->sig) {
        int rc = build_tables(b, b->ty, b->kch);
        if (rc) return rc;
        destroy_graphs(b); // captured when first replayed
        b->sig = sg;
        b->dirty = false;
    }
    // tile candidates timed on the first sweeps of this chunk whenever the
    // running set's size has moved by more than a factor of two; the tuning
    // sweeps are sweeps of the chunk (accounted for by autotune)
    int tuned = 0;
    const long cells = running_cells(b);
    if (b->autotune && (b->tuned_cells == 0 || cells > 2 * b->tuned_cells || 2 * cells < b->tuned_cells)) {
        b->phase = 0;
        int rc = autotune(b, delta, &tuned);
        if (rc) return rc;
    }
    const int rest = delta - tuned;
    int done = 0, p = 0;
    while (b->graph_iters > 0 && rest - done >= b->graph_iters) {
        int rc = graph_launch(b, p);
        if (rc) return rc;
        done += b->graph_iters;
        p ^= (b->graph_iters & 1);
    }
    for (; done < rest; ++done, p ^= 1) launch_iteration(b, p);
    HIPCHK(hipGetLastError());
    using S = SweepOf<float>;
    std::vector<int> due;
    for (int r : b->run) {
        BatchInst& bi = b->inst[r];
        S::swept(bi.m, S::get(bi.m), rest);
        S::set(bi.m, S::flipped(bi.m, S::get(bi.m), rest));
        bi.left -= rest;
        if (bi.left <= 0) due.push_back(r);
    }
    if (!due.empty()) {
        // each due model's outer step (a dozen small launches and a copy of
        // its norms) on one of the auxiliary streams, after the sweeps: the
        // steps of different models overlap instead of queueing one launch
        // at a time behind each other; every model computes what it
        // computes alone, on its own buffers
        HIPCHK(hipEventRecord(b->ev_fork, b->stream));
        for (int q = 0; q < fi_batch_s::NAUX; ++q) HIPCHK(hipStreamWaitEvent(b->aux[q], b->ev_fork, 0));
        int qi = 0;
        for (int r : due) {
            fi_model_s* m = b->inst[r].m;
            m->stream = b->aux[qi];
            qi = (qi + 1) % fi_batch_s::NAUX;
            const int rc = job_chunk_enqueue(m, b->inst[r].job);
            m->stream = b->stream;
            if (rc) return rc;
        }
        for (int q = 0; q < fi_batch_s::NAUX; ++q) HIPCHK(hipStreamSynchronize(b->aux[q]));
    }
    for (int r : due) {
        BatchInst& bi = b->inst[r];
        int rc = job_chunk_finish(bi.m, bi.job);
        bi.left = bi.job.chunk;
        bi.iters = bi.job.inner_total;
        for (int q = 0; q < 4; ++q) bi.norms[q] = bi.job.norms[q];
        if (rc == FI_ERR_NONFINITE) {
            bi.state = FI_BATCH_FAILED;
            (*n_finished)++;
            continue;
        }
        if (rc) return rc;
        if (bi.job.state == JOB_DONE) {
            bi.state = bi.job.converged ? FI_BATCH_CONVERGED : FI_BATCH_MAX_ITER;
            (*n_finished)++;
        }
    }
    return FI_OK;
}

static int step(fi_batch_s* b, int* n_finished)
{
    *n_finished = 0;
    if (b->refined == 1) return step_refined(b, n_finished);
    // the tables carry the models' clean-buffer bits: rebuild when they changed
    if (b->dirty || fp64_zbits(b) != b->zbuilt) {
        int rc = rebuild(b);
        if (rc) return rc;
    }
    if (b->run.empty()) return FI_OK;
    // iterations to the next residual check (or the iteration cap) of the nearest instance
    int delta = 1 << 30;
    for (int r : b->run) {
        const BatchInst& bi = b->inst[r];
        const int to_check = bi.ncheck - bi.iters % bi.ncheck;
        delta = std::min(delta, std::min(to_check, bi.max_iter - bi.iters));
    }
    int done = 0;
    const long cells = running_cells(b);
    if (b->autotune && (b->tuned_cells == 0 || cells > 2 * b->tuned_cells || 2 * cells < b->tuned_cells)) {
        int rc = autotune(b, delta, &done);
        if (rc) return rc;
    }
    int left = delta - done;
    while (b->graph_iters > 0 && left >= b->graph_iters) {
        int rc = graph_launch(b, b->phase);
        if (rc) return rc;
        left -= b->graph_iters;
        b->phase ^= (b->graph_iters & 1);
    }
    while (left > 0) {
        launch_iteration(b, b->phase);
        b->phase ^= 1;
        --left;
    }
    HIPCHK(hipGetLastError());
    const int adv = delta - done;
    std::vector<int> due;
    for (int r : b->run) {
        BatchInst& bi = b->inst[r];
        fi_model_s* m = bi.m;
        SweepOf<double>::swept(m, SweepOf<double>::get(m), adv);
        m->sig_cur ^= (adv & 1);
        m->v_cur ^= (adv & 1);
        m->eta_cur ^= (adv & 1);
        bi.iters += adv;
        if (bi.iters % bi.ncheck == 0) {
            int rc = residual_norms_enqueue(m);
            if (rc) return rc;
            due.push_back(r);
        }
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int r : b->run) {
        BatchInst& bi = b->inst[r];
        bool fin = false;
        if (std::find(due.begin(), due.end(), r) != due.end()) fin = finish_check(bi) != 0;
        if (!fin && bi.iters >= bi.max_iter) {
            bi.state = FI_BATCH_MAX_ITER;
            fin = true;
        }
        if (fin) {
            (*n_finished)++;
            b->dirty = true;
        }
    }
    return FI_OK;
}

} // namespace fi

extern "C" {

int fi_batch_create(fi_batch_t* out)
{
    if (!out) return fail(FI_ERR_ARG, "null output handle");
    *out = nullptr;
    fi_batch_s* b = new (std::nothrow) fi_batch_s();
    if (!b) return fail(FI_ERR_OOM, "host allocation failed");
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&b->ev[0]) == hipSuccess && hipEventCreate(&b->ev[1]) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int q = 0; q < fi_batch_s::NAUX && ok; ++q)
        ok = hipStreamCreateWithFlags(&b->aux[q], hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        for (int q = 0; q < fi_batch_s::NAUX; ++q)
            if (b->aux[q]) (void)hipStreamDestroy(b->aux[q]);
        if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
        for (int k = 0; k < 2; ++k)
            if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return fail(FI_ERR_HIP, "stream/event creation failed");
    }
    *out = b;
    return FI_OK;
}

int fi_batch_destroy(fi_batch_t b)
{
    if (!b) return FI_OK;
    (void)hipStreamSynchronize(b->stream);
    while (!b->inst.empty()) (void)detach(b, (int)b->inst.size() - 1);
    destroy_graphs(b);
    if (b->t.dev) (void)hipFree(b->t.dev);
    for (int k = 0; k < 2; ++k)
        if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
    for (int q = 0; q < fi_batch_s::NAUX; ++q) {
        (void)hipStreamSynchronize(b->aux[q]);
        (void)hipStreamDestroy(b->aux[q]);
    }
    (void)hipEventDestroy(b->ev_fork);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return FI_OK;
}

int fi_batch_add(fi_batch_t b, fi_model_t m, int max_iter, int ncheck, double tol)
{
    if (!b || !m) return fail(FI_ERR_ARG, "null argument");
    if (max_iter < 0 || ncheck < 1) return fail(FI_ERR_ARG, "bad max_iter/ncheck");
    if (m->batch) return fail(FI_ERR_ARG, "model already belongs to a batch");
    if (b->refined == 1) return fail(FI_ERR_STATE, "this batch holds refined solves");
    b->refined = 0;
    // the model's pending work completes before the batch stream takes over
    HIPCHK(hipStreamSynchronize(m->stream));
    BatchInst bi;
    bi.m = m;
    bi.home = m->stream;
    bi.max_iter = max_iter;
    bi.ncheck = ncheck;
    bi.tol = tol;
    bi.state = max_iter == 0 ? FI_BATCH_MAX_ITER : FI_BATCH_RUNNING;
    m->stream = b->stream;
    m->batch = b;
    b->inst.push_back(bi);
    b->dirty = true;
    return FI_OK;
}

int fi_batch_add_refined(fi_batch_t b, fi_model_t m, int max_outer, int max_inner, int ncheck, double inner_rtol,
                         double tol)
{
    if (!b || !m) return fail(FI_ERR_ARG, "null argument");
    if (m->batch) return fail(FI_ERR_ARG, "model already belongs to a batch");
    if (b->refined == 0) return fail(FI_ERR_STATE, "this batch holds FP64 solves");
    if (b->refined < 0) {
        // the FP32 sweep variant of the packed launches
        b->refined = 1;
        b->tx = 64;
        b->ty = 8;
        b->minw = 1;
        b->kch = 32;
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    BatchInst bi;
    bi.m = m;
    bi.home = m->stream;
    bi.refined = true;
    m->stream = b->stream;
    m->batch = b;
    int rc = job_begin(m, bi.job, max_outer, max_inner, ncheck, inner_rtol, tol);
    if (rc && rc != FI_ERR_NONFINITE) {
        (void)hipStreamSynchronize(b->stream);
        m->stream = bi.home;
        m->batch = nullptr;
        return rc;
    }
    for (int q = 0; q < 4; ++q) bi.norms[q] = bi.job.norms[q];
    bi.iters = bi.job.inner_total;
    bi.left = bi.job.chunk;
    if (rc == FI_ERR_NONFINITE) bi.state = FI_BATCH_FAILED;
    else if (bi.job.state == JOB_DONE) bi.state = bi.job.converged ? FI_BATCH_CONVERGED : FI_BATCH_MAX_ITER;
    else bi.state = FI_BATCH_RUNNING;
    b->inst.push_back(bi);
    b->dirty = true;
    return FI_OK;
}

int fi_batch_refined_counts(fi_batch_t b, fi_model_t m, int counts_out[3])
{
    if (!b || !m || !counts_out) return fail(FI_ERR_ARG, "null argument");
    int q = find_inst(b, m);
    if (q < 0) return fail(FI_ERR_ARG, "model is not in this batch");
    const BatchInst& bi = b->inst[q];
    if (!bi.refined) return fail(FI_ERR_ARG, "model is not a refined solve");
    counts_out[0] = bi.job.outer;
    counts_out[1] = bi.job.inner_total;
    counts_out[2] = bi.job.phase1;
    return FI_OK;
}

int fi_batch_remove(fi_batch_t b, fi_model_t m)
{
    if (!b || !m) return fail(FI_ERR_ARG, "null argument");
    int q = find_inst(b, m);
    if (q < 0) return fail(FI_ERR_ARG, "model is not in this batch");
    return detach(b, q);
}

int fi_batch_step(fi_batch_t b, int* n_finished)
{
    if (!b) return fail(FI_ERR_ARG, "null batch");
    int nf = 0;
    int rc = step(b, &nf);
    if (n_finished) *n_finished = nf;
    return rc;
}

int fi_batch_solve(fi_batch_t b)
{
    if (!b) return fail(FI_ERR_ARG, "null batch");
    for (;;) {
        int nf = 0;
        int rc = step(b, &nf);
        if (rc) return rc;
        bool any = false;
        for (const BatchInst& bi : b->inst) any = any || bi.state == FI_BATCH_RUNNING;
        if (!any) break;
    }
    return FI_OK;
}

int fi_batch_status(fi_batch_t b, fi_model_t m, int* state, int* iters, double norms_inf[4])
{
    if (!b || !m) return fail(FI_ERR_ARG, "null argument");
    int q = find_inst(b, m);
    if (q < 0) return fail(FI_ERR_ARG, "model is not in this batch");
    const BatchInst& bi = b->inst[q];
    if (state) *state = bi.state;
    if (iters) *iters = bi.iters;
    if (norms_inf)
        for (int k = 0; k < 4; ++k) norms_inf[k] = bi.norms[k];
    return FI_OK;
}

int fi_batch_set_option(fi_batch_t b, int key, int value)
{
    if (!b) return fail(FI_ERR_ARG, "null batch");
    if (key == FI_BOPT_VARIANT) {
        int cfg[16][3];
        const int n = fused_configs(b->refined == 1, cfg, 16);
        if (value < 0 || value >= n) return fail(FI_ERR_ARG, "fused variant must be in 0..%d", n - 1);
        b->tx = cfg[value][0];
        b->ty = cfg[value][1];
        b->minw = cfg[value][2];
    } else if (key == FI_BOPT_TILE_ORDER) {
        if (value < 0 || value > 2) return fail(FI_ERR_ARG, "tile order must be 0, 1 or 2");
        b->order = value;
    } else if (key == FI_BOPT_ZCHUNK) {
        if (value < 1) return fail(FI_ERR_ARG, "z chunk must be >= 1");
        b->kch = value;
    } else if (key == FI_BOPT_AUTOTUNE) {
        b->autotune = value ? 1 : 0;
    } else if (key == FI_BOPT_GRAPH_ITERS) {
        if (value < 0 || value > 64) return fail(FI_ERR_ARG, "graph iterations must be in 0..64");
        b->graph_iters = value & ~1; // even, so a replay keeps the table parity
    } else {
        return fail(FI_ERR_ARG, "unknown batch option %d", key);
    }
    b->dirty = true;
    return FI_OK;
}

int fi_batch_info(fi_batch_t b, int out[8])
{
    if (!b || !out) return fail(FI_ERR_ARG, "null argument");
    int running = 0;
    for (const BatchInst& bi : b->inst) running += bi.state == FI_BATCH_RUNNING;
    out[0] = (int)b->inst.size();
    out[1] = running;
    out[2] = b->t.nitems[0] + b->t.nitems[1] + b->t.nitems[2] + b->t.nitems[3];
    out[3] = (b->tx * 32 + b->ty) * 16 + b->minw;
    out[4] = b->kch * 4 + b->order;
    out[5] = b->graph_iters;
    out[6] = b->tune_runs;
    out[7] = b->rebuilds;
    return FI_OK;
}

} // extern "C"
