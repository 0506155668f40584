// refine.h -- the FP32 defect-correction solver as a resumable job, shared
// by fi_refined_solve (refine.hip) and refined batches (batch.hip).
#pragma once

#include "model.h"

namespace fi {

struct AaState {
    int depth = 0;     // columns kept (0: plain defect correction)
    int cols = 0;      // columns in use
    int have_prev = 0; // a previous correction f_{k-1} is stored
    int prev_slot = 0; // ring slot of the last applied update u_{k-1}
};

enum { JOB_PHASE1 = 0, JOB_INNER = 1, JOB_DONE = 2 };

struct RefineJob {
    int max_outer = 0, max_inner = 0, ncheck = 1;
    int mode = 1;         // 1 lagged PT, 0 correction solves
    int since_eta = 0;    // lagged PT: FP32 sweeps since the last viscosity update
    double lag_best = -1.0; // lagged PT: best outer residual since the chunk length changed
    int lag_stall = 0;    // lagged PT: sweeps since that residual last halved
    int lag_window = 0;   // lagged PT: stall limit in sweeps; then the chunk length halves (down to 8)
    long sweep_cap = 0;   // lagged PT: FP32 sweeps allowed in all (max_outer x the initial ncheck)
    double inner_rtol = 0.0, tol = 0.0;
    bool nonlinear = false, trace = false;
    int state = JOB_DONE;
    int chunk = 0;        // sweeps of the pending chunk
    int last = 0;         // sweeps of the chunk just run
    int done = 0;         // sweeps of the current phase 1 / correction solve
    int outer = 0, inner_total = 0, phase1 = 0;
    double p1_rtol = 0.0, first = -1.0, best = -1.0; // phase 1: drop, peak and best residuals
    int stall = 0, stall_window = 0; // phase 1: sweeps without halving, and the limit
    double omega = 1.0, r_prev = -1.0, r0 = 0.0;
    bool aa_restart = false;
    AaState aa;
    double norms[4] = {0, 0, 0, 0};
    bool converged = false;
    bool pending_outer = false; // the enqueued half was a lagged-PT outer step
    // lagged PT, 3D immersed boundary with traction on every face: an outer
    // step over the whole grid has run in this job, so the FP32 corrections,
    // the stress lag and the FP64 stress are 0 outside the column ranges and
    // later outer steps visit the ranges only
    bool cols_clean = false;
};

// job control: begin (may finish at once), sweeps of the pending chunk, and
// the check after the chunk has run
int job_begin(fi_model_s* m, RefineJob& j, int max_outer, int max_inner, int ncheck, double inner_rtol, double tol);
int job_after_chunk(fi_model_s* m, RefineJob& j);
// job_after_chunk in two halves: enqueue the device work the decision needs,
// then (after the stream has synchronised) decide; a batch enqueues the
// halves of all its due jobs around one synchronisation
int job_chunk_enqueue(fi_model_s* m, RefineJob& j);
int job_chunk_finish(fi_model_s* m, RefineJob& j);

// pieces of one FP32 sweep for packed launches: the sweep's kernel arguments
// (current buffers in, other buffers out), the boundary batch of one group
// and dimension, the fused interior of the FP32 sweeps, and the buffer flips
// after a sweep
StokesArgs<float> refine_args(fi_model_s* m, bool in_place);
BcBatch refine_bc_batch(fi_model_s* m, int group, int d);
bool refine_fused(fi_model_s* m, FusedGeom& fg, BoxList& shell);
void refine_flip(fi_model_s* m);
struct RefineParity {
    int sig, v, eta;
};
RefineParity refine_parity(fi_model_s* m);
void refine_set_parity(fi_model_s* m, RefineParity p);
bool refine_full(fi_model_s* m);

} // namespace fi
