// Engine internals shared by the C ABI (engine.hip) and the request server
// (serve.cpp): the device-resident time loop of one context in phases, so a
// caller can keep several contexts (streams) busy at once.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "opf_internal.hpp"

struct opf_ctx;

namespace opf {

struct RunPlan {
  std::vector<int> slots;   // slots advancing in this run
  int lo = 0, ns = 0;       // launch range [lo, lo+ns) (0 slots: nothing to do)
  int max_steps = 0;        // largest step count requested
  std::vector<int> live;    // slots still active at the last run_remaining: once they are at
                            // most half the range, iterations launch over this list only
};

// Mark `slots` active for nsteps[i] more steps with corr_fixed[i] correctors
// per step (0: to convergence).  Failed slots throw when `strict`, else are
// left out of the plan.
RunPlan run_begin(opf_ctx* c, const std::vector<int>& slots, const std::vector<int>& nsteps,
                  const std::vector<int>& corr_fixed, int diag_every, bool strict);
// Enqueue evaluation + step iterations on the context's stream (no wait).
void run_enqueue(opf_ctx* c, const RunPlan& p, long iterations);
// Wait for the stream; the most steps any slot of the plan still needs
// (and the plan's list of slots still active).
int run_remaining(opf_ctx* c, RunPlan& p);

// ---- asynchronous slot operations (the server's continuous batching) -------
// All are enqueued on the context's stream and return without waiting, so a
// request joins or leaves between two iterations of the other slots.  Host
// buffers passed in must be pinned (hipHostMalloc) and stay untouched until
// the stream has passed the operation.

hipStream_t ctx_stream(opf_ctx* c);
// double2 per slot state in the device row layout (3 components x N x Hp)
size_t state_rows(const opf_ctx* c);
// host layout changes between the coll planes (ur.Re, ur.Im, ut.Re, ...) and rows
void coll_to_rows(const opf_ctx* c, const double* const planes[6], double2* rows);
void rows_to_coll(const opf_ctx* c, const double2* rows, double* const planes[6]);
// vel_matrices for Re, alpha, dt (implicitness 0.51, constant flux), the
// state of `rows` loaded into velocity buffer `ucur` (the slot's current
// role, fixed while it is idle), vel_imposesym and the curl, then a fresh
// run of nsteps steps with corr_fixed correctors per step.  A singular
// factor leaves the slot inactive with a nonzero status.
void admit_async(opf_ctx* c, int slot, double Re, double alpha, double dt, int ucur, const double2* rows,
                 int nsteps, int corr_fixed);
// one iteration (evaluation + step kernel + control) over slots [s0, s0 + ns)
void enqueue_iteration_range(opf_ctx* c, int s0, int ns);
// the control blocks of all slots into host memory
void ctl_snapshot_async(opf_ctx* c, SlotCtl* out);
// the state of an idle slot (velocity buffer `ucur`) into host rows
void fetch_async(opf_ctx* c, int slot, int ucur, double2* rows);

}  // namespace opf
