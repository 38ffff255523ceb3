// Engine internals shared by the C ABI (engine.hip) and the request server
// (serve.cpp): the device-resident time loop of one context in phases, so a
// caller can keep several contexts (streams) busy at once.
#pragma once
#include <vector>

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

}  // namespace opf
