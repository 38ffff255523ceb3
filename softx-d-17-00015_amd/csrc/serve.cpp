// Continuous-batching server for integration requests of mixed resolution.
//
// A request is a time integration of one state for a given number of steps
// with its own Re, alpha and dt, the unit of work that steporbit
// (newton.f90:743-787) hands to the time stepper, with one corrector per
// step there.  The server keeps one context per resolution class with a
// fixed number of slots (sized by the caller against the HBM budget with
// opf_ctx_bytes_needed).  One scheduler thread
//   1. admits queued requests into free slots of their class (the factors
//      of vel_matrices and the state load run on the device),
//   2. advances every occupied slot of every class by up to `chunk` steps,
//      enqueueing all classes' launches on their own streams before waiting,
//      so the classes share the GPU,
//   3. retires finished requests (final state copied to the host) and frees
//      their slots for the queue.
// submit / poll / result only touch the request table, under a mutex; the
// contexts are used by the scheduler thread alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/opf.h"
#include "engine_internal.hpp"

namespace {

struct Request {
  long long id = 0;
  int cls = 0;
  double Re = 0, alpha = 0, dt = 0;
  int nsteps = 0, corr_fixed = 0;
  int remaining = 0;
  std::vector<double> state;   // 6 coll planes: initial, then final
  double t = 0;                // simulation time of the final state
  double t_submit = 0, t_admit = -1, t_done = -1;
  int status = 0;              // 0 ok, else an OPF_ERR_* code or a corrector failure
  bool done = false;
  int slot = -1;
};
using ReqPtr = std::shared_ptr<Request>;

struct Class {
  int N = 0, K = 0, M = 0, H = 0, cap = 0;
  opf_ctx* ctx = nullptr;
  std::vector<ReqPtr> slots;     // request in each slot (null: free)
  std::deque<ReqPtr> queue;
};

thread_local std::string s_err;

int fail(int code, const std::string& m) {
  s_err = m;
  return code;
}

}  // namespace

struct opf_server {
  std::vector<Class> cls;
  int chunk = 50;
  std::mutex mu;
  std::condition_variable cv;
  std::unordered_map<long long, ReqPtr> reqs;
  long long next_id = 1;
  long long state_steps = 0;
  int running = 0;
  bool stop = false;
  std::thread th;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // planes of a class's state
  static void planes(const Class& c, double* base, double* p[6]) {
    const size_t n = (size_t)c.N * c.H;
    for (int i = 0; i < 6; ++i) p[i] = base + i * n;
  }

  // retire a request: result recorded, slot freed (scheduler thread)
  void retire(Class& c, int slot, const ReqPtr& r, int status) {
    std::lock_guard<std::mutex> g(mu);
    r->status = status;
    r->done = true;
    r->t_done = now();
    r->slot = -1;
    c.slots[slot] = nullptr;
    --running;
  }

  void admit(Class& c, int slot, const ReqPtr& r) {
    double* p[6];
    planes(c, r->state.data(), p);
    int rc = opf_slot_setup(c.ctx, slot, r->Re, r->alpha, r->dt, 0.51, 1);
    if (rc == OPF_OK) rc = opf_put_state(c.ctx, slot, p[0], p[1], p[2], p[3], p[4], p[5], 0.0);
    {
      std::lock_guard<std::mutex> g(mu);
      r->t_admit = now();
    }
    if (rc != OPF_OK) retire(c, slot, r, rc);
  }

  void loop() {
    for (;;) {
      // 1. admission
      std::vector<std::pair<int, int>> admitted;   // (class, slot)
      {
        std::unique_lock<std::mutex> g(mu);
        for (;;) {
          if (stop) return;
          for (size_t ci = 0; ci < cls.size(); ++ci) {
            Class& c = cls[ci];
            for (int s = 0; s < c.cap && !c.queue.empty(); ++s)
              if (!c.slots[s]) {
                ReqPtr r = c.queue.front();
                c.queue.pop_front();
                c.slots[s] = r;
                r->slot = s;
                ++running;
                admitted.push_back({(int)ci, s});
              }
          }
          if (running > 0) break;
          cv.wait(g);
        }
      }
      for (auto [ci, s] : admitted) admit(cls[ci], s, cls[ci].slots[s]);

      // 2. advance every occupied slot of every class by up to `chunk` steps
      std::vector<opf::RunPlan> plans(cls.size());
      std::vector<std::vector<std::pair<int, ReqPtr>>> act(cls.size());
      for (size_t ci = 0; ci < cls.size(); ++ci) {
        Class& c = cls[ci];
        std::vector<int> sl, ns, cf;
        for (int s = 0; s < c.cap; ++s) {
          ReqPtr r = c.slots[s];
          if (!r || r->done) continue;
          sl.push_back(s);
          ns.push_back(std::min(r->remaining, chunk));
          cf.push_back(r->corr_fixed);
          act[ci].push_back({s, r});
        }
        if (sl.empty()) continue;
        try {
          plans[ci] = opf::run_begin(c.ctx, sl, ns, cf, 0, false);
          opf::run_enqueue(c.ctx, plans[ci], 2L * plans[ci].max_steps);
        } catch (const std::exception& e) {
          for (auto& [s, r] : act[ci]) retire(c, s, r, OPF_ERR_HIP);
          act[ci].clear();
          plans[ci] = opf::RunPlan{};
        }
      }
      for (size_t ci = 0; ci < cls.size(); ++ci) {
        if (plans[ci].ns == 0) continue;
        try {
          for (int round = 0; round < 64; ++round) {
            const int rem = opf::run_remaining(cls[ci].ctx, plans[ci]);
            if (rem == 0) break;
            opf::run_enqueue(cls[ci].ctx, plans[ci], 2L * rem);
          }
        } catch (const std::exception& e) {
          for (auto& [s, r] : act[ci]) retire(cls[ci], s, r, OPF_ERR_HIP);
          act[ci].clear();
        }
      }

      // 3. bookkeeping and retirement
      for (size_t ci = 0; ci < cls.size(); ++ci) {
        Class& c = cls[ci];
        for (auto& [s, r] : act[ci]) {
          int step = 0, status = 0, nc = 0;
          double t = 0, de = 0, pr0 = 0;
          if (opf_slot_status(c.ctx, s, &step, &t, &status, &nc, &de, &pr0) != OPF_OK) {
            retire(c, s, r, OPF_ERR_HIP);
            continue;
          }
          // the step counter restarts at admission (opf_put_state)
          {
            std::lock_guard<std::mutex> g(mu);
            state_steps += step - (r->nsteps - r->remaining);
            r->remaining = r->nsteps - step;
          }
          if (status != 0) {
            retire(c, s, r, OPF_ERR_CORRECTOR);
          } else if (r->remaining <= 0) {
            double* p[6];
            planes(c, r->state.data(), p);
            double tt = 0;
            const int rc = opf_get_state(c.ctx, s, p[0], p[1], p[2], p[3], p[4], p[5], &tt);
            r->t = tt;
            retire(c, s, r, rc);
          }
        }
      }
    }
  }
};

extern "C" {

const char* opf_server_last_error(void) { return s_err.c_str(); }

int opf_server_create(opf_server** out, int nclass, const int* nkm, const int* capacity, int chunk) {
  if (!out || nclass < 1 || !nkm || !capacity || chunk < 1) return fail(OPF_ERR_ARG, "bad server arguments");
  *out = nullptr;
  auto* sv = new opf_server();
  sv->chunk = chunk;
  sv->cls.resize(nclass);
  for (int i = 0; i < nclass; ++i) {
    Class& c = sv->cls[i];
    c.N = nkm[3 * i];
    c.K = nkm[3 * i + 1];
    c.M = nkm[3 * i + 2];
    c.cap = capacity[i];
    if (c.cap < 1) {
      opf_server_destroy(sv);
      return fail(OPF_ERR_ARG, "class capacity must be >= 1");
    }
    const int rc = opf_ctx_create(&c.ctx, c.N, c.K, c.M, 1, c.cap, nullptr);
    if (rc != OPF_OK) {
      const std::string m = opf_last_error();
      opf_server_destroy(sv);
      return fail(rc, "class " + std::to_string(i) + ": " + m);
    }
    int n, k, m, mp, h, cap;
    size_t bytes;
    opf_ctx_dims(c.ctx, &n, &k, &m, &mp, &h, &cap, &bytes);
    c.H = h;
    c.slots.assign(c.cap, nullptr);
  }
  sv->th = std::thread([sv] { sv->loop(); });
  *out = sv;
  return OPF_OK;
}

int opf_server_destroy(opf_server* sv) {
  if (!sv) return OPF_OK;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_all();
  if (sv->th.joinable()) sv->th.join();
  for (auto& c : sv->cls)
    if (c.ctx) opf_ctx_destroy(c.ctx);
  delete sv;
  return OPF_OK;
}

int opf_server_submit(opf_server* sv, int cls, double Re, double alpha, double dt, int nsteps, int corr_fixed,
                      const double* ur_re, const double* ur_im, const double* ut_re, const double* ut_im,
                      const double* uz_re, const double* uz_im, long long* id) {
  if (!sv || !id || cls < 0 || cls >= (int)sv->cls.size()) return fail(OPF_ERR_ARG, "bad class");
  if (nsteps < 1 || corr_fixed < 0 || !(dt > 0.0) || !(Re > 0.0) || alpha == 0.0)
    return fail(OPF_ERR_ARG, "bad request parameters");
  const double* in[6] = {ur_re, ur_im, ut_re, ut_im, uz_re, uz_im};
  for (auto* p : in)
    if (!p) return fail(OPF_ERR_ARG, "null state plane");
  auto r = std::make_shared<Request>();
  const Class& c = sv->cls[cls];
  const size_t n = (size_t)c.N * c.H;
  r->state.resize(6 * n);
  for (int i = 0; i < 6; ++i) std::memcpy(r->state.data() + i * n, in[i], n * sizeof(double));
  r->cls = cls;
  r->Re = Re;
  r->alpha = alpha;
  r->dt = dt;
  r->nsteps = nsteps;
  r->remaining = nsteps;
  r->corr_fixed = corr_fixed;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    r->id = sv->next_id++;
    r->t_submit = sv->now();
    sv->reqs[r->id] = r;
    sv->cls[cls].queue.push_back(r);
    *id = r->id;
  }
  sv->cv.notify_all();
  return OPF_OK;
}

int opf_server_poll(opf_server* sv, long long id, int* done, int* status, double* t_submit, double* t_admit,
                    double* t_done) {
  if (!sv) return fail(OPF_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) return fail(OPF_ERR_ARG, "unknown request");
  const Request& r = *it->second;
  if (done) *done = r.done ? 1 : 0;
  if (status) *status = r.status;
  if (t_submit) *t_submit = r.t_submit;
  if (t_admit) *t_admit = r.t_admit;
  if (t_done) *t_done = r.t_done;
  return OPF_OK;
}

int opf_server_result(opf_server* sv, long long id, double* ur_re, double* ur_im, double* ut_re, double* ut_im,
                      double* uz_re, double* uz_im, double* t) {
  if (!sv) return fail(OPF_ERR_ARG, "null server");
  ReqPtr r;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(OPF_ERR_ARG, "unknown request");
    if (!it->second->done) return fail(OPF_ERR_STATE, "request not finished");
    r = it->second;
    sv->reqs.erase(it);
  }
  if (r->status != OPF_OK) return fail(r->status, "request failed");
  const Class& c = sv->cls[r->cls];
  const size_t n = (size_t)c.N * c.H;
  double* outp[6] = {ur_re, ur_im, ut_re, ut_im, uz_re, uz_im};
  for (int i = 0; i < 6; ++i)
    if (outp[i]) std::memcpy(outp[i], r->state.data() + i * n, n * sizeof(double));
  if (t) *t = r->t;
  return OPF_OK;
}

int opf_server_stats(opf_server* sv, int* queued, int* running, long long* state_steps) {
  if (!sv) return fail(OPF_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->mu);
  int q = 0;
  for (auto& c : sv->cls) q += (int)c.queue.size();
  if (queued) *queued = q;
  if (running) *running = sv->running;
  if (state_steps) *state_steps = sv->state_steps;
  return OPF_OK;
}

}  // extern "C"
