// Continuous-batching server for integration requests of mixed resolution.
//
// A request is a time integration of one state for a given number of steps
// with its own Re, alpha and dt, the unit of work that steporbit
// (newton.f90:743-787) hands to the time stepper, with one corrector per
// step there.
//
// Memory.  All classes draw on one device memory pool of `pool_bytes`.  A
// class's slots come in blocks of block_slots[i]: one context (and HIP
// stream) each, whose device footprint (opf_ctx_bytes_needed) is charged to
// the pool when the block is created and returned when it is destroyed.  A
// queued request that finds no free slot of its class gets a new block if
// the pool has room for it; otherwise idle blocks of other classes are
// released first; otherwise it waits.  A request's own footprint is its
// block's share, block bytes / block_slots.
//
// Scheduling.  One scheduler thread keeps up to `depth` iterations
// (evaluation + step kernel + control) queued on every block's stream,
// each followed by an asynchronous copy of the block's control blocks into
// pinned memory.  Between two iterations it
//   1. admits queued requests into free slots: parameters, factorisation,
//      the state load, vel_imposesym and the curl are enqueued on the
//      block's stream, so the request joins at the next iteration while the
//      other slots keep running (no step boundary to wait for);
//   2. reads the control snapshots the GPU has completed: a slot that
//      reached its step count (or failed) is retired, its state copied to
//      pinned memory behind the iterations already queued, and the slot is
//      free once that copy has landed;
//   3. tops up the iterations of blocks with running slots.
// Slots advance independently (each has its own step counter and corrector
// decisions on the device), so results do not depend on what else is being
// served.  submit / poll / result only touch the request table, under a
// mutex; the contexts are used by the scheduler thread alone.
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
  std::vector<double> state;   // 6 coll planes: initial, then final
  double t = 0;                // simulation time of the final state
  double t_submit = 0, t_admit = -1, t_done = -1;
  int status = 0;              // 0 ok, else an OPF_ERR_* code
  bool done = false;
};
using ReqPtr = std::shared_ptr<Request>;

enum SlotPhase { FREE = 0, RUNNING = 1, FETCHING = 2 };

struct Snapshot {
  opf::SlotCtl* host = nullptr;   // pinned, one control block per slot
  hipEvent_t ev = nullptr;
  long seq = 0;                   // iterations enqueued before it
  bool pending = false;
};

struct Block {
  opf_ctx* ctx = nullptr;
  int cap = 0;
  size_t bytes = 0;
  std::vector<ReqPtr> req;        // request per slot
  std::vector<int> phase;         // SlotPhase per slot
  std::vector<int> ucur;          // velocity buffer role of an idle slot
  std::vector<long> admit_seq;    // snapshots taken after its admission have seq >= this
  std::vector<hipEvent_t> fetched;
  double2* in = nullptr;          // pinned state rows per slot (admission)
  double2* out = nullptr;         // pinned state rows per slot (retirement)
  size_t rows = 0;                // double2 per slot state
  std::vector<Snapshot> ring;
  int ring_head = 0;              // oldest pending snapshot
  long seq = 0;                   // iterations enqueued
  int inflight = 0;
  int running = 0;                // slots RUNNING or FETCHING
};

struct Class {
  int N = 0, K = 0, M = 0, H = 0, block_slots = 0;
  size_t block_bytes = 0;
  std::vector<std::unique_ptr<Block>> blocks;
  std::deque<ReqPtr> queue;
};

thread_local std::string s_err;

int fail(int code, const std::string& m) {
  s_err = m;
  return code;
}

void hipchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

struct opf_server {
  std::vector<Class> cls;
  size_t pool = 0, pool_used = 0;
  int depth = 3;
  std::mutex mu;
  std::condition_variable cv;
  std::unordered_map<long long, ReqPtr> reqs;
  long long next_id = 1;
  long long state_steps = 0;
  int running = 0;                // admitted, not retired
  int nblocks = 0;
  bool stop = false;
  std::string fatal;              // a scheduler failure (all requests then fail)
  std::thread th;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  static void planes(const Class& c, double* base, double* p[6]) {
    const size_t n = (size_t)c.N * c.H;
    for (int i = 0; i < 6; ++i) p[i] = base + i * n;
  }

  // ---- blocks ---------------------------------------------------------------
  Block* new_block(Class& c) {
    auto b = std::make_unique<Block>();
    b->cap = c.block_slots;
    if (opf_ctx_create(&b->ctx, c.N, c.K, c.M, 1, b->cap, nullptr) != OPF_OK)
      throw std::runtime_error(std::string("block: ") + opf_last_error());
    b->bytes = c.block_bytes;
    b->req.assign(b->cap, nullptr);
    b->phase.assign(b->cap, FREE);
    b->ucur.assign(b->cap, 0);        // the role opf_ctx_create gives every slot
    b->admit_seq.assign(b->cap, 0);
    b->fetched.assign(b->cap, nullptr);
    b->rows = opf::state_rows(b->ctx);
    hipchk(hipHostMalloc((void**)&b->in, sizeof(double2) * b->rows * b->cap, hipHostMallocDefault), "pinned in");
    hipchk(hipHostMalloc((void**)&b->out, sizeof(double2) * b->rows * b->cap, hipHostMallocDefault), "pinned out");
    b->ring.resize(depth + 1);
    for (auto& s : b->ring) {
      hipchk(hipHostMalloc((void**)&s.host, sizeof(opf::SlotCtl) * b->cap, hipHostMallocDefault), "pinned ctl");
      hipchk(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "event");
    }
    for (auto& e : b->fetched) hipchk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    Block* raw = b.get();
    c.blocks.push_back(std::move(b));
    std::lock_guard<std::mutex> g(mu);   // pool figures are read by opf_server_pool
    pool_used += c.block_bytes;
    ++nblocks;
    return raw;
  }

  void free_block(Class& c, size_t i) {
    Block* b = c.blocks[i].get();
    if (b->ctx) opf_ctx_destroy(b->ctx);   // waits for its stream
    for (auto& s : b->ring) {
      if (s.host) (void)hipHostFree(s.host);
      if (s.ev) (void)hipEventDestroy(s.ev);
    }
    for (auto e : b->fetched)
      if (e) (void)hipEventDestroy(e);
    if (b->in) (void)hipHostFree(b->in);
    if (b->out) (void)hipHostFree(b->out);
    {
      std::lock_guard<std::mutex> g(mu);
      pool_used -= b->bytes;
      --nblocks;
    }
    c.blocks.erase(c.blocks.begin() + (long)i);
  }

  // release idle blocks of other classes until `need` bytes fit the pool
  bool make_room(size_t need, const Class* keep) {
    for (auto& c : cls) {
      if (&c == keep) continue;
      for (size_t i = c.blocks.size(); i-- > 0 && pool_used + need > pool;)
        if (c.blocks[i]->running == 0) free_block(c, i);
    }
    return pool_used + need <= pool;
  }

  // ---- requests ---------------------------------------------------------------
  void finish(const ReqPtr& r, int status) {
    std::lock_guard<std::mutex> g(mu);
    r->status = status;
    r->done = true;
    r->t_done = now();
    if (status == OPF_OK) state_steps += r->nsteps;
    --running;
  }

  // 1. admission: every queued request that finds (or gets) a free slot
  bool admit_all() {
    bool any = false;
    for (auto& c : cls) {
      for (;;) {
        ReqPtr r;
        {
          std::lock_guard<std::mutex> g(mu);
          if (c.queue.empty()) break;
          r = c.queue.front();
        }
        Block* blk = nullptr;
        int slot = -1;
        for (auto& b : c.blocks) {
          for (int s = 0; s < b->cap && slot < 0; ++s)
            if (b->phase[s] == FREE) slot = s;
          if (slot >= 0) {
            blk = b.get();
            break;
          }
        }
        if (!blk) {
          if (pool_used + c.block_bytes > pool && !make_room(c.block_bytes, &c)) break;   // waits for memory
          blk = new_block(c);
          slot = 0;
        }
        {
          std::lock_guard<std::mutex> g(mu);
          c.queue.pop_front();
          r->t_admit = now();
        }
        double* p[6];
        planes(c, r->state.data(), p);
        double2* rows = blk->in + (size_t)slot * blk->rows;
        opf::coll_to_rows(blk->ctx, p, rows);
        try {
          opf::admit_async(blk->ctx, slot, r->Re, r->alpha, r->dt, blk->ucur[slot], rows, r->nsteps,
                           r->corr_fixed);
        } catch (const std::exception& e) {
          finish(r, OPF_ERR_ARG);
          continue;
        }
        blk->req[slot] = r;
        blk->phase[slot] = RUNNING;
        blk->admit_seq[slot] = blk->seq;
        ++blk->running;
        any = true;
      }
    }
    return any;
  }

  // 2. completed snapshots and fetches of one block
  bool harvest(Class& c, Block& b) {
    bool any = false;
    while (b.inflight > 0) {
      Snapshot& sn = b.ring[b.ring_head];
      const hipError_t q = hipEventQuery(sn.ev);
      if (q == hipErrorNotReady) break;
      hipchk(q, "snapshot event");
      sn.pending = false;
      b.ring_head = (b.ring_head + 1) % (int)b.ring.size();
      --b.inflight;
      any = true;
      for (int s = 0; s < b.cap; ++s) {
        if (b.phase[s] != RUNNING || sn.seq <= b.admit_seq[s]) continue;
        const opf::SlotCtl& k = sn.host[s];
        if (k.status != 0) {               // singular factor or corrector failure
          b.ucur[s] = k.ucur;
          b.phase[s] = FREE;
          --b.running;
          finish(b.req[s], k.status >= 100 ? OPF_ERR_SINGULAR : OPF_ERR_CORRECTOR);
          b.req[s] = nullptr;
        } else if (!k.active && k.step >= k.target_steps) {
          b.ucur[s] = k.ucur;
          b.req[s]->t = k.t;
          opf::fetch_async(b.ctx, s, k.ucur, b.out + (size_t)s * b.rows);
          hipchk(hipEventRecord(b.fetched[s], opf::ctx_stream(b.ctx)), "fetch event");
          b.phase[s] = FETCHING;
        }
      }
    }
    for (int s = 0; s < b.cap; ++s) {
      if (b.phase[s] != FETCHING) continue;
      const hipError_t q = hipEventQuery(b.fetched[s]);
      if (q == hipErrorNotReady) continue;
      hipchk(q, "fetch event");
      ReqPtr r = b.req[s];
      double* p[6];
      planes(c, r->state.data(), p);
      opf::rows_to_coll(b.ctx, b.out + (size_t)s * b.rows, p);
      b.req[s] = nullptr;
      b.phase[s] = FREE;
      --b.running;
      finish(r, OPF_OK);
      any = true;
    }
    return any;
  }

  // 3. keep `depth` iterations queued on a block with running slots
  bool top_up(Block& b) {
    bool any = false;
    int lo = b.cap, hi = -1;
    for (int s = 0; s < b.cap; ++s)
      if (b.phase[s] == RUNNING) {
        lo = std::min(lo, s);
        hi = std::max(hi, s);
      }
    if (hi < 0) return false;
    while (b.inflight < depth) {
      const int tail = (b.ring_head + b.inflight) % (int)b.ring.size();
      Snapshot& sn = b.ring[tail];
      opf::enqueue_iteration_range(b.ctx, lo, hi - lo + 1);
      ++b.seq;
      opf::ctl_snapshot_async(b.ctx, sn.host);
      hipchk(hipEventRecord(sn.ev, opf::ctx_stream(b.ctx)), "snapshot event");
      sn.seq = b.seq;
      sn.pending = true;
      ++b.inflight;
      any = true;
    }
    return any;
  }

  void fail_all(const std::string& why) {
    std::lock_guard<std::mutex> g(mu);
    fatal = why;
    for (auto& [id, r] : reqs)
      if (!r->done) {
        r->done = true;
        r->status = OPF_ERR_HIP;
        r->t_done = now();
      }
    for (auto& c : cls) c.queue.clear();
    running = 0;
  }

  void loop() {
    try {
      for (;;) {
        {
          std::unique_lock<std::mutex> g(mu);
          if (stop) return;
          bool queued = false;
          for (auto& c : cls) queued = queued || !c.queue.empty();
          if (!queued && running == 0) {
            cv.wait(g, [&] {
              if (stop) return true;
              for (auto& c : cls)
                if (!c.queue.empty()) return true;
              return false;
            });
            if (stop) return;
          }
        }
        bool any = admit_all();
        for (auto& c : cls)
          for (auto& b : c.blocks) {
            any = harvest(c, *b) || any;
            any = top_up(*b) || any;
          }
        if (!any) std::this_thread::sleep_for(std::chrono::microseconds(50));
      }
    } catch (const std::exception& e) {
      fail_all(e.what());
    }
  }
};

extern "C" {

const char* opf_server_last_error(void) { return s_err.c_str(); }

int opf_server_create(opf_server** out, int nclass, const int* nkm, const int* block_slots, size_t pool_bytes,
                      int depth) {
  if (!out || nclass < 1 || !nkm || !block_slots || depth < 1 || depth > 8)
    return fail(OPF_ERR_ARG, "bad server arguments");
  *out = nullptr;
  auto* sv = new opf_server();
  sv->pool = pool_bytes;
  sv->depth = depth;
  sv->cls = std::vector<Class>(nclass);   // Class holds move-only blocks: built in place
  for (int i = 0; i < nclass; ++i) {
    Class& c = sv->cls[i];
    c.N = nkm[3 * i];
    c.K = nkm[3 * i + 1];
    c.M = nkm[3 * i + 2];
    c.block_slots = block_slots[i];
    if (c.block_slots < 1 || c.N < 9 || c.K < 2 || c.M < 1) {
      delete sv;
      return fail(OPF_ERR_ARG, "class " + std::to_string(i) + ": bad shape or block size");
    }
    c.H = (2 * c.K - 1) * c.M - (c.K - 1);
    c.block_bytes = opf_ctx_bytes_needed(c.N, c.K, c.M, c.block_slots);
    if (c.block_bytes > pool_bytes) {
      delete sv;
      return fail(OPF_ERR_NOMEM, "class " + std::to_string(i) + ": one block exceeds the pool");
    }
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
    while (!c.blocks.empty()) sv->free_block(c, c.blocks.size() - 1);
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
  r->corr_fixed = corr_fixed;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    if (!sv->fatal.empty()) return fail(OPF_ERR_HIP, "server stopped: " + sv->fatal);
    r->id = sv->next_id++;
    r->t_submit = sv->now();
    sv->reqs[r->id] = r;
    sv->cls[cls].queue.push_back(r);
    ++sv->running;
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
  if (running) *running = sv->running - q;
  if (state_steps) *state_steps = sv->state_steps;
  return OPF_OK;
}

int opf_server_pool(opf_server* sv, size_t* used, size_t* limit, int* blocks) {
  if (!sv) return fail(OPF_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->mu);
  if (used) *used = sv->pool_used;
  if (limit) *limit = sv->pool;
  if (blocks) *blocks = sv->nblocks;
  return OPF_OK;
}

}  // extern "C"
