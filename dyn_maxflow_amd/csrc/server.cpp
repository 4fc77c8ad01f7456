// Streaming update server: residual graphs stay resident in HBM and update
// batches are solved as they arrive.  No reference counterpart: the reference
// drivers load one file of updates, solve it once and exit.
//
// Structure (all host threads, one device):
//   * sessions: one resident graph each (a dmf_engine with its own stream),
//     admitted only if its device footprint fits the HBM budget;
//   * one stager thread takes submitted batches in arrival order and admits
//     each into a device staging arena when its bytes fit (first fit over a
//     free list); a batch that does not fit waits until solves free space;
//   * one solver thread per session applies that session's staged batches in
//     submission order.  Solves of different sessions run concurrently on
//     their own streams, and staging copies overlap with solves.
// Every result carries its timestamps on the server clock (ms since create).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/dmf.h"
#include "abi_guard.hpp"

namespace {

struct Request {
  uint64_t ticket = 0;
  int32_t session = 0;
  std::vector<int32_t> src, dst, w;  // host copy until staged
  int64_t off = -1;                  // arena offset once staged
  int64_t bytes = 0;
  double submit_ms = 0, admit_ms = 0;
};

struct Session {
  int32_t id = 0;
  dmf_engine* eng = nullptr;
  int32_t max_batch = 0;
  dmf_params params{};  // server parameters with this graph's default kc filled in
  int64_t bytes = 0;  // device bytes held (engine incl. update buffers)
  std::deque<std::unique_ptr<Request>> ready;  // staged, in submission order
  std::thread solver;
  bool closing = false;
  int32_t pending = 0;  // submitted and not yet finished
};

// First-fit allocator over [0, size) with coalescing of freed ranges.
class Arena {
 public:
  void init(int64_t size) {
    free_.clear();
    if (size > 0) free_[0] = size;
    size_ = size;
  }
  int64_t alloc(int64_t bytes) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const int64_t off = it->first, left = it->second - bytes;
        free_.erase(it);
        if (left > 0) free_[off + bytes] = left;
        used_ += bytes;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t bytes) {
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t size() const { return size_; }
  int64_t used() const { return used_; }

 private:
  std::map<int64_t, int64_t> free_;
  int64_t size_ = 0, used_ = 0;
};

int64_t align256(int64_t x) { return (x + 255) & ~int64_t{255}; }

}  // namespace

struct dmf_server {
  dmf_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_stage;   // stager: new submission or freed arena space
  std::condition_variable cv_ready;   // solvers: staged work or shutdown
  std::condition_variable cv_done;    // waiters: a result was posted
  std::deque<std::unique_ptr<Request>> inbox;  // submitted, not yet staged
  std::unordered_map<int32_t, std::unique_ptr<Session>> sessions;
  std::unordered_map<uint64_t, dmf_result> results;
  std::unordered_set<uint64_t> outstanding;  // submitted and not yet returned by wait
  Arena arena;
  char* d_arena = nullptr;
  hipStream_t copy_stream = nullptr;
  int64_t held = 0;  // bytes of the budget in use (sessions + arena)
  int64_t peak_staged = 0;
  uint64_t next_ticket = 1;
  int32_t next_session = 1;
  uint64_t done_count = 0;
  bool stopping = false;
  std::thread stager;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void post_result(dmf_server* S, const Request& r, int status, int64_t flow, double start_ms,
                 const dmf_counters* c) {
  dmf_result res{};
  res.ticket = r.ticket;
  res.session = r.session;
  res.status = status;
  res.flow = flow;
  res.k = static_cast<int32_t>(r.src.size());
  res.submit_ms = r.submit_ms;
  res.admit_ms = r.admit_ms;
  res.start_ms = start_ms;
  res.done_ms = S->now_ms();
  if (c) {
    res.device_ms = c->device_ms;
    res.arc_visits = c->arc_visits;
    res.rounds = c->rounds;
    res.global_relabels = c->global_relabels;
  }
  S->results[r.ticket] = res;
  ++S->done_count;
}

void stager_loop(dmf_server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_stage.wait(lk, [&] {
      if (S->stopping && S->inbox.empty()) return true;
      if (S->inbox.empty()) return false;
      return S->arena.used() + S->inbox.front()->bytes <= S->arena.size();
    });
    if (S->inbox.empty()) return;  // stopping
    std::unique_ptr<Request> r = std::move(S->inbox.front());
    S->inbox.pop_front();
    r->off = S->arena.alloc(r->bytes);
    while (r->off < 0) {  // fragmented: wait for a solve to free space
      S->cv_stage.wait(lk);
      r->off = S->arena.alloc(r->bytes);
    }
    S->peak_staged = std::max(S->peak_staged, S->arena.used());
    lk.unlock();
    // copy the three arrays into the arena (this thread only)
    const size_t k = r->src.size();
    const int64_t part = align256(static_cast<int64_t>(k) * 4);
    char* base = S->d_arena + r->off;
    bool ok = true;
    if (k > 0) {
      ok = hipMemcpyAsync(base, r->src.data(), k * 4, hipMemcpyHostToDevice, S->copy_stream) == hipSuccess &&
           hipMemcpyAsync(base + part, r->dst.data(), k * 4, hipMemcpyHostToDevice, S->copy_stream) == hipSuccess &&
           hipMemcpyAsync(base + 2 * part, r->w.data(), k * 4, hipMemcpyHostToDevice, S->copy_stream) == hipSuccess &&
           hipStreamSynchronize(S->copy_stream) == hipSuccess;
    }
    lk.lock();
    r->admit_ms = S->now_ms();
    auto it = S->sessions.find(r->session);
    if (!ok || it == S->sessions.end()) {
      S->arena.release(r->off, r->bytes);
      post_result(S, *r, ok ? DMF_EINVAL : DMF_EDEVICE, 0, r->admit_ms, nullptr);
      if (it != S->sessions.end()) --it->second->pending;
      S->cv_done.notify_all();
      S->cv_stage.notify_all();
      continue;
    }
    it->second->ready.push_back(std::move(r));
    S->cv_ready.notify_all();
  }
}

void solver_loop(dmf_server* S, Session* ss) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // a closing session still finishes what the stager has not handed over
    S->cv_ready.wait(lk, [&] { return !ss->ready.empty() || (ss->closing && ss->pending == 0); });
    if (ss->ready.empty()) return;
    std::unique_ptr<Request> r = std::move(ss->ready.front());
    ss->ready.pop_front();
    lk.unlock();
    const double start = S->now_ms();
    const int32_t k = static_cast<int32_t>(r->src.size());
    const int64_t part = align256(static_cast<int64_t>(k) * 4);
    const char* base = S->d_arena + r->off;
    int64_t flow = 0;
    const int rc = dmf_engine_apply_batch_device(ss->eng, k, reinterpret_cast<const int32_t*>(base),
                                                 reinterpret_cast<const int32_t*>(base + part),
                                                 reinterpret_cast<const int32_t*>(base + 2 * part), &ss->params,
                                                 &flow);
    dmf_counters c{};
    dmf_engine_counters(ss->eng, &c);
    lk.lock();
    S->arena.release(r->off, r->bytes);
    post_result(S, *r, rc, flow, start, &c);
    --ss->pending;
    S->cv_done.notify_all();
    S->cv_stage.notify_all();
  }
}

}  // namespace

extern "C" {

void dmf_server_config_default(dmf_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->hbm_budget = 0;
  c->staging_bytes = 0;
  c->max_sessions = 64;
  c->params.variant = DMF_VARIANT_DATA;
  c->params.kc = 0;
  c->params.relabel_interval = 1;
  c->params.scheduler = DMF_SCHED_LEARNED;
  c->params.max_rounds = 0;
  c->params.mode = 2;
}

int dmf_server_create(const dmf_server_config* cfg, dmf_server** out) {
  return guarded([&]() -> int {
    if (!out) return DMF_EINVAL;
    *out = nullptr;
    if (dmf_device_count() < 1) return DMF_EDEVICE;
    auto S = std::make_unique<dmf_server>();
    if (cfg) S->cfg = *cfg;
    else dmf_server_config_default(&S->cfg);
    if (S->cfg.max_sessions <= 0) S->cfg.max_sessions = 64;
    if (S->cfg.hbm_budget < 0 || S->cfg.staging_bytes < 0) return DMF_EINVAL;
    if (hipGetDevice(&S->device) != hipSuccess) return DMF_EDEVICE;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return DMF_EDEVICE;
    if (S->cfg.hbm_budget == 0) S->cfg.hbm_budget = static_cast<int64_t>(free_b / 10 * 9);
    if (S->cfg.staging_bytes == 0) S->cfg.staging_bytes = std::max<int64_t>(S->cfg.hbm_budget / 8, 1 << 20);
    S->cfg.staging_bytes = align256(S->cfg.staging_bytes);
    if (S->cfg.staging_bytes > S->cfg.hbm_budget) return DMF_EADMIT;
    if (hipMalloc(reinterpret_cast<void**>(&S->d_arena), S->cfg.staging_bytes) != hipSuccess) return DMF_ENOMEM;
    if (hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipFree(S->d_arena);
      return DMF_EDEVICE;
    }
    S->arena.init(S->cfg.staging_bytes);
    S->held = S->cfg.staging_bytes;
    S->t0 = std::chrono::steady_clock::now();
    try {
      S->stager = std::thread(stager_loop, S.get());
    } catch (...) {  // no thread available; nothing may escape the C-ABI
      (void)hipStreamDestroy(S->copy_stream);
      (void)hipFree(S->d_arena);
      return DMF_ENOMEM;
    }
    *out = S.release();
    return DMF_OK;
  });
}

int dmf_server_open(dmf_server* S, const dmf_csr* g, int32_t s, int32_t t, int32_t max_batch,
                    int32_t* session, int64_t* static_flow) {
  return guarded([&]() -> int {
    if (!S || !g || !session || max_batch < 1) return DMF_EINVAL;
    const int64_t need = dmf_engine_footprint(g->n, g->m) + static_cast<int64_t>(max_batch) * 24;  // update slots
    auto ss = std::make_unique<Session>();  // before the reservation, which nothing below throws past
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (static_cast<int32_t>(S->sessions.size()) >= S->cfg.max_sessions) return DMF_EADMIT;
      if (S->held + need > S->cfg.hbm_budget) return DMF_EADMIT;
      S->held += need;  // reserved while the engine is built
    }
    ss->params = S->cfg.params;
    if (ss->params.kc <= 0) {
      dmf_params d;
      dmf_params_default(g, &d);
      ss->params.kc = d.kc;
    }
    int rc = dmf_engine_create(g, s, t, &ss->eng);
    if (!rc) rc = dmf_engine_reserve_updates(ss->eng, max_batch);
    int64_t f = 0;
    if (!rc) {
      dmf_params p = ss->params;
      p.scheduler = DMF_SCHED_FIXED;  // the static solve has no batch to schedule
      p.variant = DMF_VARIANT_PP;     // a push-pull static state lets every variant take the batches
      rc = dmf_engine_static_solve(ss->eng, &p, &f);
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->held -= need;
    if (rc) {
      if (ss->eng) dmf_engine_destroy(ss->eng);
      return rc;
    }
    ss->bytes = dmf_engine_device_bytes(ss->eng);
    S->held += ss->bytes;
    ss->max_batch = max_batch;
    ss->id = S->next_session++;
    Session* raw = ss.get();
    try {
      S->sessions[raw->id] = std::move(ss);
      raw->solver = std::thread(solver_loop, S, raw);
    } catch (...) {  // no memory or thread: undo the session (ss frees it if never inserted)
      S->held -= raw->bytes;
      dmf_engine_destroy(raw->eng);
      S->sessions.erase(raw->id);
      return DMF_ENOMEM;
    }
    *session = raw->id;
    if (static_flow) *static_flow = f;
    return DMF_OK;
  });
}

int dmf_server_submit(dmf_server* S, int32_t session, int32_t k, const int32_t* src, const int32_t* dst,
                      const int32_t* w, uint64_t* ticket) {
  return guarded([&]() -> int {
    if (!S || !ticket || k < 0 || (k > 0 && (!src || !dst || !w))) return DMF_EINVAL;
    auto r = std::make_unique<Request>();
    r->src.assign(src, src + k);
    r->dst.assign(dst, dst + k);
    r->w.assign(w, w + k);
    r->bytes = std::max<int64_t>(256, 3 * align256(static_cast<int64_t>(k) * 4));
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stopping) return DMF_EINVAL;
    auto it = S->sessions.find(session);
    if (it == S->sessions.end() || it->second->closing) return DMF_EINVAL;
    if (k > it->second->max_batch || r->bytes > S->arena.size()) return DMF_EADMIT;
    r->session = session;
    r->ticket = S->next_ticket++;
    r->submit_ms = S->now_ms();
    *ticket = r->ticket;
    S->outstanding.insert(r->ticket);
    ++it->second->pending;
    S->inbox.push_back(std::move(r));
    S->cv_stage.notify_all();
    return DMF_OK;
  });
}

int dmf_server_wait(dmf_server* S, uint64_t ticket, double timeout_ms, dmf_result* out) {
  return guarded([&]() -> int {
    if (!S || !out || ticket == 0) return DMF_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    if (!S->outstanding.count(ticket)) return DMF_EINVAL;  // unknown or already returned
    auto ready = [&] { return S->results.count(ticket) != 0; };
    if (timeout_ms < 0) {
      S->cv_done.wait(lk, ready);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
      return DMF_EAGAIN;
    }
    *out = S->results[ticket];
    S->results.erase(ticket);
    S->outstanding.erase(ticket);
    return DMF_OK;
  });
}

int dmf_server_get_stats(dmf_server* S, dmf_server_stats* st) {
  return guarded([&]() -> int {
    if (!S || !st) return DMF_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof(*st));
    st->sessions = static_cast<int32_t>(S->sessions.size());
    for (auto& kv : S->sessions) st->in_flight += kv.second->pending;
    st->queued = static_cast<int32_t>(S->inbox.size());
    st->completed = S->done_count;
    st->hbm_budget = S->cfg.hbm_budget;
    st->hbm_held = S->held;
    st->staged_bytes = S->arena.used();
    st->peak_staged_bytes = S->peak_staged;
    st->clock_ms = S->now_ms();
    return DMF_OK;
  });
}

int dmf_server_close(dmf_server* S, int32_t session) {
  return guarded([&]() -> int {
    if (!S) return DMF_EINVAL;
    std::unique_ptr<Session> ss;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto it = S->sessions.find(session);
      if (it == S->sessions.end()) return DMF_EINVAL;
      Session* raw = it->second.get();
      raw->closing = true;  // no new submissions; queued ones still run
      S->cv_done.wait(lk, [&] { return raw->pending == 0; });
      S->cv_ready.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      ss = std::move(S->sessions[session]);
    }
    if (ss->solver.joinable()) ss->solver.join();
    dmf_engine_destroy(ss->eng);
    std::lock_guard<std::mutex> lk(S->mu);
    S->held -= ss->bytes;
    S->sessions.erase(session);
    return DMF_OK;
  });
}

int dmf_server_destroy(dmf_server* S) {
  return guarded([&]() -> int {
    if (!S) return DMF_EINVAL;
    std::vector<int32_t> ids;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (auto& kv : S->sessions) ids.push_back(kv.first);
    }
    for (int32_t id : ids) dmf_server_close(S, id);  // finishes queued batches first
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stopping = true;
      S->cv_stage.notify_all();
    }
    if (S->stager.joinable()) S->stager.join();
    (void)hipStreamDestroy(S->copy_stream);
    (void)hipFree(S->d_arena);
    delete S;
    return DMF_OK;
  });
}

}  // extern "C"
