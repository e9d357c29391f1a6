// gu_serve.hpp -- continuous-batching query server over resident graphs.
//
// Queries (BFS with or without path counts, sampled Brandes BC, diameter
// sampling; 1..N sources each) arrive at any time.  The server
//  - admits them in arrival order while their outputs fit an HBM budget
//    (an arena carved once at start-up; a query holds its block from
//    admission until the client releases it),
//  - packs the undispatched sources of admitted queries of one graph and
//    kind into shared multi-source launches (a launch may carry sources of
//    many queries, a query may span many launches),
//  - keeps `slots` launches in flight at once, one host worker and one
//    stream per slot, and admits / packs newly arrived work whenever a slot
//    or memory frees up.
// Results land in each query's own arena block (distance / path-count /
// dependency rows, a BC vector) or, for diameter sampling, in a host int, so
// a query's result does not depend on which other queries shared its
// launches: distances, path counts and eccentricities are computed per
// source; BC sums a query's sources in source order within a launch and
// launch by launch (FP64), so only the grouping of that FP64 sum can differ.
//
// A query may ask for its results on the host: it then also holds a block
// of a second arena in pinned host memory.  A launch holding such queries
// goes, once computed, to one delivery thread, which copies its rows to
// their host blocks (and a BC vector after the query's last launch) on one
// stream while every slot keeps computing; the query is reported done only
// when its results are in host memory.  Its device block is released as
// soon as it is done.
//
// This header holds the scheduling logic only; the executor that runs a
// launch is an interface, so the same logic runs over the HIP engine in the
// library and over a CPU stand-in in the scheduler's unit test.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace gus {

// Query kinds as launched: BFS with path counts, BFS distances only,
// Brandes BC, diameter sampling.
enum Cls { C_BFS_SIGMA = 0, C_BFS = 1, C_BC = 2, C_DIAM = 3, kClasses = 4 };

constexpr int kMaxGraphs = 16;
constexpr int kOk = 0, kErrInvalid = -1, kErrCapacity = -9, kErrTimeout = -10;

// First-fit offset allocator over [0, cap) in 256-byte units.
class Arena {
 public:
  explicit Arena(size_t cap) : cap_(cap) {
    if (cap_) free_[0] = cap_;
  }
  static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~size_t(255); }
  size_t cap() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  bool alloc(size_t bytes, size_t *off) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, o = it->first + bytes;
      free_.erase(it);
      if (rest) free_[o] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes);
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
  size_t free_blocks() const { return free_.size(); }

 private:
  size_t cap_, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
};

// Runs one launch.  Row arrays hold one entry per source of the launch.
struct Exec {
  virtual ~Exec() {}
  virtual char *arena_base() = 0;
  // Called first on every server thread: the slot workers with their slot,
  // the delivery thread with the slot count.
  virtual void thread_init(int slot) { (void)slot; }
  // wrap[k] (host) receives per source 1 when one of its path counts
  // wrapped modulo 2^64
  virtual int search(int slot, int graph, const int32_t *src, int k, int32_t *const *drows,
                     uint64_t *const *srows, int32_t *ecc, int32_t *wrap) = 0;
  virtual int bc(int slot, int graph, const int32_t *src, int k, float *const *xrows,
                 double *const *targets, int32_t *wrap) = 0;
  virtual int zero(int slot, void *p, size_t bytes) = 0;
  // Host arena (pinned) and a device-to-host copy on the slot's stream that
  // returns once the bytes are in host memory.
  virtual char *host_base() { return nullptr; }
  virtual int to_host(int slot, void *dst, const void *src, size_t bytes) = 0;
  // The same copy on the one stream of the server's delivery thread.  All
  // host deliveries go through it, one after another: on MI355X one
  // stream's device-to-host copies run at 57 GB/s, four streams' at once at
  // 49 GB/s together.
  virtual int deliver(void *dst, const void *src, size_t bytes) { return to_host(0, dst, src, bytes); }
};

struct Spec {
  int graph = 0;
  int cls = C_BFS_SIGMA;
  int nsrc = 0;
  const int32_t *sources = nullptr;
  bool want_delta = false;  // BC: keep the per-source dependency rows
  bool to_host = false;     // deliver the results to the host arena
};

constexpr int kFlagWrapped = 1;  // a path count of one of the query's sources wrapped

struct Result {
  int status = 0, cls = 0, nsrc = 0, diameter = 0, launches = 0, flags = 0;
  int32_t *dist = nullptr;
  uint64_t *sigma = nullptr;
  float *delta = nullptr;
  double *bc = nullptr;
  // host copies of a query submitted with to_host (the device pointers
  // above are then null: the device block is already released)
  int32_t *h_dist = nullptr;
  uint64_t *h_sigma = nullptr;
  float *h_delta = nullptr;
  double *h_bc = nullptr;
  double submit_ms = 0, admit_ms = 0, first_launch_ms = 0, done_ms = 0;
  size_t bytes = 0;
};

struct Stats {
  int64_t submitted = 0, completed = 0, rejected = 0, launches = 0, sources = 0;
  int64_t hbm_in_use = 0, hbm_peak = 0, hbm_budget = 0;
  int32_t waiting = 0, active = 0;
  int64_t host_in_use = 0, host_peak = 0, host_budget = 0;
};

// Arena bytes a query holds: its output rows and BC vector.
inline size_t query_bytes(int cls, int64_t k, int64_t n, bool want_delta) {
  switch (cls) {
    case C_BFS_SIGMA: return Arena::round((size_t)(k * n * 4)) + Arena::round((size_t)(k * n * 8));
    case C_BFS: return Arena::round((size_t)(k * n * 4));
    case C_BC: return Arena::round((size_t)(n * 8)) + (want_delta ? Arena::round((size_t)(k * n * 4)) : 0);
    default: return Arena::round(1);
  }
}

class Server {
 public:
  Server(Exec *ex, size_t budget, int slots, int max_pack, size_t host_budget = 0)
      : ex_(ex), arena_(budget), host_arena_(host_budget), slots_(std::max(1, slots)),
        max_pack_(std::max(1, max_pack)), t0_(std::chrono::steady_clock::now()) {
    delivering_ = host_budget > 0;
    if (delivering_) deliverer_ = std::thread([this] { deliverer(); });
    for (int i = 0; i < slots_; i++) workers_.emplace_back([this, i] { worker(i); });
  }
  ~Server() { stop(); }

  // Stops the workers after their current launches, then the delivery
  // thread once it has made every delivery handed to it.
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stop_) return;
      stop_ = true;
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    cv_deliver_.notify_all();
    for (auto &t : workers_) t.join();
    {
      std::lock_guard<std::mutex> lk(mu_);
      workers_done_ = true;
    }
    cv_deliver_.notify_all();
    if (deliverer_.joinable()) deliverer_.join();
  }

  int add_graph(int64_t n, int *id) {
    std::lock_guard<std::mutex> lk(mu_);
    if (ngraphs_ >= kMaxGraphs || n < 0) return kErrInvalid;
    n_[ngraphs_] = n;
    *id = ngraphs_++;
    return kOk;
  }

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
  }

  int submit(const Spec &s, int64_t *ticket) {
    if (s.cls < 0 || s.cls >= kClasses || s.nsrc <= 0 || !s.sources || !ticket) return kErrInvalid;
    std::unique_ptr<Q> q(new Q);
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (s.graph < 0 || s.graph >= ngraphs_) return kErrInvalid;
      const int64_t n = n_[s.graph];
      for (int i = 0; i < s.nsrc; i++)
        if (s.sources[i] < 0 || s.sources[i] >= n) return kErrInvalid;
      q->graph = s.graph;
      q->cls = s.cls;
      q->nsrc = s.nsrc;
      q->want_delta = s.want_delta;
      q->to_host = s.to_host;
      q->src.assign(s.sources, s.sources + s.nsrc);
      q->n = n;
      q->bytes = query_bytes(s.cls, s.nsrc, n, s.want_delta);
      q->ticket = next_ticket_++;
      q->t_submit = now_ms();
      *ticket = q->ticket;
      stats_.submitted++;
      Q *raw = q.get();
      queries_[q->ticket] = std::move(q);
      if (raw->bytes > arena_.cap() || (raw->to_host && raw->bytes > host_arena_.cap())) {  // never fits
        raw->status = kErrCapacity;
        finish(raw);
        stats_.rejected++;
      } else {
        waiting_.push_back(raw);
        admit();
      }
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    return kOk;
  }

  // Blocks until the query is done (timeout_ms < 0: no limit).  The ticket
  // is looked up again at every wake-up: another thread may release the
  // query meanwhile (then kErrInvalid).
  int wait(int64_t ticket, int timeout_ms, Result *r) {
    std::unique_lock<std::mutex> lk(mu_);
    Q *q = nullptr;
    auto pred = [&] {
      auto it = queries_.find(ticket);
      q = it == queries_.end() ? nullptr : it->second.get();
      return !q || q->done || stop_;
    };
    if (timeout_ms < 0)
      cv_done_.wait(lk, pred);
    else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
      return kErrTimeout;
    if (!q || !q->done) return kErrInvalid;
    fill(q, r);
    q->reported = true;
    return kOk;
  }

  // Next finished query not yet returned by wait / next_done.
  int next_done(int timeout_ms, int64_t *ticket, Result *r) {
    std::unique_lock<std::mutex> lk(mu_);
    auto pred = [&] { return !done_fifo_.empty() || stop_; };
    while (true) {
      if (timeout_ms < 0)
        cv_done_.wait(lk, pred);
      else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return kErrTimeout;
      if (done_fifo_.empty()) return kErrInvalid;  // stopped
      int64_t t = done_fifo_.front();
      done_fifo_.pop_front();
      auto it = queries_.find(t);
      if (it == queries_.end() || it->second->reported) continue;
      it->second->reported = true;
      *ticket = t;
      fill(it->second.get(), r);
      return kOk;
    }
  }

  int poll(int64_t ticket, int *done) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = queries_.find(ticket);
    if (it == queries_.end()) return kErrInvalid;
    *done = it->second->done ? 1 : 0;
    return kOk;
  }

  // Frees a finished query's arena block; waiting queries may be admitted.
  int release(int64_t ticket) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = queries_.find(ticket);
      if (it == queries_.end() || !it->second->done) return kErrInvalid;
      Q *q = it->second.get();
      release_blocks(q);
      queries_.erase(it);
      admit();
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    return kOk;
  }

  Stats stats() {
    std::lock_guard<std::mutex> lk(mu_);
    Stats s = stats_;
    s.hbm_in_use = (int64_t)arena_.used();
    s.hbm_peak = (int64_t)arena_.peak();
    s.hbm_budget = (int64_t)arena_.cap();
    s.waiting = (int32_t)waiting_.size();
    s.host_in_use = (int64_t)host_arena_.used();
    s.host_peak = (int64_t)host_arena_.peak();
    s.host_budget = (int64_t)host_arena_.cap();
    int a = 0;
    for (auto &kv : queries_) a += kv.second->admitted && !kv.second->done;
    s.active = a;
    return s;
  }

 private:
  struct Q {
    int64_t ticket = 0;
    int graph = 0, cls = 0, nsrc = 0;
    int64_t n = 0;
    bool want_delta = false, to_host = false;
    std::vector<int32_t> src;
    size_t bytes = 0, off = 0, host_off = 0;
    bool dev_held = false, host_held = false;
    int32_t *dist = nullptr;
    uint64_t *sigma = nullptr;
    float *delta = nullptr;
    double *bc = nullptr;
    int32_t *h_dist = nullptr;
    uint64_t *h_sigma = nullptr;
    float *h_delta = nullptr;
    double *h_bc = nullptr;
    bool bc_zeroed = false;
    int diameter = 0;
    int next = 0, done_src = 0, inflight = 0, launches = 0, status = 0, flags = 0;
    uint64_t seq = 0;
    bool admitted = false, done = false, reported = false;
    double t_submit = 0, t_admit = -1, t_first = -1, t_done = -1;
  };
  struct Item {
    Q *q;
    int first, count;
  };
  struct Pack {
    int graph = 0, cls = 0, nsrc = 0;
    std::vector<Item> items;
  };
  // A finished launch whose rows some of its queries want on the host: the
  // delivery thread copies them, then completes the launch's items.
  struct Delivery {
    std::vector<Item> items;
    int cls = 0, rc = 0;
    size_t n = 0;
    std::vector<int32_t> ecc, wrap;
  };

  void fill(const Q *q, Result *r) const {
    r->status = q->status;
    r->cls = q->cls;
    r->nsrc = q->nsrc;
    r->diameter = q->diameter;
    r->launches = q->launches;
    r->flags = q->flags;
    r->dist = q->dist;
    r->sigma = q->sigma;
    r->delta = q->delta;
    r->bc = q->bc;
    r->h_dist = q->h_dist;
    r->h_sigma = q->h_sigma;
    r->h_delta = q->h_delta;
    r->h_bc = q->h_bc;
    r->submit_ms = q->t_submit;
    r->admit_ms = q->t_admit;
    r->first_launch_ms = q->t_first;
    r->done_ms = q->t_done;
    r->bytes = q->bytes;
  }

  // Caller holds mu_.
  void release_blocks(Q *q) {
    if (q->dev_held) arena_.release(q->off, q->bytes);
    if (q->host_held) host_arena_.release(q->host_off, q->bytes);
    q->dev_held = q->host_held = false;
  }

  // Caller holds mu_.  A query delivered to the host gives its device block
  // back at once (its device pointers become null).
  void finish(Q *q) {
    q->done = true;
    q->t_done = now_ms();
    done_fifo_.push_back(q->ticket);
    if (q->status == 0) stats_.completed++;
    if (q->to_host && q->dev_held) {
      arena_.release(q->off, q->bytes);
      q->dev_held = false;
      q->dist = nullptr;
      q->sigma = nullptr;
      q->delta = nullptr;
      q->bc = nullptr;
      admit();
    }
  }

  // FIFO admission under the budget (caller holds mu_).
  void admit() {
    while (!waiting_.empty()) {
      Q *q = waiting_.front();
      size_t off = 0, hoff = 0;
      if (!arena_.alloc(q->bytes, &off)) break;
      if (q->to_host && !host_arena_.alloc(q->bytes, &hoff)) {
        arena_.release(off, q->bytes);
        break;
      }
      waiting_.pop_front();
      q->off = off;
      q->host_off = hoff;
      q->dev_held = true;
      q->host_held = q->to_host;
      q->admitted = true;
      q->t_admit = now_ms();
      q->seq = seq_++;
      char *b = ex_->arena_base() + off;
      char *hb = q->to_host ? ex_->host_base() + hoff : nullptr;
      const size_t n = (size_t)q->n, k = (size_t)q->nsrc;
      // the host block has the device block's layout
      switch (q->cls) {
        case C_BFS_SIGMA:
          q->dist = (int32_t *)b;
          q->sigma = (uint64_t *)(b + Arena::round(k * n * 4));
          if (hb) {
            q->h_dist = (int32_t *)hb;
            q->h_sigma = (uint64_t *)(hb + Arena::round(k * n * 4));
          }
          break;
        case C_BFS:
          q->dist = (int32_t *)b;
          if (hb) q->h_dist = (int32_t *)hb;
          break;
        case C_BC:
          q->bc = (double *)b;
          if (q->want_delta) q->delta = (float *)(b + Arena::round(n * 8));
          if (hb) {
            q->h_bc = (double *)hb;
            if (q->want_delta) q->h_delta = (float *)(hb + Arena::round(n * 8));
          }
          break;
        default: break;
      }
      active_[q->graph * kClasses + q->cls].push_back(q);
    }
  }

  bool dispatchable(const Q *q) const {
    return q->next < q->nsrc && (q->cls != C_BC || q->inflight == 0);
  }

  // Oldest dispatchable query picks the (graph, kind) of the next launch;
  // the launch then takes sources of that kind's queries in admission order.
  bool pick(Pack *p) {
    Q *best = nullptr;
    for (int i = 0; i < ngraphs_ * kClasses; i++) {
      for (Q *q : active_[i]) {
        if (!dispatchable(q)) continue;
        if (!best || q->seq < best->seq) best = q;
        break;
      }
    }
    if (!best) return false;
    std::deque<Q *> &dq = active_[best->graph * kClasses + best->cls];
    p->graph = best->graph;
    p->cls = best->cls;
    p->nsrc = 0;
    p->items.clear();
    const double t = now_ms();
    for (Q *q : dq) {
      if (p->nsrc == max_pack_) break;
      if (!dispatchable(q)) continue;
      const int take = std::min(max_pack_ - p->nsrc, q->nsrc - q->next);
      p->items.push_back(Item{q, q->next, take});
      q->next += take;
      q->inflight++;
      q->launches++;
      if (q->t_first < 0) q->t_first = t;
      p->nsrc += take;
    }
    dq.erase(std::remove_if(dq.begin(), dq.end(), [](Q *q) { return q->next == q->nsrc; }), dq.end());
    stats_.launches++;
    stats_.sources += p->nsrc;
    return true;
  }

  void worker(int slot) {
    ex_->thread_init(slot);
    Pack p;
    std::vector<int32_t> src, ecc, wrap;
    std::vector<int32_t *> dr;
    std::vector<uint64_t *> sr;
    std::vector<float *> xr;
    std::vector<double *> bt;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] { return stop_ || pick(&p); });
        if (stop_) return;
      }
      const size_t n = (size_t)n_[p.graph];
      src.clear();
      dr.clear();
      sr.clear();
      xr.clear();
      bt.clear();
      ecc.assign((size_t)p.nsrc, 0);
      wrap.assign((size_t)p.nsrc, 0);
      int rc = 0;
      for (const Item &it : p.items) {
        Q *q = it.q;
        for (int j = it.first; j < it.first + it.count; j++) {
          src.push_back(q->src[(size_t)j]);
          if (q->dist) dr.push_back(q->dist + (size_t)j * n);
          if (q->sigma) sr.push_back(q->sigma + (size_t)j * n);
          if (p.cls == C_BC) {
            xr.push_back(q->delta ? q->delta + (size_t)j * n : nullptr);
            bt.push_back(q->bc);
          }
        }
        // a BC query is in one launch at a time, so this worker owns its vector
        if (p.cls == C_BC && !q->bc_zeroed) {
          rc = rc ? rc : ex_->zero(slot, q->bc, n * 8);
          q->bc_zeroed = true;
        }
      }
      if (rc == 0) {
        switch (p.cls) {
          case C_BFS_SIGMA:
            rc = ex_->search(slot, p.graph, src.data(), p.nsrc, dr.data(), sr.data(), nullptr, wrap.data());
            break;
          case C_BFS:
            rc = ex_->search(slot, p.graph, src.data(), p.nsrc, dr.data(), nullptr, nullptr, wrap.data());
            break;
          case C_BC: {
            bool any_delta = false;
            for (float *x : xr) any_delta |= x != nullptr;
            rc = ex_->bc(slot, p.graph, src.data(), p.nsrc, any_delta ? xr.data() : nullptr, bt.data(),
                         wrap.data());
            break;
          }
          default:
            rc = ex_->search(slot, p.graph, src.data(), p.nsrc, nullptr, nullptr, ecc.data(), wrap.data());
            break;
        }
      }
      // results wanted on the host: those items go to the delivery thread,
      // which copies them while this slot runs its next launch (the rows
      // belong to the queries' own blocks, which the next launch does not
      // touch) and completes them after the copies; the launch's other
      // items complete now
      Delivery dl{{}, p.cls, rc, n, {}, {}};
      std::vector<Item> now;
      std::vector<int32_t> now_ecc, now_wrap;
      size_t pos = 0;
      for (const Item &it : p.items) {
        const bool host = rc == 0 && delivering_ && it.q->to_host;
        std::vector<Item> &to = host ? dl.items : now;
        std::vector<int32_t> &e = host ? dl.ecc : now_ecc, &w = host ? dl.wrap : now_wrap;
        to.push_back(it);
        e.insert(e.end(), ecc.begin() + pos, ecc.begin() + pos + it.count);
        w.insert(w.end(), wrap.begin() + pos, wrap.begin() + pos + it.count);
        pos += (size_t)it.count;
      }
      if (!now.empty()) complete(now, p.cls, now_ecc, now_wrap, rc);
      if (!dl.items.empty()) {
        {
          std::lock_guard<std::mutex> lk(mu_);
          deliveries_.push_back(std::move(dl));
        }
        cv_deliver_.notify_one();
      }
    }
  }

  // The launch's items are done (rc: its status): per-query results,
  // failures, and queries whose last sources these were finish.
  void complete(const std::vector<Item> &items, int cls, const std::vector<int32_t> &ecc,
                const std::vector<int32_t> &wrap, int rc) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      int pos = 0;
      for (const Item &it : items) {
        Q *q = it.q;
        if (cls == C_DIAM)
          for (int j = 0; j < it.count; j++) q->diameter = std::max(q->diameter, ecc[(size_t)(pos + j)]);
        for (int j = 0; j < it.count; j++)
          if (wrap[(size_t)(pos + j)]) q->flags |= kFlagWrapped;
        pos += it.count;
        q->inflight--;
        q->done_src += it.count;
        if (rc && !q->status) {
          q->status = rc;
          // sources not yet launched are dropped with the query
          q->done_src += q->nsrc - q->next;
          q->next = q->nsrc;
          std::deque<Q *> &dq = active_[q->graph * kClasses + q->cls];
          dq.erase(std::remove(dq.begin(), dq.end(), q), dq.end());
        }
        if (q->done_src == q->nsrc && q->inflight == 0 && !q->done) finish(q);
      }
    }
    cv_done_.notify_all();
    cv_work_.notify_all();
  }

  // Copies finished launches' rows to the host arena, in launch order, on
  // the executor's delivery stream: this launch's rows of the queries that
  // want them, and a BC vector once its query's last launch is done (a BC
  // query is in one launch at a time and its sources go out in order).
  void deliverer() {
    ex_->thread_init(slots_);
    while (true) {
      Delivery dl;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_deliver_.wait(lk, [&] { return workers_done_ || !deliveries_.empty(); });
        if (deliveries_.empty()) return;
        dl = std::move(deliveries_.front());
        deliveries_.pop_front();
      }
      int rc = dl.rc;
      const size_t n = dl.n;
      for (const Item &it : dl.items) {
        Q *q = it.q;
        if (rc || !q->to_host) continue;
        const size_t j0 = (size_t)it.first, c = (size_t)it.count;
        if (q->dist) rc = ex_->deliver(q->h_dist + j0 * n, q->dist + j0 * n, c * n * 4);
        if (!rc && q->sigma) rc = ex_->deliver(q->h_sigma + j0 * n, q->sigma + j0 * n, c * n * 8);
        if (!rc && q->delta) rc = ex_->deliver(q->h_delta + j0 * n, q->delta + j0 * n, c * n * 4);
        if (!rc && q->bc && it.first + it.count == q->nsrc) rc = ex_->deliver(q->h_bc, q->bc, n * 8);
      }
      complete(dl.items, dl.cls, dl.ecc, dl.wrap, rc);
    }
  }

  Exec *ex_;
  Arena arena_, host_arena_;
  const int slots_, max_pack_;
  const std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_, cv_deliver_;
  bool stop_ = false;
  int ngraphs_ = 0;
  int64_t n_[kMaxGraphs] = {};
  int64_t next_ticket_ = 1;
  uint64_t seq_ = 0;
  std::deque<Q *> waiting_;
  std::deque<Q *> active_[kMaxGraphs * kClasses];
  std::deque<int64_t> done_fifo_;
  std::unordered_map<int64_t, std::unique_ptr<Q>> queries_;
  Stats stats_;
  std::vector<std::thread> workers_;
  std::deque<Delivery> deliveries_;
  bool delivering_ = false;    // a host arena: the delivery thread runs
  bool workers_done_ = false;  // every worker has exited (under mu_)
  std::thread deliverer_;
};

}  // namespace gus
