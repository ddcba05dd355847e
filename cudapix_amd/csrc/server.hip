// server.hip -- pi(x) server: continuous batching over the streaming launch,
// or (CPX_SERVER_BATCHES) one batch launch per admitted FIFO prefix.
//
// A dispatcher thread owns an executor context (tables resident across
// requests; the caller's tuning snapshotted at creation) and at most one
// streaming launch (stream.hip) at a time.  Requests wait in a FIFO; when a
// cohort slot is free and the device's backlog of unclaimed tiles has
// drained below one tile per block (or the oldest request has waited
// max_wait_ms), the dispatcher plans the longest FIFO prefix that fits the
// slot's share of the HBM budget as one cohort and publishes it into the
// running launch, which starts it as soon as blocks free up -- next to the
// tail of the cohorts before it.  Completed cohorts report through the pinned
// mailbox.  A launch closes when the server goes idle, and at a boundary:
// when a request needs larger tables or slot arenas (they are reallocated
// only while no launch runs).  (The reference has no serving path: its CLI
// runs one term per process, src/main.cpp:23-88.)
#include "engine.hpp"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

static double srv_now_ms()
{
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  uint64_t ticket;
  cpx_u128 x;
  double t_arrive;
  bool cancelled = false;
};

// A cohort slot: its entries sit at the slot's share of every arena array.
struct SSlot {
  bool busy = false;
  uint64_t seq = 0, launch = 0, q_base = 0;
  std::vector<Req> reqs;
  uint64_t need_bytes = 0;
  double t_publish = 0;
};

struct cpx_server {
  cpx_ctx *lane = nullptr;
  cpx_server_opts opts{};
  uint32_t nslots = 4;
  uint64_t budget = 0;             // device bytes for the arena
  double arena_mult = 8;           // an arena holds this many cohorts like the one that sized it
  SSlot slot[CPX_STREAM_SLOTS];
  BatchSlot arena;                 // every cohort's plan and out arrays, pinned plan image, pinned results
  Caps cursor{};                   // where the next cohort's entries go (reset when a launch closes)
  ResultDev *d_res = nullptr;      // device address of arena.h_res
  const char *d_src = nullptr;     // device address of arena.h_plan
  StreamMailbox *mb = nullptr;     // pinned: what the launch reports
  StreamPub *pub = nullptr;        // pinned: what the host publishes (the kernel polls it)
  const StreamPub *d_pubp = nullptr;   // its device address
  StreamMailbox *d_mb = nullptr;       // the mailbox's device address
  StreamCtl *d_ctl = nullptr;
  BatchDev *d_bd = nullptr, h_bd{};  // the launch's arena record in device memory (and its source)
  hipEvent_t ev_end = nullptr;
  SegShape shape{512, 32};
  uint32_t grid = 0;
  double tick_per_ms = 1e5;        // device wall clock
  bool running = false;
  uint64_t launch_no = 0, seq = 0;
  uint64_t pub_copy = 0, pub_main = 0, pub_cohorts = 0;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req> pending;
  std::deque<cpx_result> done;
  std::unordered_map<uint64_t, int> state;     // 0 queued, 1 running (finished: erased when polled)
  uint64_t next_ticket = 1;
  uint64_t inflight_bytes = 0, max_inflight = 0;
  uint64_t submitted = 0, completed = 0, rejected = 0, cancelled = 0, batches = 0, joined = 0;
  std::vector<double> lat;
  double t0 = 0, t_first = -1, t_last = 0;
  bool stop = false;
  uint32_t planning = 0;           // cohorts taken off the FIFO and not yet published
  int fatal = 0;                   // a launch failed: every later request gets this code
};

static void finish(cpx_server *s, const Req &r, cpx_u128 pi, int status, double t_admit, double t_done,
                   uint64_t launch = 0, uint64_t cohort = 0, double d0 = 0, double d1 = 0)
{
  cpx_result o{};
  o.ticket = r.ticket;
  o.x = r.x;
  o.pi = pi;
  o.status = status;
  o.t_arrive_ms = r.t_arrive - s->t0;
  o.t_admit_ms = t_admit - s->t0;
  o.t_done_ms = t_done - s->t0;
  o.launch = launch;
  o.cohort = cohort;
  o.dev_first_ms = d0;
  o.dev_last_ms = d1;
  s->done.push_back(o);
  s->state[r.ticket] = 2;
  if (status == CPX_OK) {
    s->completed++;
    s->lat.push_back(t_done - r.t_arrive);
    s->t_last = std::max(s->t_last, t_done);
  } else if (status == CPX_ENOMEM) {
    s->rejected++;
  }
}

template <typename T>
static T mb_load(const T *p)
{
  return __atomic_load_n(p, __ATOMIC_ACQUIRE);
}
template <typename T>
static void mb_store(T *p, T v)
{
  __atomic_store_n(p, v, __ATOMIC_RELEASE);
}

// Results of every cohort the device has completed (called with mu held).
static void collect(cpx_server *s)
{
  const double now = srv_now_ms();
  for (uint32_t j = 0; j < s->nslots; j++) {
    SSlot &S = s->slot[j];
    if (!S.busy || mb_load(&s->mb->done_seq[j]) != S.seq) continue;
    const ResultDev *res = (const ResultDev *)s->arena.h_res + S.q_base;
    const double d0 = (double)mb_load(&s->mb->t_first[j]) / s->tick_per_ms;
    const double d1 = (double)mb_load(&s->mb->t_last[j]) / s->tick_per_ms;
    for (size_t i = 0; i < S.reqs.size(); i++) {
      const bool ok = res[i].status == 1;
      finish(s, S.reqs[i], {res[i].pi[0], res[i].pi[1]}, ok ? CPX_OK : CPX_EDEVICE, S.t_publish, now,
             S.launch, S.seq, d0, d1);
    }
    S.reqs.clear();
    S.busy = false;
    s->inflight_bytes -= S.need_bytes;
  }
}

static bool srv_debug()
{
  static const bool on = getenv("CPX_SERVER_DEBUG") != nullptr;
  return on;
}

// Debug dump of a launch's device state after it ended (CPX_SERVER_DEBUG).
static void dump_launch(cpx_server *s, const char *why)
{
  if (!srv_debug()) return;
  StreamCtl ctl;
  (void)hipMemcpy(&ctl, s->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost);
  const StreamPub &pub = *s->pub;
  fprintf(stderr, "[srv] %s: launch %llu main_head %llu main_end %llu completed %llu cohorts %llu close %llu err %llu mb.err %u claimed %llu\n",
          why, (unsigned long long)s->launch_no, (unsigned long long)ctl.main_head, (unsigned long long)pub.main_end,
          (unsigned long long)ctl.completed, (unsigned long long)pub.cohorts, (unsigned long long)pub.close,
          (unsigned long long)ctl.error, s->mb->error, (unsigned long long)s->mb->main_claimed);
  for (uint32_t j = 0; j < s->nslots; j++)
    fprintf(stderr, "[srv]   slot %u busy %d seq %llu | pub seq %llu range %llu+%llu n_main %llu n_fin %llu n_comb %llu | main_done %u fin_done %u comb_done %u fin_head %llx comb_head %llx | done_seq %llu\n",
            j, (int)s->slot[j].busy, (unsigned long long)s->slot[j].seq, (unsigned long long)pub.rec[j].seq,
            (unsigned long long)(pub.rec[j].main_range >> CPX_STREAM_NBITS),
            (unsigned long long)(pub.rec[j].main_range & ((1ull << CPX_STREAM_NBITS) - 1)),
            (unsigned long long)pub.rec[j].n_main, (unsigned long long)pub.rec[j].n_fin, (unsigned long long)pub.rec[j].n_comb,
            ctl.main_done[j], ctl.fin_done[j], ctl.comb_done[j], (unsigned long long)ctl.fin_head[j],
            (unsigned long long)ctl.comb_head[j], (unsigned long long)s->mb->done_seq[j]);
}

static bool any_busy(const cpx_server *s)
{
  for (uint32_t j = 0; j < s->nslots; j++)
    if (s->slot[j].busy) return true;
  return false;
}

// Fails every in-flight cohort (a launch ended without completing them).
static void fail_inflight(cpx_server *s, int code)
{
  const double now = srv_now_ms();
  for (uint32_t j = 0; j < s->nslots; j++) {
    SSlot &S = s->slot[j];
    if (!S.busy) continue;
    for (const Req &r : S.reqs) finish(s, r, {0, 0}, code, S.t_publish, now);
    S.reqs.clear();
    S.busy = false;
    s->inflight_bytes -= S.need_bytes;
  }
}

static void reset_mailbox(cpx_server *s)
{
  StreamPub *p = s->pub;
  mb_store(&p->close, (uint64_t)0);
  mb_store(&p->copy_end, (uint64_t)0);
  mb_store(&p->main_end, (uint64_t)0);
  mb_store(&p->cohorts, (uint64_t)0);
  for (uint32_t j = 0; j < CPX_STREAM_SLOTS; j++) {
    mb_store(&p->rec[j].copy_range, (uint64_t)0);
    mb_store(&p->rec[j].main_range, (uint64_t)0);
  }
  mb_store(&s->mb->main_claimed, (uint64_t)0);
  mb_store(&s->mb->error, 0u);
  s->pub_copy = s->pub_main = s->pub_cohorts = 0;
  s->cursor = Caps{};
}

// Ends the running launch once its published cohorts are complete (mu held
// on entry and exit; released while the device finishes).
static void close_launch(cpx_server *s, std::unique_lock<std::mutex> &lk)
{
  if (!s->running) return;
  mb_store(&s->pub->close, (uint64_t)1);
  lk.unlock();
  const hipError_t e = hipEventSynchronize(s->ev_end);
  lk.lock();
  s->running = false;
  collect(s);
  const uint32_t err = mb_load(&s->mb->error);
  if (any_busy(s) || err) dump_launch(s, "close left cohorts");
  if (e != hipSuccess || err) {
    s->fatal = e != hipSuccess ? CPX_EDEVICE : 0;
    fail_inflight(s, CPX_EDEVICE);
  }
  reset_mailbox(s);
  s->cv_done.notify_all();
}

// The arena's pinned addresses as the device sees them (after it grew).
static int refresh_arena_addrs(cpx_server *s)
{
  void *dp = nullptr;
  hipError_t e = hipHostGetDevicePointer(&dp, s->arena.h_res, 0);
  if (e != hipSuccess) return hip_fail(s->lane, e, "hipHostGetDevicePointer");
  s->d_res = (ResultDev *)dp;
  e = hipHostGetDevicePointer(&dp, s->arena.h_plan, 0);
  if (e != hipSuccess) return hip_fail(s->lane, e, "hipHostGetDevicePointer");
  s->d_src = (const char *)dp;
  s->pub->image = s->d_src;
  s->pub->arena = (char *)s->arena.d_plan;
  return CPX_OK;
}

static int start_launch(cpx_server *s)
{
  cpx_ctx *c = s->lane;
  StreamArgs a{};
  a.bd = stream_batchdev(c, s->arena, s->d_res);
  a.bd_mem = s->d_bd;
  a.ctl = s->d_ctl;
  a.pub = s->d_pubp;
  a.mb = s->d_mb;
  a.idle_ticks = (uint64_t)(CPX_STREAM_IDLE_MS * s->tick_per_ms);
  a.wait_ticks = (uint64_t)(CPX_STREAM_WAIT_MS * s->tick_per_ms);
  a.nslots = s->nslots;
  s->h_bd = a.bd;
  hipError_t e = hipMemcpyAsync(s->d_bd, &s->h_bd, sizeof(BatchDev), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->d_ctl, 0, sizeof(StreamCtl), c->stream);
  // phase heads "not ready"
  if (e == hipSuccess) e = hipMemsetAsync(s->d_ctl->fin_head, 0xff, sizeof(s->d_ctl->fin_head), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->d_ctl->comb_head, 0xff, sizeof(s->d_ctl->comb_head), c->stream);
  if (e == hipSuccess) e = launch_stream(a, s->shape, s->grid, c->stream);
  if (e == hipSuccess) e = hipEventRecord(s->ev_end, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "streaming launch");
  s->running = true;
  s->launch_no++;
  return CPX_OK;
}

static int publish(cpx_server *s, uint32_t j, const CohortInfo &ci, std::vector<Req> &reqs, double now)
{
  SSlot &S = s->slot[j];
  StreamRec &r = s->pub->rec[j];
  const uint64_t seq = s->seq + 1;
  const uint64_t ncopy = (ci.copy_bytes + CPX_STREAM_COPY_BYTES - 1) / CPX_STREAM_COPY_BYTES;
  // the record first, then the ends that cover it (copy units before main tiles)
  r.seq = seq;
  r.n_main = ci.n_main;
  r.n_fin = ci.n_fin;
  r.n_comb = ci.nq;
  r.q_base = ci.q_base;
  r.tile_base = ci.tile_base;
  r.fin_base = ci.fin_base;
  r.nranges = ci.nranges;
  for (uint32_t k = 0; k < ci.nranges; k++) { r.off[k] = ci.off[k]; r.len[k] = ci.len[k]; }
  r.copy_bytes = ci.copy_bytes;
  mb_store(&r.copy_range, stream_pack(s->pub_copy, ncopy));
  mb_store(&r.main_range, stream_pack(s->pub_main, ci.n_main));
  mb_store(&s->pub->copy_end, s->pub_copy + ncopy);
  mb_store(&s->pub->main_end, s->pub_main + ci.n_main);
  mb_store(&s->pub->cohorts, s->pub_cohorts + 1);
  if (srv_debug())
    fprintf(stderr, "[srv] publish seq %llu slot %u n_main %u n_fin %u nq %u at %llu (launch running %d #%llu)\n",
            (unsigned long long)seq, j, ci.n_main, ci.n_fin, ci.nq, (unsigned long long)s->pub_main, (int)s->running,
            (unsigned long long)s->launch_no);
  s->seq = seq;
  s->pub_copy += ncopy;
  s->pub_main += ci.n_main;
  s->pub_cohorts++;
  s->cursor = ci.next;
  S.busy = true;
  S.seq = seq;
  S.launch = s->running ? s->launch_no : s->launch_no + 1;
  S.reqs.swap(reqs);
  S.need_bytes = ci.need_bytes;
  S.q_base = ci.q_base;
  S.t_publish = now;
  s->inflight_bytes += ci.need_bytes;
  s->max_inflight = std::max(s->max_inflight, s->inflight_bytes);
  s->batches++;
  return CPX_OK;
}

// Plans and publishes one cohort from the FIFO head, if the policy says so.
// Returns true when it did something (mu held on entry and exit).
static bool admit(cpx_server *s, std::unique_lock<std::mutex> &lk)
{
  if (s->pending.empty()) return false;
  uint32_t j = s->nslots;
  for (uint32_t i = 0; i < s->nslots; i++)
    if (!s->slot[i].busy) { j = i; break; }
  if (j == s->nslots) return false;
  const double now = srv_now_ms();
  if (s->running && any_busy(s)) {
    const uint64_t claimed = mb_load(&s->mb->main_claimed);
    const uint64_t backlog = s->pub_main > claimed ? s->pub_main - claimed : 0;
    if (backlog >= s->grid && now - s->pending.front().t_arrive < s->opts.max_wait_ms) return false;
  }
  // candidates: the FIFO prefix (cancelled ones dropped, invalid ones answered)
  std::vector<Req> cand;
  std::vector<HostQuery> hq;
  while (!s->pending.empty() && cand.size() < s->opts.max_batch) {
    Req r = s->pending.front();
    s->pending.pop_front();
    if (r.cancelled) { s->state.erase(r.ticket); continue; }
    HostQuery h;
    const int rc = make_host_query(s->lane, ((u128)r.x.hi << 64) | r.x.lo, 0, 0, &h);
    if (rc) { finish(s, r, {0, 0}, rc, now, now); continue; }
    s->state[r.ticket] = 1;
    cand.push_back(r);
    hq.push_back(h);
  }
  if (cand.empty()) { s->cv_done.notify_all(); return true; }
  s->planning++;                     // drain() waits for these while they are off the FIFO
  auto give_back = [&](size_t from) {
    for (size_t i = cand.size(); i > from; i--) {
      s->state[cand[i - 1].ticket] = 0;
      s->pending.push_front(cand[i - 1]);
    }
    cand.resize(from);
    hq.resize(from);
  };
  for (;;) {
    // planning runs without the lock (only this thread touches the lane and the slots)
    lk.unlock();
    CohortInfo ci;
    const bool grow = !s->running;
    int rc = stream_plan(s->lane, hq, s->arena, s->cursor, grow, s->budget, &ci, s->arena_mult);
    if (rc == CPX_EAGAIN && ci.why == CohortInfo::TABLES && grow) {
      rc = stream_grow_tables(s->lane, hq);
      if (!rc) rc = stream_plan(s->lane, hq, s->arena, s->cursor, true, s->budget, &ci, s->arena_mult);
    }
    if (!rc && grow) rc = refresh_arena_addrs(s);
    lk.lock();
    if (rc == CPX_OK) {
      const bool join = s->running && any_busy(s);
      rc = publish(s, j, ci, cand, srv_now_ms());
      s->planning--;
      if (rc) {
        s->fatal = rc;
        const double t = srv_now_ms();
        for (const Req &r : cand) finish(s, r, {0, 0}, rc, t, t);
        s->cv_done.notify_all();
        return true;
      }
      s->joined += join ? 1 : 0;
      if (!s->running) {
        const int r3 = start_launch(s);
        if (r3) { s->fatal = r3; fail_inflight(s, r3); }
      }
      return true;
    }
    if (rc == CPX_EAGAIN && ci.why == CohortInfo::CAPS && cand.size() > 1) {
      give_back(cand.size() / 2);      // a smaller cohort may fit the arena's remaining space
      continue;
    }
    if (rc == CPX_EAGAIN) {
      // larger tables or arenas: finish the running launch, then grow
      give_back(0);
      s->planning--;
      close_launch(s, lk);
      return true;
    }
    if (rc == CPX_ENOMEM && cand.size() > 1) {       // too many at once: halve the cohort
      give_back(cand.size() / 2);
      continue;
    }
    // alone too large for the budget, or a device error
    const double t = srv_now_ms();
    for (const Req &r : cand) finish(s, r, {0, 0}, rc, t, t);
    s->planning--;
    s->cv_done.notify_all();
    return true;
  }
}

static void dispatcher(cpx_server *s)
{
  std::unique_lock<std::mutex> lk(s->mu);
  double idle_since = -1;
  for (;;) {
    const bool busy = any_busy(s);
    if (!busy && s->pending.empty()) {
      if (s->stop) break;
      if (s->running) {
        // idle: close the launch after a short grace period
        const double now = srv_now_ms();
        if (idle_since < 0) idle_since = now;
        if (now - idle_since > 2.0) { close_launch(s, lk); idle_since = -1; continue; }
        s->cv_work.wait_for(lk, std::chrono::microseconds(200));
        continue;
      }
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
      continue;
    }
    idle_since = -1;
    collect(s);
    if (s->running && hipEventQuery(s->ev_end) == hipSuccess) {
      // the launch ended on its own (a bounded wait or the idle limit)
      s->running = false;
      collect(s);
      dump_launch(s, "launch ended on its own");
      fail_inflight(s, CPX_EDEVICE);
      reset_mailbox(s);
    }
    if (s->fatal) {
      while (!s->pending.empty()) {
        Req r = s->pending.front();
        s->pending.pop_front();
        if (!r.cancelled) finish(s, r, {0, 0}, s->fatal, srv_now_ms(), srv_now_ms());
      }
    }
    s->cv_done.notify_all();
    if (admit(s, lk)) continue;
    s->cv_work.wait_for(lk, std::chrono::microseconds(100));
  }
  close_launch(s, lk);
}

// Batch mode: the dispatcher takes the longest FIFO prefix (at most
// max_batch requests whose planned bytes fit the budget) and runs it as one
// batch launch on the executor context; requests that arrive meanwhile
// queue for the next.  Nothing stays resident on the device between batches.
static void batch_dispatcher(cpx_server *s)
{
  std::vector<Req> grp;
  std::vector<cpx_u128> xs, out;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
    if (s->pending.empty()) {
      if (s->stop) break;
      continue;
    }
    grp.clear();
    uint64_t gbytes = 0;
    while (!s->pending.empty() && grp.size() < s->opts.max_batch) {
      const Req r = s->pending.front();
      if (r.cancelled) {                 // its result was reported by the cancel
        s->pending.pop_front();
        continue;
      }
      HostQuery h;
      int rc = make_host_query(s->lane, ((u128)r.x.hi << 64) | r.x.lo, 0, 0, &h);
      const uint64_t b = rc == CPX_OK ? plan_bytes_estimate(s->lane, h) : 0;
      if (rc == CPX_OK && b > s->budget) rc = CPX_ENOMEM;
      if (rc != CPX_OK) {
        s->pending.pop_front();
        const double now = srv_now_ms();
        finish(s, r, {0, 0}, rc, now, now);
        continue;
      }
      if (!grp.empty() && gbytes + b > s->budget) break;
      gbytes += b;
      grp.push_back(r);
      s->state[r.ticket] = 1;
      s->pending.pop_front();
    }
    s->cv_done.notify_all();
    if (grp.empty()) continue;
    s->planning++;                       // drain() waits for it
    s->batches++;
    const uint64_t no = ++s->launch_no;
    s->inflight_bytes = gbytes;
    s->max_inflight = std::max(s->max_inflight, gbytes);
    lk.unlock();
    const double t_admit = srv_now_ms();
    xs.resize(grp.size());
    out.assign(grp.size(), cpx_u128{0, 0});
    for (size_t i = 0; i < grp.size(); i++) xs[i] = grp[i].x;
    // P2 by lookups only where the resident pi table reaches: growing it
    // would stall the queue (tables for max_x are built at creation)
    cpx_batch_opts bo{1, 0, 1};
    const int rc = cpx_pi_batch(s->lane, xs.data(), xs.size(), out.data(), &bo, nullptr);
    const double t_done = srv_now_ms();
    lk.lock();
    s->inflight_bytes = 0;
    for (size_t i = 0; i < grp.size(); i++) finish(s, grp[i], out[i], rc, t_admit, t_done, no, no);
    s->planning--;
    s->cv_done.notify_all();
  }
}

extern "C" {

int cpx_server_create(cpx_ctx *c, const cpx_server_opts *o, cpx_server **out)
{
  if (!c || !out) return CPX_EINVAL;
  *out = nullptr;
  cpx_server *s = new cpx_server();
  if (o) s->opts = *o;
  if (s->opts.max_batch == 0) s->opts.max_batch = 256;
  if (s->opts.max_wait_ms <= 0) s->opts.max_wait_ms = 2.0;
  if (s->opts.mode != CPX_SERVER_STREAM && s->opts.mode != CPX_SERVER_BATCHES) {
    delete s;
    return CPX_EINVAL;
  }
  s->nslots = s->opts.slots ? std::min<uint32_t>(s->opts.slots, CPX_STREAM_SLOTS) : 4;
  s->budget = s->opts.hbm_budget ? s->opts.hbm_budget : (64ull << 30);
  auto fail = [&](int rc) {
    if (s->lane) cpx_ctx_destroy(s->lane);
    if (s->mb) (void)hipHostFree(s->mb);
    if (s->pub) (void)hipHostFree(s->pub);
    if (s->d_ctl) (void)hipFree(s->d_ctl);
    if (s->d_bd) (void)hipFree(s->d_bd);
    if (s->ev_end) (void)hipEventDestroy(s->ev_end);
    delete s;
    return rc;
  };
  // the executor context: the caller's options and tuning, taken under its lock
  cpx_opts lo;
  {
    std::lock_guard<std::recursive_mutex> g(c->mu);
    lo = c->opts;
    lo.fresh_tables = 0;
    lo.use_graphs = s->opts.mode == CPX_SERVER_BATCHES ? 1 : 0;
    lo.hbm_budget = 0;
    int rc = cpx_ctx_create(&lo, &s->lane);
    if (rc) return fail(rc);
    s->lane->shape = c->shape;
    s->lane->shape_tuned = c->shape_tuned;
    s->lane->alpha_scale = c->alpha_scale;
  }
  cpx_ctx *lc = s->lane;
  if (s->opts.mode == CPX_SERVER_BATCHES) {
    // tables (with P2 lookups) for the largest x announced
    const u128 mx = ((u128)s->opts.max_x.hi << 64) | s->opts.max_x.lo;
    if (mx >= CPX_SMALL_X) {
      HostQuery h;
      int rc = make_host_query(lc, mx, 0, 0, &h);
      std::vector<HostQuery> v(1, h);
      if (!rc) rc = stream_grow_tables(lc, v, true);
      if (rc) return fail(rc);
    }
    s->t0 = srv_now_ms();
    s->th = std::thread(batch_dispatcher, s);
    *out = s;
    return CPX_OK;
  }
  if (!stream_shape_supported(lc->shape)) lc->shape = SegShape{512, 32};
  s->shape = lc->shape;
  {
    hipError_t e = hipSetDevice(lc->device);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->mb, sizeof(StreamMailbox), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->pub, sizeof(StreamPub), hipHostMallocCoherent);
    void *dp = nullptr;
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dp, s->pub, 0);
    s->d_pubp = (const StreamPub *)dp;
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dp, s->mb, 0);
    s->d_mb = (StreamMailbox *)dp;
    if (e == hipSuccess) e = hipMalloc((void **)&s->d_ctl, sizeof(StreamCtl));
    if (e == hipSuccess) e = hipMalloc((void **)&s->d_bd, sizeof(BatchDev));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_end, hipEventDisableTiming);
    int rate_khz = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, lc->device);
    if (e != hipSuccess) return fail(hip_fail(lc, e, "server setup"));
    if (rate_khz > 0) s->tick_per_ms = (double)rate_khz;
  }
  memset(s->mb, 0, sizeof(StreamMailbox));
  memset(s->pub, 0, sizeof(StreamPub));
  const int nb = stream_blocks_per_cu(s->shape);
  if (nb < 1) return fail(ctx_fail(lc, CPX_EDEVICE, "streaming kernel does not fit a CU"));
  s->grid = (uint32_t)(nb * lc->cu_count);
  // tables (with P2 lookups) for the largest x announced, and slot arenas
  // sized for a cohort of max_batch queries spread log-uniformly up to it
  const u128 mx = ((u128)s->opts.max_x.hi << 64) | s->opts.max_x.lo;
  if (mx >= CPX_SMALL_X) {
    HostQuery h;
    int rc = make_host_query(lc, mx, 0, 0, &h);
    std::vector<HostQuery> v(1, h);
    if (!rc) rc = stream_grow_tables(lc, v, true);
    if (rc) return fail(rc);
    std::vector<HostQuery> syn;
    const double l1 = std::log10((double)mx), l0 = std::min(8.0, l1);
    const uint32_t n = s->opts.max_batch;
    for (uint32_t i = 0; i < n; i++) {
      const double lx = n > 1 ? l0 + (l1 - l0) * i / (n - 1) : l1;
      const u128 x = i + 1 == n ? mx : (u128)std::pow(10.0, lx);
      if (make_host_query(lc, x, 0, 0, &h) == CPX_OK) syn.push_back(h);
    }
    while (!syn.empty()) {
      CohortInfo ci;
      rc = stream_plan(lc, syn, s->arena, Caps{}, true, s->budget, &ci, s->arena_mult);
      if (rc == CPX_OK || syn.size() == 1) break;
      syn.resize(syn.size() / 2);
    }
    if (s->arena.d_plan && (rc = refresh_arena_addrs(s))) return fail(rc);
  }
  s->t0 = srv_now_ms();
  s->th = std::thread(dispatcher, s);
  *out = s;
  return CPX_OK;
}

int cpx_server_submit(cpx_server *s, cpx_u128 x, uint64_t *ticket)
{
  if (!s) return CPX_EINVAL;
  if (x.hi >> 36) return CPX_ERANGE;          // x above 2^100
  Req r;
  r.x = x;
  r.t_arrive = srv_now_ms();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r.ticket = s->next_ticket++;
    s->state[r.ticket] = 0;
    s->pending.push_back(r);
    s->submitted++;
    if (s->t_first < 0) s->t_first = r.t_arrive;
  }
  if (ticket) *ticket = r.ticket;
  s->cv_work.notify_one();
  return CPX_OK;
}

int cpx_server_cancel(cpx_server *s, uint64_t ticket)
{
  if (!s) return CPX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->state.find(ticket);
  if (it == s->state.end()) return CPX_ENOTFOUND;
  if (it->second != 0) return CPX_ESTATE;     // already running or finished
  for (Req &r : s->pending)
    if (r.ticket == ticket && !r.cancelled) {
      r.cancelled = true;
      s->cancelled++;
      cpx_result o{};
      o.ticket = ticket;
      o.x = r.x;
      o.status = CPX_ECANCELED;
      o.t_arrive_ms = r.t_arrive - s->t0;
      o.t_admit_ms = o.t_done_ms = srv_now_ms() - s->t0;
      s->done.push_back(o);
      it->second = 2;
      s->cv_done.notify_all();
      return CPX_OK;
    }
  return CPX_ESTATE;
}

int cpx_server_poll(cpx_server *s, cpx_result *out, size_t max, size_t *n)
{
  if (!s || !n || (!out && max)) return CPX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  size_t k = 0;
  while (k < max && !s->done.empty()) {
    out[k] = s->done.front();
    s->done.pop_front();
    auto it = s->state.find(out[k].ticket);
    if (it != s->state.end() && it->second == 2) s->state.erase(it);
    k++;
  }
  *n = k;
  return CPX_OK;
}

int cpx_server_drain(cpx_server *s)
{
  if (!s) return CPX_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    bool live = false;
    for (const Req &r : s->pending)
      if (!r.cancelled) { live = true; break; }
    return !live && !s->planning && !any_busy(s);
  });
  return CPX_OK;
}

int cpx_server_stats_get(cpx_server *s, cpx_server_stats *st)
{
  if (!s || !st) return CPX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->rejected = s->rejected;
  st->cancelled = s->cancelled;
  st->batches = s->batches;
  st->max_inflight_bytes = s->max_inflight;
  st->launches = s->launch_no;
  st->joined = s->joined;
  if (s->arena.d_plan && s->arena.d_out) st->arena_bytes = s->arena.L.plan_bytes + s->arena.L.out_bytes;
  if (!s->lat.empty()) {
    std::vector<double> v = s->lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      size_t i = (size_t)std::min<double>(v.size() - 1, std::max(0.0, q * v.size() - 1e-9));
      return v[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    double sum = 0;
    for (double d : v) sum += d;
    st->mean_ms = sum / v.size();
    double span = s->t_last - s->t_first;
    st->qps = span > 0 ? (double)s->completed / (span / 1000.0) : 0;
  }
  return CPX_OK;
}

int cpx_server_destroy(cpx_server *s)
{
  if (!s) return CPX_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  BatchSlot &A = s->arena;
  if (A.d_plan) (void)hipFree(A.d_plan);
  if (A.d_out) (void)hipFree(A.d_out);
  if (A.h_plan) (void)hipHostFree(A.h_plan);
  if (A.h_res) (void)hipHostFree(A.h_res);
  if (s->mb) (void)hipHostFree(s->mb);
  if (s->pub) (void)hipHostFree(s->pub);
  if (s->d_ctl) (void)hipFree(s->d_ctl);
  if (s->d_bd) (void)hipFree(s->d_bd);
  if (s->ev_end) (void)hipEventDestroy(s->ev_end);
  if (s->lane) cpx_ctx_destroy(s->lane);
  delete s;
  return CPX_OK;
}

// Submit n requests at the given offsets (ms after the call) from this
// thread, sleeping between arrivals; returns when the last one is queued.
int cpx_server_replay(cpx_server *s, const cpx_u128 *xs, const double *t_ms, size_t n,
                      uint64_t *tickets)
{
  if (!s || (!xs && n) || (!t_ms && n)) return CPX_EINVAL;
  double base = srv_now_ms();
  for (size_t i = 0; i < n; i++) {
    double due = base + t_ms[i];
    for (;;) {
      double now = srv_now_ms();
      if (now >= due) break;
      double left = due - now;
      if (left > 0.3) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((left - 0.2) * 1000)));
      else std::this_thread::yield();
    }
    uint64_t tk = 0;
    int rc = cpx_server_submit(s, xs[i], &tk);
    if (rc) return rc;
    if (tickets) tickets[i] = tk;
  }
  return CPX_OK;
}

}  // extern "C"
