// server.hip -- continuous-batching pi(x) server.
//
// Requests wait in a FIFO.  An executor lane (own context, own HIP stream,
// own resident tables kept across requests) repeatedly admits the longest
// FIFO prefix whose predicted sieve-state bytes fit the HBM admission budget
// and runs it as one batch; arrivals meanwhile queue up for the next one.
// One lane: two lanes' persistent segment kernels on the device at once
// stretch each other (at 2000 requests/s, p50 85 ms with two lanes, 50 ms
// with one).  A request larger than the whole budget is rejected with
// CPX_ENOMEM; cancelled requests are dropped before admission.
// (The reference has no serving path: its CLI runs one term per process,
// src/main.cpp:23-88.)
#include "engine.hpp"
#include <algorithm>
#include <chrono>
#include <condition_variable>
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
  uint64_t bytes;
  double t_arrive;
  bool cancelled = false;
};

struct cpx_server {
  cpx_ctx *parent = nullptr;
  cpx_server_opts opts{};
  std::vector<cpx_ctx *> lanes;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req> pending;
  std::deque<cpx_result> done;
  std::unordered_map<uint64_t, int> state;     // 0 queued, 1 running, 2 done
  uint64_t next_ticket = 1;
  uint64_t inflight_bytes = 0, budget = 0;
  uint64_t submitted = 0, completed = 0, rejected = 0, cancelled = 0, batches = 0;
  uint64_t max_inflight = 0;
  uint64_t running = 0;
  std::vector<double> lat;
  double t0 = 0, t_first = -1, t_last = 0;
  bool stop = false;
};

static void finish(cpx_server *s, const Req &r, cpx_u128 pi, int status, double t_admit, double t_done)
{
  cpx_result o;
  o.ticket = r.ticket;
  o.x = r.x;
  o.pi = pi;
  o.status = status;
  o.t_arrive_ms = r.t_arrive - s->t0;
  o.t_admit_ms = t_admit - s->t0;
  o.t_done_ms = t_done - s->t0;
  s->done.push_back(o);
  s->state[r.ticket] = 2;
  if (status == CPX_OK) {
    s->completed++;
    s->lat.push_back(t_done - r.t_arrive);
    s->t_last = std::max(s->t_last, t_done);
  }
}

static void lane_loop(cpx_server *s, cpx_ctx *ctx)
{
  std::vector<Req> grp;
  std::vector<cpx_u128> xs, out;
  for (;;) {
    grp.clear();
    uint64_t gbytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
      if (s->stop && s->pending.empty()) return;
      // a lone request may wait briefly for company (bounded by max_wait_ms)
      if (s->pending.size() == 1 && s->opts.max_wait_ms > 0 && s->running > 0) {
        s->cv_work.wait_for(lk, std::chrono::microseconds((int64_t)(s->opts.max_wait_ms * 1000)),
                            [&] { return s->stop || s->pending.size() > 1; });
      }
      while (!s->pending.empty() && grp.size() < s->opts.max_batch) {
        Req &r = s->pending.front();
        if (r.cancelled) { s->state[r.ticket] = 2; s->pending.pop_front(); continue; }
        if (r.bytes > s->budget) {          // can never fit
          Req rr = r;
          s->pending.pop_front();
          s->rejected++;
          finish(s, rr, {0, 0}, CPX_ENOMEM, srv_now_ms(), srv_now_ms());
          s->cv_done.notify_all();
          continue;
        }
        if (s->inflight_bytes + gbytes + r.bytes > s->budget) break;
        gbytes += r.bytes;
        grp.push_back(r);
        s->state[r.ticket] = 1;
        s->pending.pop_front();
      }
      if (grp.empty()) {
        s->cv_done.notify_all();            // drain() re-checks (cancelled entries were dropped)
        if (!s->pending.empty())            // memory held by other lanes: wait for a release
          s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
        continue;
      }
      s->inflight_bytes += gbytes;
      s->max_inflight = std::max(s->max_inflight, s->inflight_bytes);
      s->running++;
      s->batches++;
    }
    double t_admit = srv_now_ms();
    xs.resize(grp.size());
    out.resize(grp.size());
    for (size_t i = 0; i < grp.size(); i++) xs[i] = grp[i].x;
    // the P2 table only where it already reaches: a growing table would
    // stall the stream (and is not in the admission budget)
    cpx_batch_opts bo{1, 0, 1};
    int rc = cpx_pi_batch(ctx, xs.data(), xs.size(), out.data(), &bo, nullptr);
    double t_done = srv_now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->inflight_bytes -= gbytes;
      s->running--;
      for (size_t i = 0; i < grp.size(); i++) finish(s, grp[i], out[i], rc, t_admit, t_done);
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

extern "C" {

int cpx_server_create(cpx_ctx *c, const cpx_server_opts *o, cpx_server **out)
{
  if (!c || !out) return CPX_EINVAL;
  *out = nullptr;
  cpx_server *s = new cpx_server();
  s->parent = c;
  if (o) s->opts = *o;
  if (s->opts.max_batch == 0) s->opts.max_batch = 256;
  if (s->opts.max_wait_ms < 0) s->opts.max_wait_ms = 0;
  s->budget = s->opts.hbm_budget ? s->opts.hbm_budget : c->opts.hbm_budget / 2;
  if (s->budget == 0) s->budget = 16ull << 30;
  const int L = 1;                     // executor lanes (see the header)
  for (int i = 0; i < L; i++) {
    cpx_opts lo = c->opts;
    lo.fresh_tables = 0;
    lo.use_graphs = 1;
    lo.hbm_budget = 0;
    cpx_ctx *lc = nullptr;
    int rc = cpx_ctx_create(&lo, &lc);
    if (rc) {
      for (cpx_ctx *q : s->lanes) cpx_ctx_destroy(q);
      delete s;
      return rc;
    }
    lc->shape = c->shape;
    lc->shape_tuned = c->shape_tuned;
    s->lanes.push_back(lc);
  }
  s->t0 = srv_now_ms();
  for (cpx_ctx *lc : s->lanes) s->threads.emplace_back(lane_loop, s, lc);
  *out = s;
  return CPX_OK;
}

int cpx_server_submit(cpx_server *s, cpx_u128 x, uint64_t *ticket)
{
  if (!s) return CPX_EINVAL;
  HostQuery h;
  int rc = make_host_query(s->parent, ((u128)x.hi << 64) | x.lo, 0, 0, &h);
  if (rc) return rc;
  Req r;
  r.x = x;
  r.bytes = plan_bytes_estimate(s->parent, h);
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
      return CPX_OK;
    }
  return CPX_ESTATE;
}

int cpx_server_poll(cpx_server *s, cpx_result *out, size_t max, size_t *n)
{
  if (!s || !n || (!out && max)) return CPX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  size_t k = 0;
  while (k < max && !s->done.empty()) { out[k++] = s->done.front(); s->done.pop_front(); }
  *n = k;
  return CPX_OK;
}

int cpx_server_drain(cpx_server *s)
{
  if (!s) return CPX_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    if (!s->pending.empty() || s->running) return false;
    return true;
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
  for (std::thread &t : s->threads) t.join();
  for (cpx_ctx *c : s->lanes) cpx_ctx_destroy(c);
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
