// capi.hip -- extern "C" entry points of libcudapix_amd.so (declared in
// include/cudapix.h; the reference interface each one replaces is cited there).
#include "engine.hpp"
#include <chrono>
#include <cstring>
#include <random>


static double now_ms()
{
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static inline u128 to_u128(cpx_u128 v) { return ((u128)v.hi << 64) | v.lo; }
static inline cpx_u128 from_u128(u128 v) { return {(uint64_t)v, (uint64_t)(v >> 64)}; }
static inline i128 get_i128(const uint64_t a[2]) { return (i128)(((u128)a[1] << 64) | a[0]); }
static inline cpx_i128 to_ci(const uint64_t a[2]) { return {a[0], a[1]}; }

#define CTX_GUARD(c)                         \
  if (!(c)) return CPX_EINVAL;               \
  std::lock_guard<std::recursive_mutex> lk_((c)->mu)

#define HIPCHK(call)                                     \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hip_fail(c, e_, #call); \
  } while (0)

extern "C" {

void cpx_opts_default(cpx_opts *o)
{
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->device = 0;
  o->c = 6;
  o->alpha = 0;
  o->use_graphs = 1;
}

const char *cpx_strerror(int code)
{
  switch (code) {
    case CPX_OK: return "ok";
    case CPX_EINVAL: return "invalid argument";
    case CPX_ENOMEM: return "out of device memory or over the HBM budget";
    case CPX_EDEVICE: return "device (HIP) error";
    case CPX_ERANGE: return "argument out of the supported range";
    case CPX_EAGAIN: return "no result ready";
    case CPX_ECANCELED: return "request cancelled";
    case CPX_ENOTFOUND: return "unknown ticket";
    case CPX_ESTATE: return "invalid state";
    default: return "unknown error";
  }
}

const char *cpx_ctx_last_error(const cpx_ctx *c) { return c ? c->err.c_str() : "null context"; }

int cpx_ctx_create(const cpx_opts *o, cpx_ctx **out)
{
  if (!out) return CPX_EINVAL;
  *out = nullptr;
  cpx_opts op;
  if (o) op = *o; else cpx_opts_default(&op);
  if (op.c == 0) op.c = 6;
  if (op.c > 8) return CPX_EINVAL;
  cpx_ctx *c = new cpx_ctx();
  c->opts = op;
  c->device = op.device;
  hipError_t e = hipSetDevice(op.device);
  if (e != hipSuccess) { int r = hip_fail(c, e, "hipSetDevice"); delete c; return r; }
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    int r = hip_fail(c, e, "hipStreamCreate"); delete c; return r;
  }
  if ((e = hipStreamCreateWithFlags(&c->up_stream, hipStreamNonBlocking)) != hipSuccess) {
    int r = hip_fail(c, e, "hipStreamCreate"); cpx_ctx_destroy(c); return r;
  }
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, op.device)) != hipSuccess) {
    int r = hip_fail(c, e, "hipGetDeviceProperties"); cpx_ctx_destroy(c); return r;
  }
  c->cu_count = prop.multiProcessorCount;
  if (op.hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) c->opts.hbm_budget = (uint64_t)(fr * 0.9);
  }
  if (op.seg_threads && op.seg_kwords) {
    SegShape s{(int)op.seg_threads, (int)op.seg_kwords};
    if (!seg_shape_supported(s)) { cpx_ctx_destroy(c); return CPX_EINVAL; }
    c->shape = s;
    c->shape_tuned = true;
  }
  *out = c;
  return CPX_OK;
}

int cpx_ctx_destroy(cpx_ctx *c)
{
  if (!c) return CPX_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->up_stream) (void)hipStreamSynchronize(c->up_stream);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_slots(c);
  free_tables(c);
  if (c->small.p) (void)hipFree(c->small.p);
  if (c->small_primes.p) (void)hipFree(c->small_primes.p);
  if (c->wheel.p) (void)hipFree(c->wheel.p);
  if (c->winfo.p) (void)hipFree(c->winfo.p);
  if (c->up_stream) (void)hipStreamDestroy(c->up_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CPX_OK;
}

int cpx_choose_y(cpx_ctx *c, cpx_u128 x, uint64_t *y)
{
  CTX_GUARD(c);
  if (!y) return CPX_EINVAL;
  *y = choose_y_host(c, to_u128(x));
  return CPX_OK;
}

uint64_t cpx_ctx_device_bytes(const cpx_ctx *c)
{
  if (!c) return 0;
  return c->table_bytes + slot_bytes(c);
}

// -------------------------------------------------------------- one query

static int solve_one(cpx_ctx *c, u128 x, uint64_t y, uint32_t cc, uint32_t mask, ResultDev *r,
                     HostQuery *hq_out = nullptr, double *ms = nullptr)
{
  double t0 = now_ms();
  HIPCHK(hipSetDevice(c->device));
  HostQuery hq;
  int rc = make_host_query(c, x, y, cc, &hq);
  if (rc) return rc;
  std::vector<HostQuery> v{hq};
  std::vector<ResultDev> res;
  if ((rc = run_queries(c, v, mask, 0, res, nullptr, false))) return rc;
  *r = res[0];
  if (hq_out) *hq_out = hq;
  if (ms) *ms = now_ms() - t0;
  return CPX_OK;
}

static void fill_terms(cpx_terms *t, const ResultDev &r, const HostQuery &h, double ms)
{
  if (!t) return;
  memset(t, 0, sizeof(*t));
  t->pi = {r.pi[0], r.pi[1]};
  t->ordinary = to_ci(r.ordinary);
  t->trivial = to_ci(r.trivial);
  t->easy = to_ci(r.easy);
  t->hard = to_ci(r.hard);
  t->p2 = to_ci(r.p2);
  t->y = h.y; t->z = h.z; t->c = h.c; t->pi_y = r.pi_y;
  t->alpha = h.cbrtx ? (double)h.y / (double)h.cbrtx : 0;
  t->ms = ms;
}

int cpx_pi_u128(cpx_ctx *c, cpx_u128 x, uint64_t y, uint32_t cc, cpx_u128 *out, cpx_terms *terms)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  ResultDev r; HostQuery h; double ms = 0;
  int rc = solve_one(c, to_u128(x), y, cc, TERM_ALL, &r, &h, &ms);
  if (rc) return rc;
  *out = {r.pi[0], r.pi[1]};
  fill_terms(terms, r, h, ms);
  return CPX_OK;
}

int cpx_pi_u64(cpx_ctx *c, uint64_t x, uint64_t y, uint32_t cc, uint64_t *out, cpx_terms *terms)
{
  if (!out) return CPX_EINVAL;
  cpx_u128 o;
  int rc = cpx_pi_u128(c, {x, 0}, y, cc, &o, terms);
  if (rc) return rc;
  *out = o.lo;
  return CPX_OK;
}

static int term_i128(cpx_ctx *c, u128 x, uint64_t y, uint32_t cc, uint32_t mask, int which,
                     i128 *out)
{
  ResultDev r;
  int rc = solve_one(c, x, y, cc, mask, &r);
  if (rc) return rc;
  const uint64_t *src = which == 0 ? r.ordinary : which == 1 ? r.trivial : which == 2 ? r.easy
                      : which == 3 ? r.hard : r.p2;
  *out = get_i128(src);
  return CPX_OK;
}

int cpx_p2_u128(cpx_ctx *c, cpx_u128 x, uint64_t y, cpx_u128 *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  i128 v;
  int rc = term_i128(c, to_u128(x), y, 0, TERM_P2, 4, &v);
  if (rc) return rc;
  *out = from_u128((u128)v);
  return CPX_OK;
}

int cpx_p2_u64(cpx_ctx *c, uint64_t x, uint64_t y, uint64_t *out)
{
  if (!out) return CPX_EINVAL;
  cpx_u128 o;
  int rc = cpx_p2_u128(c, {x, 0}, y, &o);
  if (!rc) *out = o.lo;
  return rc;
}

int cpx_ordinary_u128(cpx_ctx *c, cpx_u128 x, uint64_t y, uint16_t cc, cpx_i128 *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  i128 v;
  int rc = term_i128(c, to_u128(x), y, cc, TERM_ORD, 0, &v);
  if (rc) return rc;
  *out = {(uint64_t)v, (uint64_t)((u128)v >> 64)};
  return CPX_OK;
}

int cpx_ordinary_u64(cpx_ctx *c, uint64_t x, uint64_t y, uint16_t cc, int64_t *out)
{
  if (!out) return CPX_EINVAL;
  cpx_i128 o;
  int rc = cpx_ordinary_u128(c, {x, 0}, y, cc, &o);
  if (!rc) *out = (int64_t)o.lo;
  return rc;
}

int cpx_s2_trivial_u128(cpx_ctx *c, cpx_u128 x, uint64_t y, cpx_u128 *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  i128 v;
  int rc = term_i128(c, to_u128(x), y, 0, TERM_TRIV, 1, &v);
  if (rc) return rc;
  *out = from_u128((u128)v);
  return CPX_OK;
}

int cpx_s2_trivial_u64(cpx_ctx *c, uint64_t x, uint64_t y, uint64_t *out)
{
  if (!out) return CPX_EINVAL;
  cpx_u128 o;
  int rc = cpx_s2_trivial_u128(c, {x, 0}, y, &o);
  if (!rc) *out = o.lo;
  return rc;
}

int cpx_s3_u128(cpx_ctx *c, cpx_u128 x, uint64_t y, uint32_t cc, cpx_i128 *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  i128 v;
  int rc = term_i128(c, to_u128(x), y, cc, TERM_HARD, 3, &v);
  if (rc) return rc;
  *out = {(uint64_t)v, (uint64_t)((u128)v >> 64)};
  return CPX_OK;
}

int cpx_s3_u64(cpx_ctx *c, uint64_t x, uint64_t y, uint32_t cc, int64_t *out)
{
  if (!out) return CPX_EINVAL;
  cpx_i128 o;
  int rc = cpx_s3_u128(c, {x, 0}, y, cc, &o);
  if (!rc) *out = (int64_t)o.lo;
  return rc;
}

int cpx_v_u128(cpx_ctx *c, cpx_u128 x, uint64_t y, cpx_i128 *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  i128 v;
  int rc = term_i128(c, to_u128(x), y, 0, TERM_EASY, 2, &v);
  if (rc) return rc;
  *out = {(uint64_t)v, (uint64_t)((u128)v >> 64)};
  return CPX_OK;
}

int cpx_v_u64(cpx_ctx *c, uint64_t x, uint64_t y, int64_t *out)
{
  if (!out) return CPX_EINVAL;
  cpx_i128 o;
  int rc = cpx_v_u128(c, {x, 0}, y, &o);
  if (!rc) *out = (int64_t)o.lo;
  return rc;
}

int cpx_v_split_u64(cpx_ctx *c, uint64_t x, uint64_t y, int64_t *va, int64_t *vb)
{
  CTX_GUARD(c);
  if (!va || !vb) return CPX_EINVAL;
  i128 a, b;
  int rc = term_i128(c, (u128)x, y, 0, TERM_EASY | TERM_EASY_A, 2, &a);
  if (rc) return rc;
  rc = term_i128(c, (u128)x, y, 0, TERM_EASY | TERM_EASY_B, 2, &b);
  if (rc) return rc;
  *va = (int64_t)a;
  *vb = (int64_t)b;
  return CPX_OK;
}

// pi(v) from the resident table (v <= pitab_limit), read on the host.
static int pi_read(cpx_ctx *c, uint64_t v, uint64_t *out)
{
  PiEntry e;
  HIPCHK(hipMemcpyAsync(&e, c->pitab.p + (v >> 6), sizeof(e), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const uint32_t k = (uint32_t)(((v & 63) + 1) >> 1);
  *out = (uint64_t)(v >= 2) + e.count + (uint64_t)__builtin_popcount(e.bits & (uint32_t)((1ull << k) - 1));
  return CPX_OK;
}

int cpx_s2_sigma_u128(cpx_ctx *c, cpx_u128 xx, uint64_t y, cpx_sigma_terms *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  const u128 x = to_u128(xx);
  HostQuery h;
  int rc = make_host_query(c, x, y, 0, &h);
  if (rc) return rc;
  if (h.small) return ctx_fail(c, CPX_ERANGE, "x below 2^14 has no easy leaves to partition");
  y = h.y;
  const uint64_t q4 = h.qrtx, c3 = h.cbrtx;
  if ((x / q4) >> 64) return ctx_fail(c, CPX_ERANGE, "x / x^(1/4) must fit 64 bits");
  const uint64_t sz = cpx_isqrt128(x / y);
  // the primes to sqrt(x/p) for the smallest p above x^(1/4), pi(n) to sqrt(x)
  const uint64_t P = std::max<uint64_t>(c3, cpx_isqrt128(x / q4)) + 1;
  if ((rc = ensure_tables(c, std::max<uint64_t>(h.sqrtx, P), P, 1))) return rc;
  uint64_t a, b, cc, d;
  if ((rc = pi_read(c, y, &a)) || (rc = pi_read(c, c3, &b)) || (rc = pi_read(c, q4, &cc)) ||
      (rc = pi_read(c, sz, &d)))
    return rc;
  const uint64_t j0 = cc, np = b > cc ? b - cc : 0;   // primes[j] for j in [pi(q4), pi(c3))
  i128 A = 0, s4 = 0, s5 = 0, s6 = 0;
  if (np) {
    i128 *part = nullptr;
    HIPCHK(hipMalloc((void **)&part, np * 4 * sizeof(i128)));
    std::vector<i128> hp(np * 4);
    hipError_t e = launch_sigma(x, y, sz, c->primes.p, c->pitab.p, j0, np, part, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hp.data(), part, np * 4 * sizeof(i128), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return hip_fail(c, e, "sigma partition");
    for (uint64_t i = 0; i < np; i++) {
      A += hp[4 * i]; s4 += hp[4 * i + 1]; s5 += hp[4 * i + 2]; s6 += hp[4 * i + 3];
    }
  }
  const i128 pa = (i128)a, pb = (i128)b, pc = (i128)cc, pd = (i128)d;
  const i128 sg[6] = {(pa - pb) * (pa - pb - 1) / 2,
                      pa * (pc * (pc - 3) / 2 - pd * (pd - 3) / 2),
                      pb * (pb - 1) * (2 * pb - 1) / 6 - pb - pc * (pc - 1) * (2 * pc - 1) / 6 + pc,
                      pa * s4, s5, s6};
  auto put = [](i128 v) { return cpx_i128{(uint64_t)v, (uint64_t)((u128)v >> 64)}; };
  out->a = put(A);
  for (int i = 0; i < 6; i++) out->sigma[i] = put(sg[i]);
  out->s2 = put(A + sg[1] + sg[2] + sg[3] + 2 * sg[4] + sg[5]);
  out->y = y;
  // exact unless a prime lies in (x^(1/4), x / y^2]: the first prime above x^(1/4)
  const u128 lim = x / ((u128)y * y);
  out->exact = !(np && (u128)c->h_primes[j0] <= lim);
  out->pad = 0;
  return CPX_OK;
}

int cpx_div_x_by_y(cpx_ctx *c, cpx_u128 x, uint64_t *d_y, size_t len)
{
  CTX_GUARD(c);
  if (!d_y && len) return CPX_EINVAL;
  HIPCHK(launch_div_x_by_y(x.lo, x.hi, d_y, len, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CPX_OK;
}

// ------------------------------------------------------- primes / tables

int cpx_count_primes(cpx_ctx *c, uint64_t lo, uint64_t hi, uint64_t *out)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  *out = 0;
  if (hi <= lo) return CPX_OK;
  HIPCHK(hipSetDevice(c->device));
  unsigned long long *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(unsigned long long)));
  int rc = CPX_OK;
  unsigned long long h = 0;
  // decided once: the sieve branch builds tables that may then reach hi
  const bool from_table = hi - 1 <= c->pitab_limit && c->pitab.p;
  do {
    if (from_table) {
      hipError_t e = launch_count_range(c->pitab.p, lo, hi, d, c->stream);
      if (e) { rc = hip_fail(c, e, "count_range"); break; }
    } else {
      uint64_t r = cpx_isqrt128((u128)hi) + 1;
      if ((rc = ensure_tables(c, r, r, 64))) break;
      hipError_t e = launch_zero_u32((uint32_t *)d, 2, c->stream);
      if (!e) e = launch_count_primes_sieve(lo, hi, c->primes.p, (uint32_t)c->nprimes, d, c->stream);
      if (e) { rc = hip_fail(c, e, "count_primes_sieve"); break; }
    }
    hipError_t e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    if (e) { rc = hip_fail(c, e, "count copy"); break; }
    // the sieve counts odd primes >= 3; add the prime 2
    if (!from_table && lo <= 2 && hi > 2) h += 1;
  } while (0);
  (void)hipFree(d);
  if (!rc) *out = h;
  return rc;
}

int cpx_get_d_pitable(cpx_ctx *c, uint32_t hi, uint32_t **d_out)
{
  CTX_GUARD(c);
  if (!d_out) return CPX_EINVAL;
  *d_out = nullptr;
  HIPCHK(hipSetDevice(c->device));
  int rc = ensure_tables(c, hi, 64, 64);
  if (rc) return rc;
  uint64_t n = hi / 2;
  uint32_t *d = nullptr;
  HIPCHK(hipMalloc(&d, std::max<uint64_t>(n, 1) * 4));
  hipError_t e = launch_expand_pitab_u32(c->pitab.p, d, 0, n, 0, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { (void)hipFree(d); return hip_fail(c, e, "expand pitab"); }
  *d_out = d;
  return CPX_OK;
}

int cpx_free(cpx_ctx *c, void *p)
{
  if (!c) return CPX_EINVAL;
  if (p) (void)hipFree(p);
  return CPX_OK;
}

int cpx_copy_to_host(cpx_ctx *c, void *dst, const void *d_src, size_t bytes)
{
  CTX_GUARD(c);
  if (bytes && (!dst || !d_src)) return CPX_EINVAL;
  if (!bytes) return CPX_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CPX_OK;
}

struct cpx_pitable {
  cpx_ctx *ctx;
  uint64_t base, bottom, range, pi_base;
  uint32_t *d = nullptr;
  bool pi_set = false;          // pi_base given by the caller (set_pi_base): carried, not counted
};

// Primes in [lo, hi) from the resident pi table (which reaches hi).
static int pitable_count(cpx_ctx *c, uint64_t lo, uint64_t hi, uint64_t *out)
{
  *out = 0;
  if (hi <= lo) return CPX_OK;
  unsigned long long *d;
  HIPCHK(hipMalloc(&d, 8));
  hipError_t e = launch_count_range(c->pitab.p, lo, hi, d, c->stream);
  unsigned long long h = 0;
  if (!e) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e) return hip_fail(c, e, "pi table count");
  *out = h;
  return CPX_OK;
}

static int pitable_fill(cpx_pitable *t)
{
  cpx_ctx *c = t->ctx;
  int rc = ensure_tables(c, t->base + t->range + 2, 64, 64);
  if (rc) return rc;
  // the table holds pi(n) - pi(base) (src/pitable.cu:196-218 writes the
  // window's own prime index); pi_base is reported beside it, counted unless
  // the caller set it
  uint64_t pb = 0;
  if (t->base >= 2 && (rc = pitable_count(c, 0, t->base, &pb))) return rc;
  if (!t->pi_set) t->pi_base = pb;
  HIPCHK(launch_expand_pitab_u32(c->pitab.p, t->d, t->base, t->range / 2 + 2, pb, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CPX_OK;
}

int cpx_pitable_create(cpx_ctx *c, uint64_t base, uint64_t bottom, uint64_t range, cpx_pitable **out)
{
  CTX_GUARD(c);
  if (!out || range == 0 || (base & 1)) return CPX_EINVAL;
  *out = nullptr;
  cpx_pitable *t = new cpx_pitable{c, base, bottom, range, 0, nullptr};
  hipError_t e = hipMalloc(&t->d, (range / 2 + 2) * 4);
  if (e) { delete t; return hip_fail(c, e, "hipMalloc(pitable)"); }
  *out = t;
  return CPX_OK;
}

int cpx_pitable_destroy(cpx_pitable *t)
{
  if (!t) return CPX_EINVAL;
  if (t->d) (void)hipFree(t->d);
  delete t;
  return CPX_OK;
}

int cpx_pitable_current(cpx_pitable *t, uint32_t **d)
{
  if (!t || !d) return CPX_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(t->ctx->mu);
  int rc = pitable_fill(t);
  if (!rc) *d = t->d;
  return rc;
}

int cpx_pitable_next_up(cpx_pitable *t, uint32_t **d)
{
  if (!t || !d) return CPX_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(t->ctx->mu);
  if (t->pi_set) {                 // a caller's pi_base moves up by the window's primes
    int rc = ensure_tables(t->ctx, t->base + t->range + 2, 64, 64);
    uint64_t n = 0;
    if (!rc) rc = pitable_count(t->ctx, t->base, t->base + t->range, &n);
    if (rc) return rc;
    t->pi_base += n;
  }
  t->base += t->range;
  return cpx_pitable_current(t, d);
}

int cpx_pitable_next_down(cpx_pitable *t, uint32_t **d)
{
  if (!t || !d) return CPX_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(t->ctx->mu);
  const uint64_t b0 = t->base;
  if (t->base >= t->bottom + t->range) t->base -= t->range;
  else { t->range = t->base - t->bottom; t->base = t->bottom; }
  if (t->range == 0) return CPX_ERANGE;
  if (t->pi_set) {                 // ... and down by the new window's
    uint64_t n = 0;
    const int rc = pitable_count(t->ctx, t->base, b0, &n);
    if (rc) return rc;
    t->pi_base -= n;
  }
  return cpx_pitable_current(t, d);
}

int cpx_pitable_set_range(cpx_pitable *t, uint64_t range)
{
  if (!t || range == 0) return CPX_EINVAL;
  cpx_ctx *c = t->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  uint32_t *d = nullptr;
  HIPCHK(hipMalloc(&d, (range / 2 + 2) * 4));
  if (t->d) (void)hipFree(t->d);
  t->d = d;
  t->range = range;
  return CPX_OK;
}

int cpx_pitable_set_pi_base(cpx_pitable *t, uint64_t pi_base)
{
  if (!t) return CPX_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(t->ctx->mu);
  t->pi_base = pi_base;
  t->pi_set = true;
  return CPX_OK;
}

int cpx_pitable_max_range(cpx_ctx *c, uint64_t *range)
{
  CTX_GUARD(c);
  if (!range) return CPX_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  *range = (uint64_t)(fr / 2.5) & ~1ull;      // a table of range/2 words then takes 80 % of the free bytes
  return CPX_OK;
}

int cpx_pitable_info(cpx_pitable *t, uint64_t *base, uint64_t *range, uint64_t *pi_base)
{
  if (!t) return CPX_EINVAL;
  if (base) *base = t->base;
  if (range) *range = t->range;
  if (pi_base) *pi_base = t->pi_base;
  return CPX_OK;
}

struct cpx_phi {
  cpx_ctx *ctx;
  uint32_t max;
  uint32_t *d_lpf = nullptr;     // odd k = 2i + 1 < max
  uint64_t n = 0;
  uint32_t p_max = 0;            // the prime list ends at this bound
  uint32_t *d_plist = nullptr;   // the prime list (made on first request)
  std::vector<uint32_t> h_plist;
};

int cpx_phi_create_pmax(cpx_ctx *c, uint32_t max, uint32_t p_max, cpx_phi **out)
{
  CTX_GUARD(c);
  if (!out || max < 3) return CPX_EINVAL;
  *out = nullptr;
  HIPCHK(hipSetDevice(c->device));
  if (p_max == 0) p_max = (uint32_t)cpx_isqrt128((u128)max);
  uint64_t r = std::max<uint64_t>(cpx_isqrt128((u128)max + 16384) + 2, (uint64_t)p_max + 2);
  int rc = ensure_tables(c, r, r, 64);
  if (rc) return rc;
  cpx_phi *p = new cpx_phi{c, max, nullptr, 0, p_max, nullptr, {}};
  p->n = ((uint64_t)max + 1) / 2;
  uint64_t nalloc = (p->n + 4095) / 4096 * 4096;
  hipError_t e = hipMalloc(&p->d_lpf, nalloc * 4);
  if (!e) e = launch_mu_lpf_split(p->d_lpf, nullptr, nalloc, c->primes.p, (uint32_t)c->nprimes, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { if (p->d_lpf) (void)hipFree(p->d_lpf); delete p; return hip_fail(c, e, "Phi lpf"); }
  *out = p;
  return CPX_OK;
}

int cpx_phi_create(cpx_ctx *c, uint32_t max, cpx_phi **out)
{
  return cpx_phi_create_pmax(c, max, 0, out);
}

int cpx_phi_destroy(cpx_phi *p)
{
  if (!p) return CPX_EINVAL;
  if (p->d_lpf) (void)hipFree(p->d_lpf);
  if (p->d_plist) (void)hipFree(p->d_plist);
  delete p;
  return CPX_OK;
}

int cpx_phi_lpf(cpx_phi *p, uint32_t **d_lpf, uint64_t *len)
{
  if (!p || !d_lpf) return CPX_EINVAL;
  *d_lpf = p->d_lpf;
  if (len) *len = p->n;
  return CPX_OK;
}

// The sieving-prime list of the reference's Phi: src/phi.cu:44-45 indexes it
// as the continuation of its own table of the first twelve primes (a > 12
// reads h_primeList[a - 12]), so it runs from the 13th prime, 41, to p_max.
int cpx_phi_prime_list(cpx_phi *p, uint32_t **d_list, const uint32_t **h_list, uint32_t *len)
{
  if (!p) return CPX_EINVAL;
  cpx_ctx *c = p->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (!p->d_plist) {
    std::vector<uint32_t> v;
    for (uint64_t i = 0; i < c->nprimes && c->h_primes[i] <= p->p_max; i++)
      if (c->h_primes[i] >= 41) v.push_back(c->h_primes[i]);
    uint32_t *d = nullptr;
    HIPCHK(hipMalloc(&d, std::max<size_t>(v.size(), 1) * 4));
    hipError_t e = v.empty() ? hipSuccess
                             : hipMemcpyAsync(d, v.data(), v.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    if (e) { (void)hipFree(d); return hip_fail(c, e, "phi prime list"); }
    p->d_plist = d;
    p->h_plist.swap(v);
  }
  if (d_list) *d_list = p->d_plist;
  if (h_list) *h_list = p->h_plist.data();
  if (len) *len = (uint32_t)p->h_plist.size();
  return CPX_OK;
}

int cpx_phi_mod_cache(cpx_phi *p, uint32_t pi_max, uint64_t *out)
{
  if (!p || (pi_max > 1 && !out)) return CPX_EINVAL;
  cpx_ctx *c = p->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (pi_max > 16) return ctx_fail(c, CPX_ERANGE, "pi_max above 16 (the primorial leaves 64 bits)");
  static const uint32_t sp[16] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};
  uint64_t w = 1, tot = 1;
  for (uint32_t k = 1; k < pi_max; k++) {
    w *= sp[k - 1];
    tot *= sp[k - 1] - 1;
    // phi(p_1...p_k, k) = prod (p_i - 1); counted by the table where it reaches
    if (w < p->max) {
      uint64_t v = 0;
      const int rc = cpx_phi_eval(p, (uint32_t)w, k, &v);
      if (rc) return rc;
      if (v != tot) return ctx_fail(c, CPX_EDEVICE, "phi of a primorial disagrees with its totient");
    }
    out[k - 1] = tot;
  }
  return CPX_OK;
}

int cpx_phi_eval(cpx_phi *p, uint32_t x, uint32_t a, uint64_t *out)
{
  if (!p || !out) return CPX_EINVAL;
  cpx_ctx *c = p->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (x >= p->max) return ctx_fail(c, CPX_ERANGE, "x must be below the Phi table size");
  if (a == 0) { *out = x; return CPX_OK; }
  if (a > c->nprimes) return ctx_fail(c, CPX_ERANGE, "a beyond the resident prime list");
  if (x == 0) { *out = 0; return CPX_OK; }
  uint32_t pa = c->h_primes[a - 1];
  unsigned long long *d;
  HIPCHK(hipMalloc(&d, 8));
  hipError_t e = launch_zero_u32((uint32_t *)d, 2, c->stream);
  // odd k in [3, x]: indices 1 .. (x-1)/2
  uint64_t i1 = (uint64_t)(x - 1) / 2 + 1;
  if (!e) e = launch_phi_count_lpf(p->d_lpf, 1, i1, pa, d, c->stream);
  unsigned long long h = 0;
  if (!e) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e) return hip_fail(c, e, "phi count");
  *out = 1 + h;                                    // k = 1
  return CPX_OK;
}

int cpx_phi_range(cpx_phi *p, uint32_t num, uint32_t a, uint32_t **d_out)
{
  if (!p || !d_out) return CPX_EINVAL;
  cpx_ctx *c = p->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  *d_out = nullptr;
  uint64_t n = num / 2;
  if (n > p->n) return ctx_fail(c, CPX_ERANGE, "range beyond the Phi table");
  if (a == 0 || a > c->nprimes) return ctx_fail(c, CPX_ERANGE, "a out of range");
  uint32_t pa = c->h_primes[a - 1];
  uint32_t *d = nullptr;
  HIPCHK(hipMalloc(&d, std::max<uint64_t>(n, 1) * 4));
  hipError_t e = launch_phi_prefix(p->d_lpf, n, pa, d, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { (void)hipFree(d); return hip_fail(c, e, "phi range"); }
  *d_out = d;
  return CPX_OK;
}

int cpx_gen_d_lpf(cpx_ctx *c, uint64_t bottom, uint64_t top, uint32_t **d_out)
{
  CTX_GUARD(c);
  if (!d_out || (bottom & 1) || top <= bottom || top >= (1ull << 32)) return CPX_EINVAL;
  *d_out = nullptr;
  uint64_t r = cpx_isqrt128((u128)top + 16384) + 2;
  int rc = ensure_tables(c, r, r, 64);
  if (rc) return rc;
  uint64_t n = (top - bottom) / 2, nalloc = (n + 4095) / 4096 * 4096;
  uint32_t *d = nullptr;
  HIPCHK(hipMalloc(&d, nalloc * 4));
  hipError_t e = launch_mu_lpf_range(d, nullptr, bottom / 2, nalloc, c->primes.p, (uint32_t)c->nprimes, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { (void)hipFree(d); return hip_fail(c, e, "gen_d_lpf"); }
  *d_out = d;
  return CPX_OK;
}

// Windows of 64-bit n end below this (a sieve segment past top and the
// sieving-prime limit sqrt(top) stay inside 64 and 32 bits).
static const uint64_t LPF64_TOP_MAX = ~0ull - (1ull << 34);

int cpx_gen_d_mu(cpx_ctx *c, uint64_t bottom, uint64_t top, int8_t **d_out)
{
  CTX_GUARD(c);
  if (!d_out || (bottom & 1) || top <= bottom || top > LPF64_TOP_MAX) return CPX_EINVAL;
  *d_out = nullptr;
  uint64_t r = cpx_isqrt128((u128)top + 16384) + 2;
  int rc = ensure_tables(c, r, r, 64);
  if (rc) return rc;
  uint64_t n = (top - bottom) / 2, nalloc = (n + 4095) / 4096 * 4096;
  int8_t *d = nullptr;
  HIPCHK(hipMalloc(&d, nalloc));
  // 32-bit cofactors while the window's last segment stays below 2^32
  hipError_t e = bottom + 2 * nalloc < (1ull << 32)
                     ? launch_mu_lpf_range(nullptr, d, bottom / 2, nalloc, c->primes.p, (uint32_t)c->nprimes, c->stream)
                     : launch_mu_lpf64_range(nullptr, d, bottom / 2, nalloc, c->primes.p, (uint32_t)c->nprimes,
                                             c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { (void)hipFree(d); return hip_fail(c, e, "gen_d_mu"); }
  *d_out = d;
  return CPX_OK;
}

int cpx_gen_d_lpf64(cpx_ctx *c, uint64_t bottom, uint64_t top, uint64_t **d_out)
{
  CTX_GUARD(c);
  if (!d_out || (bottom & 1) || top <= bottom || top > LPF64_TOP_MAX) return CPX_EINVAL;
  *d_out = nullptr;
  uint64_t r = cpx_isqrt128((u128)top + 16384) + 2;
  int rc = ensure_tables(c, r, r, 64);
  if (rc) return rc;
  uint64_t n = (top - bottom) / 2, nalloc = (n + 4095) / 4096 * 4096;
  uint64_t *d = nullptr;
  HIPCHK(hipMalloc(&d, nalloc * 8));
  hipError_t e = launch_mu_lpf64_range(d, nullptr, bottom / 2, nalloc, c->primes.p, (uint32_t)c->nprimes, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { (void)hipFree(d); return hip_fail(c, e, "gen_d_lpf64"); }
  *d_out = d;
  return CPX_OK;
}

// Ranges above this would let a sieve segment's end (2^20 past hi) or the
// sieving-prime limit sqrt(hi) leave 64 and 32 bits.
static const uint64_t RANGE_HI_MAX = ~0ull - (1ull << 34);
// Below this hi the primes come from the pi table, grown to hi if needed.
static const uint64_t RANGE_TABLE_MAX = 1ull << 24;

// Primes of [lo, hi) into a new device array (u64, or u32 when out32) by the
// segmented bit sieve: sieving primes to sqrt(hi) only, one 2^20-integer
// segment per block, per-segment counts scanned on the host, then each
// segment's primes written at its offset.
static int range_primes(cpx_ctx *c, uint64_t lo, uint64_t hi, bool out32, void **d_out, size_t *len)
{
  if (hi > RANGE_HI_MAX) return ctx_fail(c, CPX_ERANGE, "prime range above 2^64 - 2^34");
  const uint64_t base = lo & ~63ull;
  const uint64_t nseg = range_sieve_segments(base, hi);
  const uint64_t nwords = (hi - base + 63) / 64;
  const uint64_t r = cpx_isqrt128((u128)hi + (1ull << 21)) + 1;     // covers the last segment's end
  int rc = ensure_tables(c, r, r, 64);
  if (rc) return rc;
  uint32_t *words = nullptr, *cnt = nullptr;
  uint64_t *off = nullptr;
  void *out = nullptr;
  const size_t esz = out32 ? 4 : 8;
  hipError_t e = hipMalloc(&words, nwords * 4);
  if (!e) e = hipMalloc(&cnt, nseg * 4);
  if (!e) e = hipMalloc(&off, nseg * 8);
  if (!e) e = launch_bit_sieve(base, lo, hi, nwords, c->primes.p, (uint32_t)c->nprimes, false, words, cnt,
                               c->stream);
  std::vector<uint32_t> hc(nseg);
  std::vector<uint64_t> ho(nseg);
  if (!e) e = hipMemcpyAsync(hc.data(), cnt, nseg * 4, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  uint64_t total = 0;
  for (uint64_t b = 0; b < nseg; b++) { ho[b] = total; total += hc[b]; }
  const bool has2 = lo <= 2 && hi > 2;
  const uint64_t n = total + (has2 ? 1 : 0);
  if (!e) e = hipMalloc(&out, std::max<uint64_t>(n, 1) * esz);
  if (!e) e = hipMemcpyAsync(off, ho.data(), nseg * 8, hipMemcpyHostToDevice, c->stream);
  if (!e) e = launch_bit_extract(base, hi, words, nwords, off, (char *)out + (has2 ? esz : 0), out32, c->stream);
  const uint64_t two = 2;
  if (!e && has2) e = hipMemcpyAsync(out, &two, esz, hipMemcpyHostToDevice, c->stream);   // little-endian
  if (!e) e = hipStreamSynchronize(c->stream);
  if (words) (void)hipFree(words);
  if (cnt) (void)hipFree(cnt);
  if (off) (void)hipFree(off);
  if (e) {
    if (out) (void)hipFree(out);
    (void)hipGetLastError();
    return hip_fail(c, e, "prime range sieve");
  }
  *d_out = out;
  *len = n;
  return CPX_OK;
}

// Primes of [lo, hi) from the pi table (which then reaches hi).
static int table_primes(cpx_ctx *c, uint64_t lo, uint64_t hi, uint64_t **d_out, size_t *len)
{
  int rc = ensure_tables(c, hi + 1, 64, 64);
  if (rc) return rc;
  uint64_t cnt = 0;
  if ((rc = cpx_count_primes(c, lo, hi, &cnt))) return rc;
  uint64_t *d = nullptr;
  HIPCHK(hipMalloc(&d, std::max<uint64_t>(cnt, 1) * 8));
  hipError_t e = launch_extract_range_u64(c->pitab.p, lo, hi, d, std::max<uint64_t>(cnt, 1), c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (e) { (void)hipFree(d); return hip_fail(c, e, "getDevicePrimes"); }
  *d_out = d;
  *len = cnt;
  return CPX_OK;
}

int cpx_get_device_primes(cpx_ctx *c, uint64_t lo, uint64_t hi, uint64_t **d_out, size_t *len)
{
  CTX_GUARD(c);
  if (!d_out || !len) return CPX_EINVAL;
  *d_out = nullptr; *len = 0;
  if (hi <= lo) return CPX_OK;
  HIPCHK(hipSetDevice(c->device));
  if (hi <= RANGE_TABLE_MAX || (c->pitab.p && hi - 1 <= c->pitab_limit))
    return table_primes(c, lo, hi, d_out, len);
  void *d = nullptr;
  int rc = range_primes(c, lo, hi, false, &d, len);
  if (!rc) *d_out = (uint64_t *)d;
  return rc;
}

int cpx_get_device_primes32(cpx_ctx *c, uint64_t lo, uint64_t hi, uint32_t **d_out, size_t *len)
{
  CTX_GUARD(c);
  if (!d_out || !len) return CPX_EINVAL;
  *d_out = nullptr; *len = 0;
  if (hi > (1ull << 32)) return ctx_fail(c, CPX_ERANGE, "getDevicePrimes32: primes above 2^32");
  if (hi <= lo) return CPX_OK;
  HIPCHK(hipSetDevice(c->device));
  void *d = nullptr;
  int rc = range_primes(c, lo, hi, true, &d, len);
  if (!rc) *d_out = (uint32_t *)d;
  return rc;
}

int cpx_get_host_primes(cpx_ctx *c, uint64_t lo, uint64_t hi, uint64_t *out, size_t cap, size_t *len)
{
  CTX_GUARD(c);
  if (!len) return CPX_EINVAL;
  *len = 0;
  if (hi <= lo) return CPX_OK;
  if (!out) {
    uint64_t n = 0;
    int rc = cpx_count_primes(c, lo, hi, &n);
    if (!rc) *len = n;
    return rc;
  }
  uint64_t *d = nullptr;
  size_t n = 0;
  int rc = cpx_get_device_primes(c, lo, hi, &d, &n);
  if (rc) return rc;
  *len = n;
  hipError_t e = n <= cap ? hipMemcpy(out, d, n * 8, hipMemcpyDeviceToHost) : hipSuccess;
  (void)hipFree(d);
  if (e) return hip_fail(c, e, "getHostPrimes copy");
  return n <= cap ? CPX_OK : ctx_fail(c, CPX_EINVAL, "getHostPrimes: output buffer too small");
}

int cpx_gen_device_bit_sieve(cpx_ctx *c, uint64_t lo, uint64_t hi, uint32_t **d_words, size_t *nwords)
{
  CTX_GUARD(c);
  if (!d_words || !nwords || (lo & 63)) return CPX_EINVAL;
  *d_words = nullptr; *nwords = 0;
  if (hi <= lo) return CPX_OK;
  if (hi > RANGE_HI_MAX) return ctx_fail(c, CPX_ERANGE, "bit sieve above 2^64 - 2^34");
  HIPCHK(hipSetDevice(c->device));
  const uint64_t nw = (hi - lo + 63) / 64, nseg = range_sieve_segments(lo, hi);
  const uint64_t r = cpx_isqrt128((u128)hi + (1ull << 21)) + 1;
  int rc = ensure_tables(c, r, r, 64);
  if (rc) return rc;
  uint32_t *w = nullptr, *cnt = nullptr;
  hipError_t e = hipMalloc(&w, nw * 4);
  if (!e) e = hipMalloc(&cnt, nseg * 4);
  if (!e) e = launch_bit_sieve(lo, lo, hi, nw, c->primes.p, (uint32_t)c->nprimes, true, w, cnt, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (cnt) (void)hipFree(cnt);
  if (e) {
    if (w) (void)hipFree(w);
    (void)hipGetLastError();
    return hip_fail(c, e, "genDeviceBitSieve");
  }
  *d_words = w;
  *nwords = nw;
  return CPX_OK;
}

// --------------------------------------------------------------- batch

static int batch_impl(cpx_ctx *c, const cpx_u128 *xs, size_t n, cpx_u128 *out, cpx_terms *terms,
                      const cpx_batch_opts *bo, cpx_batch_stats *st)
{
  if ((!xs || !out) && n) return CPX_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  if (bo && bo->tune && !c->shape_tuned) {
    uint32_t tt, kk;
    int rc = cpx_autotune(c, xs, n, &tt, &kk);
    if (rc) return rc;
  }
  std::vector<HostQuery> hq(n);
  for (size_t i = 0; i < n; i++) {
    int rc = make_host_query(c, to_u128(xs[i]), 0, 0, &hq[i]);
    if (rc) return rc;
  }
  std::vector<ResultDev> res;
  bool graph = bo ? bo->use_graph != 0 : true;
  cpx_batch_stats local{};
  int rc = run_queries(c, hq, TERM_ALL, bo ? bo->p2_table : 0, res, st ? &local : nullptr, graph);
  if (rc) return rc;
  for (size_t i = 0; i < n; i++) {
    out[i] = {res[i].pi[0], res[i].pi[1]};
    if (terms) fill_terms(&terms[i], res[i], hq[i], 0);
  }
  if (st) *st = local;
  return CPX_OK;
}

int cpx_pi_batch(cpx_ctx *c, const cpx_u128 *xs, size_t n, cpx_u128 *out,
                 const cpx_batch_opts *bo, cpx_batch_stats *st)
{
  CTX_GUARD(c);
  return batch_impl(c, xs, n, out, nullptr, bo, st);
}

int cpx_pi_batch_terms(cpx_ctx *c, const cpx_u128 *xs, size_t n, cpx_u128 *out, cpx_terms *terms,
                       const cpx_batch_opts *bo, cpx_batch_stats *st)
{
  CTX_GUARD(c);
  return batch_impl(c, xs, n, out, terms, bo, st);
}

// ----------------------------------------------------- pipelined batches

int cpx_pi_batch_submit(cpx_ctx *c, const cpx_u128 *xs, size_t n, const cpx_batch_opts *bo,
                        uint64_t *ticket)
{
  CTX_GUARD(c);
  if (!xs || !n || !ticket) return CPX_EINVAL;
  *ticket = 0;
  HIPCHK(hipSetDevice(c->device));
  bool idle = true;
  for (const BatchSlot &S : c->slot) idle = idle && !S.busy;
  if (bo && bo->tune && !c->shape_tuned && idle) {
    uint32_t tt, kk;
    int rc = cpx_autotune(c, xs, n, &tt, &kk);
    if (rc) return rc;
  }
  std::vector<HostQuery> hq(n);
  for (size_t i = 0; i < n; i++) {
    int rc = make_host_query(c, to_u128(xs[i]), 0, 0, &hq[i]);
    if (rc) return rc;
  }
  uint32_t si = 0;
  int rc = submit_queries(c, hq, TERM_ALL, bo ? bo->p2_table : 0, bo ? bo->use_graph != 0 : true, &si);
  if (rc) return rc;
  *ticket = c->slot[si].ticket;
  return CPX_OK;
}

int cpx_pi_batch_collect(cpx_ctx *c, uint64_t ticket, cpx_u128 *out, cpx_batch_stats *st)
{
  CTX_GUARD(c);
  if (!out) return CPX_EINVAL;
  uint32_t si = CPX_SLOTS;
  for (uint32_t i = 0; i < CPX_SLOTS; i++)
    if (c->slot[i].busy && c->slot[i].ticket == ticket) si = i;
  if (si == CPX_SLOTS) return CPX_ENOTFOUND;
  HIPCHK(hipSetDevice(c->device));
  std::vector<ResultDev> res;
  cpx_batch_stats local{};
  int rc = collect_queries(c, si, res, st ? &local : nullptr);
  if (rc) return rc;
  for (size_t i = 0; i < res.size(); i++) out[i] = {res[i].pi[0], res[i].pi[1]};
  if (st) *st = local;
  return CPX_OK;
}

// ------------------------------------------------------------ autotuner
// Times one calibration batch per candidate and keeps the fastest, in three
// passes: the segment shape, then a factor on the alpha model, then the
// number of wheel primes c.  Every candidate must return the same pi(x) for
// every query (bit for bit), or the tuner fails.
int cpx_autotune(cpx_ctx *c, const cpx_u128 *xs, size_t n, uint32_t *threads, uint32_t *kwords)
{
  CTX_GUARD(c);
  if (!xs || !n) return CPX_EINVAL;
  static const SegShape cands[] = {{64, 32}, {64, 64}, {64, 128}, {128, 32}, {128, 64}, {256, 32},
                                   {256, 64}, {256, 128}, {512, 32}, {512, 64}};
  static const double scales[] = {0.15, 0.2, 0.25, 0.3, 0.4, 0.5, 0.6, 0.7, 0.85, 1.0, 1.2, 1.45};
  static const uint32_t wheels[] = {6, 7, 8};
  std::vector<HostQuery> hq(n);
  auto make = [&]() -> int {
    for (size_t i = 0; i < n; i++) {
      int rc = make_host_query(c, to_u128(xs[i]), 0, 0, &hq[i]);
      if (rc) return rc;
    }
    return CPX_OK;
  };
  std::vector<ResultDev> ref, res;
  bool have_ref = false;
  // time one configuration (after a warm-up run for tables and arena)
  auto trial = [&](double &ms) -> int {
    int rc = make();
    if (!rc) rc = run_queries(c, hq, TERM_ALL, 0, res, nullptr, false);
    if (rc) return rc;
    const double t0 = now_ms();
    if ((rc = run_queries(c, hq, TERM_ALL, 0, res, nullptr, false))) return rc;
    ms = now_ms() - t0;
    if (!have_ref) { ref = res; have_ref = true; }
    for (size_t i = 0; i < n; i++)
      if (memcmp(res[i].pi, ref[i].pi, 16) != 0)
        return ctx_fail(c, CPX_EDEVICE, "autotune: candidates disagree");
    return CPX_OK;
  };
  double best_ms = 1e300, ms = 0;
  SegShape best = c->shape;
  for (const SegShape &s : cands) {
    if (!seg_shape_supported(s) || seg_blocks_per_cu(s) < 1) continue;
    c->shape = s;
    int rc = trial(ms);
    if (rc) return rc;
    if (ms < best_ms) { best_ms = ms; best = s; }
  }
  c->shape = best;
  if (c->opts.alpha <= 0) {
    double best_a = c->alpha_scale;
    best_ms = 1e300;
    for (double a : scales) {
      c->alpha_scale = a;
      int rc = trial(ms);
      if (rc) return rc;
      if (ms < best_ms) { best_ms = ms; best_a = a; }
    }
    c->alpha_scale = best_a;
  }
  {
    uint32_t best_c = c->opts.c;
    best_ms = 1e300;
    for (uint32_t w : wheels) {
      c->opts.c = w;
      int rc = trial(ms);
      if (rc) return rc;
      if (ms < best_ms) { best_ms = ms; best_c = w; }
    }
    c->opts.c = best_c;
  }
  c->shape_tuned = true;
  c->generation++;
  if (threads) *threads = (uint32_t)best.threads;
  if (kwords) *kwords = (uint32_t)best.kwords;
  return CPX_OK;
}

// Set the tuning the autotuner would pick (for replaying a tuned
// configuration without re-timing it).
int cpx_set_tuning(cpx_ctx *c, uint32_t threads, uint32_t kwords, double alpha_scale, uint32_t wheel_c)
{
  CTX_GUARD(c);
  const SegShape s{(int)threads, (int)kwords};
  if (!seg_shape_supported(s)) return ctx_fail(c, CPX_EINVAL, "unsupported segment shape");
  if (!(alpha_scale > 0.0 && alpha_scale < 64.0)) return ctx_fail(c, CPX_EINVAL, "alpha_scale must be in (0, 64)");
  if (wheel_c < 1 || wheel_c > 8) return ctx_fail(c, CPX_EINVAL, "c must be in 1..8");
  c->shape = s;
  c->alpha_scale = alpha_scale;
  c->opts.c = wheel_c;
  c->shape_tuned = true;
  c->generation++;
  return CPX_OK;
}

int cpx_get_tuning(cpx_ctx *c, uint32_t *threads, uint32_t *kwords, double *alpha_scale, uint32_t *wheel_c)
{
  CTX_GUARD(c);
  if (threads) *threads = (uint32_t)c->shape.threads;
  if (kwords) *kwords = (uint32_t)c->shape.kwords;
  if (alpha_scale) *alpha_scale = c->alpha_scale;
  if (wheel_c) *wheel_c = c->opts.c;
  return CPX_OK;
}

uint64_t cpx_request_bytes(cpx_ctx *c, cpx_u128 x)
{
  if (!c) return 0;
  std::lock_guard<std::recursive_mutex> g(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return 0;
  HostQuery h;
  if (make_host_query(c, to_u128(x), 0, 0, &h)) return 0;
  std::vector<HostQuery> v(1, h);
  uint64_t b = 0;
  if (stream_need_bytes(c, v, &b)) return 0;
  return b;
}

// ------------------------------------------------------------ selftests

int cpx_selftest_arith(uint64_t seed, uint64_t n)
{
  std::mt19937_64 g(seed);
  int fails = 0;
  for (uint64_t i = 0; i < n; i++) {
    int bits = 1 + (int)(g() % 100);
    u128 a = ((u128)g() << 64) | g();
    if (bits < 128) a &= (((u128)1 << bits) - 1);
    uint64_t d = g() >> (g() % 64);
    if (d == 0) d = 1;
    uint64_t r;
    u128 q = cpx_div128_64(a, d, &r);
    if (q != a / d || r != (uint64_t)(a % d)) fails++;
    if ((a / d) >> 64 == 0) {
      if (cpx_div128_to64(a, d) != (uint64_t)(a / d)) fails++;
    }
    uint64_t a64 = (uint64_t)a, d53 = d >> 11;
    if (d53 == 0) d53 = 1;
    if (a64 / d53 < (1ull << 50)) {
      if (cpx_div_smallq(a64, d53) != a64 / d53) fails++;
    }
    uint32_t p = (uint32_t)(g() | 1);
    if (cpx_mod64_32(a64, p) != a64 % p) fails++;
    u128 xr = a >> (g() % 64);
    uint64_t s = cpx_isqrt128(xr);
    if (!((u128)s * s <= xr && ((u128)s + 1) * ((u128)s + 1) > xr)) fails++;
    uint64_t cb = cpx_icbrt128(xr);
    if (!((u128)cb * cb * cb <= xr && ((u128)cb + 1) * (cb + 1) * (cb + 1) > xr)) fails++;
  }
  return fails;
}

}  // extern "C"

int cpx_selftest_wave(cpx_ctx *c, uint64_t seed, uint32_t nwaves, uint32_t *bad)
{
  CTX_GUARD(c);
  if (!bad || nwaves == 0 || nwaves > (1u << 20)) return CPX_EINVAL;
  *bad = 0;
  HIPCHK(hipSetDevice(c->device));
  std::mt19937_64 g(seed);
  std::vector<uint32_t> h((size_t)nwaves * 64);
  for (auto &v : h) v = (uint32_t)g() >> (g() % 32);      // values up to 2^32 exercise the wrap
  uint32_t *d = nullptr;
  HIPCHK(hipMalloc(&d, (h.size() + 1) * sizeof(uint32_t)));
  int rc = CPX_OK;
  do {
    hipError_t e = hipMemcpyAsync(d, h.data(), h.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (!e) e = launch_zero_u32(d + h.size(), 1, c->stream);
    if (!e) e = launch_selftest_wave(d, nwaves, d + h.size(), c->stream);
    uint32_t nb = 0;
    if (!e) e = hipMemcpyAsync(&nb, d + h.size(), 4, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    if (e) { rc = hip_fail(c, e, "selftest_wave"); break; }
    *bad = nb;
  } while (0);
  (void)hipFree(d);
  return rc;
}

// ------------------------------------------------------------ peak probes
hipError_t probe_lds(uint32_t blocks, uint32_t iters, uint32_t *sink, hipStream_t s);
hipError_t probe_valu(uint32_t blocks, uint32_t iters, uint32_t *sink, hipStream_t s);

int cpx_probe_peaks(cpx_ctx *c, double *lds_bytes_per_s, double *int32_ops_per_s)
{
  CTX_GUARD(c);
  if (!lds_bytes_per_s || !int32_ops_per_s) return CPX_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  const uint32_t blocks = 8u * (uint32_t)std::max(c->cu_count, 1);
  const uint32_t lds_iters = 4096, valu_iters = 16384;
  uint32_t *sink = nullptr;
  HIPCHK(hipMalloc(&sink, blocks * sizeof(uint32_t)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CPX_OK;
  do {
    hipError_t e = hipEventCreate(&e0);
    if (!e) e = hipEventCreate(&e1);
    float ms = 0;
    // LDS: 8 x 16-byte reads per thread per iteration (plus one warm-up launch)
    if (!e) e = probe_lds(blocks, lds_iters, sink, c->stream);
    if (!e) e = hipEventRecord(e0, c->stream);
    if (!e) e = probe_lds(blocks, lds_iters, sink, c->stream);
    if (!e) e = hipEventRecord(e1, c->stream);
    if (!e) e = hipEventSynchronize(e1);
    if (!e) e = hipEventElapsedTime(&ms, e0, e1);
    if (e) { rc = hip_fail(c, e, "probe_lds"); break; }
    *lds_bytes_per_s = (double)blocks * 256.0 * lds_iters * 8.0 * 16.0 / (ms * 1e-3);
    // VALU: 24 int32 instructions per lane per iteration (8 x add, shift, xor)
    if (!e) e = probe_valu(blocks, valu_iters, sink, c->stream);
    if (!e) e = hipEventRecord(e0, c->stream);
    if (!e) e = probe_valu(blocks, valu_iters, sink, c->stream);
    if (!e) e = hipEventRecord(e1, c->stream);
    if (!e) e = hipEventSynchronize(e1);
    if (!e) e = hipEventElapsedTime(&ms, e0, e1);
    if (e) { rc = hip_fail(c, e, "probe_valu"); break; }
    *int32_ops_per_s = (double)blocks * 256.0 * valu_iters * 24.0 / (ms * 1e-3);
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(sink);
  return rc;
}
