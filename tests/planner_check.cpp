// planner_check.cpp -- host-only check of the batch planner (CPU test, no GPU).
//
// Compiles the engine's host code (engine.hip, host pass only) with every
// kernel launcher stubbed out, fakes a context whose host prime list and
// table limits are what the resident tables would hold, and for each batch
// read from stdin runs the planner and the plan packing into an image of
// the plan arena exactly as large as its layout.  Built with AddressSanitizer
// by tests/test_planner_host.py, so an out-of-bounds host access in planning
// or packing fails the run; on top of that every index the kernels will form
// from the plan is checked against the arena capacities.
//
// stdin: one batch per line, "[pass=<bytes>] mask p2mode x[:y] x[:y] ..."
// (decimal x < 2^100; pass= sets cpx_opts.pass_bytes, so a single query
// whose hard rows exceed it is planned in passes over segment ranges, whose
// windows are checked as the kernels will index them).
// stdout: one line per batch, "ok <queries> <tiles> <plan_bytes> <out_bytes>
// <estimate_bytes_sum> <passes>", or "FAIL <what>" (and exit status 1).
#define CPX_NO_ROCTX 1
#include "../cudapix_amd/csrc/engine.hip"
#include <cstdio>
#include <iostream>
#include <sstream>

// ---- launchers: planning never launches anything
#define STUB(sig) sig { std::abort(); }
STUB(hipError_t launch_mlpf(int32_t *, uint64_t, const uint32_t *, uint32_t, hipStream_t))
STUB(hipError_t launch_recip(const uint32_t *, double *, uint64_t, hipStream_t))
STUB(hipError_t launch_setup(BatchDev, uint32_t, uint32_t, hipStream_t))
STUB(hipError_t launch_wheel(PiEntry *, uint32_t, uint32_t, hipStream_t))
STUB(hipError_t launch_combine(BatchDev, uint32_t, hipStream_t))
STUB(hipError_t launch_trivial(BatchDev, uint32_t, hipStream_t))
STUB(hipError_t launch_ordinary(BatchDev, uint32_t, hipStream_t))
STUB(hipError_t launch_segments(BatchDev, SegShape, uint32_t, hipStream_t))
STUB(hipError_t launch_zero_u32(uint32_t *, size_t, hipStream_t))
STUB(int seg_blocks_per_cu(SegShape))
STUB(hipError_t launch_small_sieve(PiEntry *, uint64_t, hipStream_t))
STUB(hipError_t launch_scan_entries(PiEntry *, uint64_t, uint32_t *, uint64_t *, hipStream_t))
STUB(hipError_t launch_hard_finalize(BatchDev, uint32_t, hipStream_t))
STUB(hipError_t launch_extract_primes(const PiEntry *, uint64_t, uint32_t *, uint64_t, hipStream_t))
STUB(hipError_t launch_pitab_segments(PiEntry *, uint64_t, const uint32_t *, uint32_t, uint32_t *,
                                      uint64_t *, hipStream_t))

static std::vector<uint32_t> sieve_primes(uint64_t n)
{
  std::vector<uint8_t> comp(n + 1, 0);
  std::vector<uint32_t> pr;
  for (uint64_t i = 2; i <= n; i++) {
    if (comp[i]) continue;
    pr.push_back((uint32_t)i);
    for (uint64_t j = i * i; j <= n; j += i) comp[j] = 1;
  }
  return pr;
}

static u128 parse_u128(const std::string &s)
{
  u128 v = 0;
  for (char ch : s) v = v * 10 + (u128)(ch - '0');
  return v;
}

#define CHECK(cond, what)                                                    \
  do {                                                                       \
    if (!(cond)) { std::printf("FAIL %s\n", what); std::exit(1); }           \
  } while (0)

// What ensure_tables would make resident for these needs (the same limits,
// without the growth headroom of a warm context).
static void fake_tables(cpx_ctx *c, uint64_t R, uint64_t P, uint64_t Y)
{
  if (Y < 64) Y = 64;
  uint64_t Pm = std::max<uint64_t>(P, cpx_isqrt128((u128)Y + 16384) + 1);
  Pm = std::max<uint64_t>(Pm, Y);
  uint64_t Rm = std::max<uint64_t>({R, Pm, (uint64_t)1 << 14});
  if (c->prime_limit < Pm) {
    c->h_primes = sieve_primes(Pm);
    c->prime_limit = Pm;
    c->nprimes = c->h_primes.size();
  }
  uint64_t nent = ((Rm + 1 + 63) / 64 + 8191) / 8192 * 8192;
  c->pitab_limit = std::max<uint64_t>(c->pitab_limit, nent * 64 - 1);
  c->mlpf_limit = std::max<uint64_t>(c->mlpf_limit, (Y + 1) / 2 * 2 + 1);
}

static void check_plan(const cpx_ctx *c, const Plan &pl, const Layout &L, const Caps &k, bool passes)
{
  const uint64_t nq = pl.q.size();
  for (size_t i = 0; i < nq; i++) {
    const QueryDev &d = pl.q[i];
    const QPlan &qp = pl.qps[i];
    CHECK(d.part_off + qp.parts <= k.parts, "parts");
    CHECK(d.stage_off + d.nstage <= k.stages, "stages");
    CHECK(d.hard_row0 + d.hard_nseg <= k.rows, "rows");
    CHECK(d.p2_row0 + d.p2_nseg <= k.p2rows, "p2 rows");
    CHECK(d.mrow_base + d.hard_nmrow <= k.mrows, "m rows");
    if (!passes) {                    // (in passes the T and M arrays hold one range's rows)
      CHECK(d.tm_off + qp.tm <= k.tm, "T entries");
      CHECK(d.m_off + qp.mm <= k.mm, "M entries");
    }
    CHECK(qp.rowinfo.size() == d.hard_nseg, "row count");
    CHECK(d.n_ord_parts + d.n_triv_parts + d.n_easy_parts + d.n_hard_parts + d.n_p2_parts == qp.parts,
          "part split");
    if (d.p2_nseg) CHECK(d.p2_seg0 + d.p2_nseg <= d.hard_nseg, "P2 segments inside the tiles");
    uint32_t prev_act = ~0u;
    for (uint64_t s = 0; s < d.hard_nseg; s++) {
      const HardRowInfo &ri = qp.rowinfo[s];
      CHECK(ri.act <= d.nstage, "act <= stages");
      CHECK(ri.act <= prev_act, "act non-increasing");
      prev_act = ri.act;
      CHECK(ri.toff + ri.act <= qp.tm, "T row inside the query");
      CHECK(ri.nrep >= 1 && ri.nrep <= 64, "replicas");
      CHECK(ri.mrow + ri.nrep <= d.hard_nmrow, "M row inside the query");
      CHECK(ri.moff + (uint64_t)ri.nrep * ri.act <= qp.mm, "M entries inside the query");
    }
    for (const BlockDesc &b : qp.ord) CHECK(b.query == i && b.part < d.n_ord_parts && b.end <= (d.y + 1) / 2, "ord desc");
    for (const BlockDesc &b : qp.easy) CHECK(b.query == i && b.end <= c->nprimes, "easy desc");
    for (const BlockDesc &b : qp.p2d) CHECK(b.query == i && b.end <= c->nprimes, "p2 desc");
    for (const BlockDesc &b : qp.triv) CHECK(b.query == i && b.end <= c->nprimes, "triv desc");
    for (const BlockDesc &b : qp.fin) CHECK(b.query == i && b.end <= d.nstage, "fin desc");
    if (d.nstage) CHECK(d.c + d.nstage <= c->nprimes, "stage primes in the prime list");
  }
  for (const GroupDev &g : pl.groups) {
    CHECK(g.n >= 1 && g.n <= CPX_GROUP, "group size");
    for (uint32_t m = 0; m < g.n; m++) CHECK(g.q[m] < nq, "group member");
  }
  for (const TileDev &t : pl.tiles) {
    if (t.group == CPX_TILE_EASY) { CHECK(t.seg < pl.neasy, "easy tile"); continue; }
    if (t.group == CPX_TILE_P2) { CHECK(t.seg < pl.np2d, "p2 tile"); continue; }
    CHECK(t.group < pl.groups.size(), "tile group");
    CHECK(t.rep < t.nrep, "tile replica");
    const GroupDev &g = pl.groups[t.group];
    uint64_t nseg = 0;
    for (uint32_t m = 0; m < g.n; m++) {
      const QueryDev &d = pl.q[g.q[m]];
      nseg = std::max<uint64_t>(nseg, d.hard_nseg);
      if (t.seg < d.hard_nseg) CHECK(pl.qps[g.q[m]].rowinfo[t.seg].nrep == t.nrep, "member replicas");
    }
    CHECK(t.seg < nseg, "tile segment");
  }
  CHECK(pl.tiles.size() <= k.tiles && pl.fin.size() <= k.fin && pl.groups.size() <= k.groups, "list caps");
  CHECK(L.plan_bytes > 0, "layout");
}

// A run in passes: the ranges cover the segments in order; every tile sits
// in the pass of its segment (term tiles in the first), once; and every row
// a pass's tiles and carry blocks touch, T at tm_off + toff - pass_toff + b
// and M at m_off + moff - pass_moff + r * act + b (b < act, r < nrep), lies
// inside the arrays sized for the largest range.
static void check_passes(const Plan &pl, const PassPlan &pp, const Caps &k)
{
  CHECK(pl.qps.size() == 1, "passes: one query");
  const QueryDev &d = pl.q[0];
  const QPlan &qp = pl.qps[0];
  const size_t np = pp.n();
  CHECK(np >= 2 && pp.s.front() == 0 && pp.s.back() == d.hard_nseg, "passes: cover the segments");
  CHECK(pp.toff.size() == np && pp.moff.size() == np && pp.tstart.size() == np && pp.tcnt.size() == np,
        "passes: per-pass records");
  uint64_t ntile = 0;
  for (size_t p = 0; p < np; p++) {
    CHECK(pp.s[p] < pp.s[p + 1], "passes: ranges ascend");
    CHECK(pp.tstart[p] == ntile, "passes: tiles contiguous");
    ntile += pp.tcnt[p];
    for (uint64_t s = pp.s[p]; s < pp.s[p + 1]; s++) {
      const HardRowInfo &ri = qp.rowinfo[s];
      CHECK(ri.toff >= pp.toff[p] && d.tm_off + ri.toff - pp.toff[p] + ri.act <= k.tm, "passes: T window");
      CHECK(ri.moff >= pp.moff[p] && d.m_off + ri.moff - pp.moff[p] + (uint64_t)ri.nrep * ri.act <= k.mm,
            "passes: M window");
    }
    for (uint64_t i = pp.tstart[p]; i < pp.tstart[p] + pp.tcnt[p]; i++) {
      const TileDev &t = pl.tiles[i];
      if (t.group == CPX_TILE_EASY || t.group == CPX_TILE_P2) CHECK(p == 0, "passes: term tiles first");
      else CHECK(t.seg >= pp.s[p] && t.seg < pp.s[p + 1], "passes: tile in its range");
    }
  }
  CHECK(ntile == pl.tiles.size(), "passes: every tile once");
  CHECK(k.base >= d.nstage, "passes: stage totals");
}

int main()
{
  cpx_ctx ctx;
  cpx_ctx *c = &ctx;
  c->opts.c = 6;
  c->opts.alpha = 0;
  c->opts.fresh_tables = 0;
  c->shape = {512, 32};
  std::string line;
  while (std::getline(std::cin, line)) {
    if (line.empty()) continue;
    std::istringstream in(line);
    uint32_t mask, p2mode;
    c->opts.pass_bytes = 0;
    if (line.compare(0, 5, "pass=") == 0) {
      std::string pb;
      in >> pb;
      c->opts.pass_bytes = (uint64_t)parse_u128(pb.substr(5));
    }
    in >> mask >> p2mode;
    std::vector<HostQuery> hq;
    std::string tok;
    while (in >> tok) {
      const size_t colon = tok.find(':');
      const u128 x = parse_u128(tok.substr(0, colon));
      const uint64_t y = colon == std::string::npos ? 0 : (uint64_t)parse_u128(tok.substr(colon + 1));
      HostQuery h;
      int rc = make_host_query(c, x, y, 0, &h);
      CHECK(rc == CPX_OK, "make_host_query");
      hq.push_back(h);
    }
    choose_p2tab(c, hq, mask, p2mode);
    uint64_t R = 0, P = 0, Y = 0;
    for (const HostQuery &h : hq) plan_needs(c, h, mask, c->shape.seg_bits(), &R, &P, &Y);
    fake_tables(c, R, P, Y);
    Plan pl;
    CHECK(build_plan(c, hq, mask, pl) == CPX_OK, "build_plan");
    PassPlan pp;
    const bool passes = plan_passes(c, pl, pp);
    Caps need = needed_caps(pl);
    if (passes) {                     // as submit_impl asks for the arena
      need.tm = pp.tm_max;
      need.mm = pp.mm_max;
      need.base = pl.q[0].nstage;
    }
    Caps k{};                         // and as a fresh slot's arena grows to it
    grown_caps(Caps{}, need, k);
    const Layout L = make_layout(k);
    std::vector<char> image(L.plan_bytes);
    pack_plan(L, pl, image.data());
    check_plan(c, pl, L, k, passes);
    if (passes) check_passes(pl, pp, k);
    uint64_t est = 0;
    for (const HostQuery &h : hq) est += plan_bytes_estimate(c, h);
    const Layout Ln = make_layout(need);   // the bytes the plan itself needs
    std::printf("ok %zu %zu %zu %zu %llu %zu\n", hq.size(), pl.tiles.size(), Ln.plan_bytes, Ln.out_bytes,
                (unsigned long long)est, passes ? pp.n() : (size_t)1);
  }
  return 0;
}
