"""Batched queries: each result equals the same x solved alone (tolerance 0),
against the GPU brute count where it reaches, graphs and direct launches
agree, and all segment shapes agree.  (Served queries: test_gpu_server.py.)"""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def c3(n=None):
    with open(os.path.join(GOLD, "queries_c3.json")) as f:
        xs = json.load(f)["x"]
    return xs if n is None else xs[:n]


def test_batch_equals_alone(ctx, oracle):
    xs = sorted(c3(64))[:48] + [10 ** 9, 10 ** 12, 16384, 5, 10 ** 5]
    got = ctx.pi_batch(xs)
    alone = [ctx.pi(x) for x in xs]
    assert got == alone
    for x, v in zip(xs, got):
        if x <= 3 * 10 ** 9:
            want = oracle.count_primes(0, x + 1)
            assert v == want, x


def test_batch_c3_sample_against_gpu_brute_count(ctx):
    """A stratified sample of the 512-query workload above 3e9 (one x per
    decade up to 1e13) against the plain GPU segmented-Eratosthenes count."""
    xs = c3()
    pick = []
    for lo, hi in ((3 * 10 ** 9, 10 ** 10), (10 ** 10, 10 ** 11), (10 ** 11, 10 ** 12),
                   (10 ** 12, 10 ** 13)):
        inside = sorted(x for x in xs if lo < x <= hi)
        pick += [inside[0], inside[len(inside) // 2]]
    got = ctx.pi_batch(pick)
    for x, v in zip(pick, got):
        print("case brute x=%d" % x, flush=True)
        assert v == ctx.countPrimes(0, x + 1), x


def test_full_c3_batch_equals_plain_solves(ctx):
    """All 512 queries in one persistent launch, each equal to its plain
    one-at-a-time solve."""
    xs = c3()
    assert len(xs) == 512
    got = ctx.pi_batch(xs)
    assert got == [ctx.pi(x) for x in xs]


def test_batch_graph_vs_direct_and_repeat(ctx):
    xs = c3(128)
    a = ctx.pi_batch(xs, use_graph=True)
    b = ctx.pi_batch(xs, use_graph=False)
    c = ctx.pi_batch(xs, use_graph=True)
    assert a == b == c
    # a smaller batch after a larger one replays the same graph
    d = ctx.pi_batch(xs[:7], use_graph=True)
    assert d == a[:7]


def test_batch_terms_match_oracle(ctx, oracle):
    xs = [x for x in c3() if x < 3 * 10 ** 10][:12]
    terms = ctx.pi_batch(xs, terms=True)
    for x, t in zip(xs, terms):
        ref = oracle.pi_dr(x, t["y"], int(t["c"]))
        for k in ("ordinary", "trivial", "easy", "hard", "p2", "pi"):
            assert t[k] == ref[k], (x, k)


def test_pipelined_batches(oracle):
    """Two batches in flight: each collected result equals the batch solved
    with one synchronous call; tickets collect in any order; a third
    submission, or a synchronous batch while both slots are busy, is refused
    with CPX_ESTATE; an unknown ticket gives CPX_ENOTFOUND."""
    import cudapix_amd as C
    with C.Context(device=0) as ctx:     # its own slots: a failure leaves no batch in flight elsewhere
        _pipelined(C, ctx, oracle)


def _pipelined(C, ctx, oracle):
    xs = c3()
    b1, b2, b3 = sorted(xs[:40]), xs[40:100], [10 ** 9, 2 ** 32, 16384, 5] + xs[100:120]
    want = [ctx.pi_batch(b) for b in (b1, b2, b3)]
    t1 = ctx.pi_batch_submit(b1)
    t2 = ctx.pi_batch_submit(b2)
    with pytest.raises(C.CpxError) as e:
        ctx.pi_batch_submit(b3)
    assert e.value.code == -8
    with pytest.raises(C.CpxError) as e:
        ctx.pi_batch(b3)
    assert e.value.code == -8
    out = (C.CpxU128 * 4)()
    assert ctx._L.cpx_pi_batch_collect(ctx.h, C.U64(t2 + 1000), out, None) == -7
    assert ctx.pi_batch_collect(t2) == want[1]          # out of order
    t3 = ctx.pi_batch_submit(b3)
    r1, st = ctx.pi_batch_collect(t1, stats=True)
    assert r1 == want[0] and st["ms_device"] > 0 and st["tiles"] > 0
    assert ctx.pi_batch_collect(t3) == want[2]
    assert ctx._L.cpx_pi_batch_collect(ctx.h, C.U64(t3), out, None) == -7   # collected already
    for x, v in zip(b3, want[2]):
        if x <= 3 * 10 ** 9:
            assert v == oracle.count_primes(0, x + 1), x
    # a stream of batches of varied sizes, pipelined two deep, also after a
    # larger batch grew the slots
    batches = [b2, b1[:3], b3, b1, b2[:1]]
    assert list(ctx.pi_batches(batches)) == [ctx.pi_batch(b) for b in batches]
    assert list(ctx.pi_batches(batches, use_graph=False)) == [ctx.pi_batch(b) for b in batches]
    # a consumer that stops after the first result leaves no batch in flight
    gen = ctx.pi_batches(batches)
    first = next(gen)                    # two submitted, the first collected
    gen.close()                          # the generator collects the second
    t1, t2 = ctx.pi_batch_submit(b3), ctx.pi_batch_submit(b1)   # both slots free again
    assert ctx.pi_batch_collect(t1) == want[2] and ctx.pi_batch_collect(t2) == want[0]
    assert first == ctx.pi_batch(batches[0])
    # a submission that grows the resident tables while the other slot's
    # batch is still queued (the old tables are freed only after it ran)
    big = [10 ** 16, 10 ** 16 + 1, 3 * 10 ** 15]
    ta = ctx.pi_batch_submit(b1)
    tb = ctx.pi_batch_submit(big)
    assert ctx.pi_batch_collect(ta) == want[0]
    got = ctx.pi_batch_collect(tb)
    assert got[0] == 279238341033925
    assert got == [ctx.pi(x) for x in big]


def test_p2_lookup_equals_p2_segments():
    """P2 from lookups in the resident pi table (a batch of >= 8 queries)
    equals P2 counted in the sieve's P2 segments (p2_table=2), term by term,
    also in one batch with both kinds of query (p2_table=1 after the table
    reached half of them)."""
    import cudapix_amd as C
    xs = [x for x in c3() if 10 ** 11 < x < 10 ** 14][:12]
    assert len(xs) >= 8
    with C.Context(device=0) as a:
        looked = a.pi_batch(xs, terms=True, p2_table=0)
        sieved = a.pi_batch(xs, terms=True, p2_table=2)
        for x, t, u in zip(xs, looked, sieved):
            assert t["y"] == u["y"]
            for k in ("p2", "pi"):
                assert t[k] == u[k], (x, k)
    with C.Context(device=0) as b:
        small = sorted(xs)[:8]
        b.pi_batch(small, p2_table=0)             # the table reaches the smaller half
        mixed = b.pi_batch(xs, terms=True, p2_table=1)
        assert [t["pi"] for t in mixed] == [t["pi"] for t in looked]
        assert [t["p2"] for t in mixed] == [t["p2"] for t in looked]


@pytest.mark.parametrize("shape", [(64, 32), (64, 64), (64, 128), (128, 32), (128, 64), (256, 32),
                                   (256, 64), (512, 32)])
def test_segment_shapes_agree(oracle, shape):
    import cudapix_amd as C
    xs = [10 ** 12, 123456789012, 10 ** 13 + 7]
    with C.Context(device=0, seg_threads=shape[0], seg_kwords=shape[1]) as c:
        got = c.pi_batch(xs)
    assert got[0] == 37607912018
    assert got[2] == oracle.pi_dr(xs[2], oracle.choose_y(xs[2], 2.0), 6)["pi"]


def test_invalid_arguments(ctx):
    import cudapix_amd as C
    with pytest.raises(C.CpxError):
        ctx.pi(10 ** 12, y=10)                  # y below cbrt(x)
    with pytest.raises(C.CpxError):
        ctx.pi(10 ** 12, c=9)
    with pytest.raises(C.CpxError):
        ctx.pi(1 << 101)
