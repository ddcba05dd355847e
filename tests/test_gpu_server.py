"""Served queries through the server (tolerance 0): each
result equals the same x solved alone; requests submitted while a launch runs
join that launch and their tiles run beside the earlier cohort's on the
device; a slice of the served workload (x to 1e17 at 2000 requests/s) is
value-checked, in both server modes; admission rejects a request larger than a slot's share of the
HBM budget and cancels queued ones; the server serves with the tuning of the
context it was created from."""
import json
import os
import random
import time

import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def c4(lam):
    with open(os.path.join(GOLD, "queries_c4.json")) as f:
        st = json.load(f)["streams"][str(lam)]
    return st["x"], st["t_ms"]


def served(srv, xs, ts):
    tickets = srv.replay(xs, ts)
    srv.drain()
    res = {r["ticket"]: r for r in srv.poll(len(xs) + 16)}
    return tickets, res


def test_server_mix_equals_alone(ctx):
    import cudapix_amd as C
    rnd = random.Random(17)
    xs = [int(10 ** (8 + 5 * rnd.random())) for _ in range(120)] + [5, 16383, 16384]
    ts = sorted(rnd.random() * 300.0 for _ in xs)
    alone = {x: ctx.pi(x) for x in set(xs)}
    srv = C.Server(ctx)
    tickets, res = served(srv, xs, ts)
    st = srv.stats()
    srv.close()
    assert len(res) == len(xs)
    for tk, x in zip(tickets, xs):
        r = res[tk]
        assert r["status"] == 0 and r["x"] == x and r["pi"] == alone[x], r
        assert r["t_arrive_ms"] <= r["t_admit_ms"] <= r["t_done_ms"]
    assert st["completed"] == len(xs) and st["p99_ms"] >= st["p50_ms"] > 0
    assert st["launches"] >= 1 and st["batches"] >= 1


def test_requests_join_the_running_launch(ctx):
    """A large query runs; requests submitted a few ms later are published
    into the same streaming launch (a later cohort), their first tile starts
    on the device before the large query's last tile ends, and every result
    equals its solve alone."""
    import cudapix_amd as C
    big = 10 ** 17
    rnd = random.Random(5)
    xs = [int(10 ** (9 + 5 * rnd.random())) for _ in range(40)]
    want = {x: ctx.pi(x) for x in xs + [big]}
    srv = C.Server(ctx, max_wait_ms=1.0, max_x=big)
    t_big = srv.submit(big)
    time.sleep(0.004)
    later = []
    for x in xs:
        later.append(srv.submit(x))
        time.sleep(0.0002)
    srv.drain()
    res = {r["ticket"]: r for r in srv.poll(100)}
    st = srv.stats()
    srv.close()
    rb = res[t_big]
    assert rb["status"] == 0 and rb["pi"] == want[big]
    for tk, x in zip(later, xs):
        r = res[tk]
        assert r["status"] == 0 and r["pi"] == want[x], r
    lr = [res[t] for t in later]
    print("big", rb, "later cohorts", sorted({r["cohort"] for r in lr}), "stats", st, flush=True)
    # joined mid-launch: the same launch, later cohorts
    assert all(r["launch"] == rb["launch"] for r in lr)
    assert all(r["cohort"] > rb["cohort"] for r in lr)
    assert st["joined"] >= 1
    # overlapped on the device: a later cohort began before the big one's last tile ended
    assert min(r["dev_first_ms"] for r in lr) < rb["dev_last_ms"]


def test_served_stream_slice_c4(ctx):
    """The first 300 requests of the 2000 requests/s stream (x log-uniform to
    1e17), each served value equal to the same x solved alone."""
    import cudapix_amd as C
    xs, ts = c4(2000)
    xs, ts = xs[:300], ts[:300]
    srv = C.Server(ctx, max_x=max(xs))
    tickets, res = served(srv, xs, ts)
    st = srv.stats()
    srv.close()
    print("served slice", st, flush=True)
    assert max(xs) > 10 ** 16
    for tk, x in zip(tickets, xs):
        r = res[tk]
        assert r["status"] == 0, r
        assert r["pi"] == ctx.pi(x), x


def test_batch_mode_server_c4_slice(ctx):
    """CPX_SERVER_BATCHES: the same 300-request slice (x to 1e17) plus tiny x,
    each value equal to the same x solved alone, every request served by one
    of the server's own batch launches; with a one-byte budget every request
    is rejected with CPX_ENOMEM."""
    import cudapix_amd as C
    xs, ts = c4(2000)
    xs, ts = xs[:300] + [5, 16383], ts[:300] + [ts[299] + 1.0, ts[299] + 2.0]
    srv = C.Server(ctx, max_x=max(xs), mode=C.SERVER_BATCHES)
    tickets, res = served(srv, xs, ts)
    st = srv.stats()
    srv.close()
    print("batch-mode slice", st, flush=True)
    for tk, x in zip(tickets, xs):
        r = res[tk]
        assert r["status"] == 0, r
        assert r["pi"] == ctx.pi(x), x
        assert 1 <= r["launch"] <= st["launches"]
    assert st["launches"] == st["batches"] >= 2 and st["joined"] == 0
    assert st["completed"] == len(xs)
    srv = C.Server(ctx, hbm_budget=1, mode=C.SERVER_BATCHES)
    tickets, res = served(srv, [10 ** 9, 10 ** 12], [0.0, 1.0])
    st = srv.stats()
    srv.close()
    assert all(res[t]["status"] == -2 for t in tickets) and st["rejected"] == 2


def test_server_budget_rejects_and_cancels(ctx):
    import cudapix_amd as C
    small, big = 10 ** 9, 10 ** 15
    nb_small, nb_big = ctx.request_bytes(small), ctx.request_bytes(big)
    assert 0 < 64 * nb_small < nb_big
    share = nb_big // 2                                # a slot's share: below the big request
    srv = C.Server(ctx, hbm_budget=2 * share, slots=2, max_batch=4)
    t_big = srv.submit(big)
    ts = [srv.submit(small + i) for i in range(12)]
    c = srv.cancel(ts[-1])
    srv.drain()
    res = {r["ticket"]: r for r in srv.poll(100)}
    st = srv.stats()
    srv.close()
    assert res[t_big]["status"] == -2                  # CPX_ENOMEM: never fits a slot
    if c == 0:
        assert res[ts[-1]]["status"] == -6             # CPX_ECANCELED
    for t in ts[:-1]:
        assert res[t]["status"] == 0 and res[t]["pi"] == ctx.pi(res[t]["x"])
    assert st["rejected"] == 1 and st["arena_bytes"] <= 2 * share


def test_server_uses_the_creators_tuning(oracle):
    """A server made from a context tuned to alpha factor 1.45 plans with that
    tuning (the per-query y), and its values equal the tuned context's."""
    import cudapix_amd as C
    with C.Context(device=0) as t:
        th, kw = t.tuning()["seg_threads"], t.tuning()["seg_kwords"]
        t.set_tuning(th, kw, 1.45, 6)
        xs = [10 ** 11 + 7 * k for k in range(20)] + [10 ** 13, 3 * 10 ** 14]
        want = [t.pi(x) for x in xs]
        ys = [t.pi(x, terms=True)["y"] for x in xs]
        untuned = C.Context(device=0)
        assert any(untuned.pi(x, terms=True)["y"] != y for x, y in zip(xs, ys))
        untuned.close()
        srv = C.Server(t)
        tickets, res = served(srv, xs, [0.05 * k for k in range(len(xs))])
        srv.close()
    assert [res[tk]["pi"] for tk in tickets] == want
    for x, w in zip(xs, want):
        if x <= 10 ** 12:
            assert w == oracle.pi_dr(x, oracle.choose_y(x, 2.0), 6)["pi"]
