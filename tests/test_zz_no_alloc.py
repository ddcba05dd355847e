"""The no-allocation batch entry point, cpx_pi_batch_noalloc, called through
the C ABI and checked against cpx_pi_batch on the same x (tolerance 0: exact
integers): two calls in a row, the second smaller, so the second runs over
the first one's leftovers in the shared device array; CPX_EAGAIN while the
resident tables do not cover a batch; CPX_ENOMEM for a batch larger than the
array."""
import ctypes
import random

import pytest

pytestmark = pytest.mark.gpu

CPX_ENOMEM, CPX_EAGAIN = -2, -5


def noalloc(ctx, xs):
    import cudapix_amd as C
    fn = ctx._L.cpx_pi_batch_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(C.CpxU128), ctypes.c_size_t,
                   ctypes.POINTER(C.CpxU128)]
    n = len(xs)
    arr = (C.CpxU128 * max(n, 1))(*[C.u128(v) for v in xs])
    out = (C.CpxU128 * max(n, 1))()
    rc = fn(ctx.h, arr, n, out)
    return rc, [C.from_u128(out[i]) for i in range(n)]


def test_noalloc_matches_batch_on_two_calls(ctx):
    rnd = random.Random(23)
    first = [int(10 ** (9 + 7 * rnd.random())) for _ in range(48)] + [10 ** 16, 5, 16383, 16384]
    second = [int(10 ** (8 + 4 * rnd.random())) for _ in range(9)]
    # the allocating path first: its results, and tables that cover both lists
    want1 = ctx.pi_batch(first)
    want2 = ctx.pi_batch(second)
    print("case noalloc first call, %d queries" % len(first), flush=True)
    rc, got1 = noalloc(ctx, first)
    assert rc == 0 and got1 == want1
    print("case noalloc second call, %d queries" % len(second), flush=True)
    rc, got2 = noalloc(ctx, second)
    assert rc == 0 and got2 == want2


def test_noalloc_rejects_what_it_cannot_hold(ctx):
    import cudapix_amd as C
    # a context whose tables cover nothing yet
    fresh = C.Context(device=0)
    try:
        print("case noalloc before tables", flush=True)
        rc, _ = noalloc(fresh, [10 ** 12])
        assert rc == CPX_EAGAIN
        want = fresh.pi_batch([10 ** 12, 10 ** 11])
        rc, got = noalloc(fresh, [10 ** 12, 10 ** 11])
        assert rc == 0 and got == want
    finally:
        fresh.close()
    # ~25 MB of state per x = 1e16: forty of them exceed the 128 MiB array
    ctx.pi_batch([10 ** 16])
    print("case noalloc oversize batch", flush=True)
    rc, _ = noalloc(ctx, [10 ** 16] * 40)
    assert rc == CPX_ENOMEM
    rc, got = noalloc(ctx, [10 ** 16, 10 ** 13])
    assert rc == 0 and got == ctx.pi_batch([10 ** 16, 10 ** 13])
