"""The serving loop on the GPU: requests of mixed sizes arrive while others
are in flight, are admitted against a device-memory budget, and every
result is the same as the request's result alone (bit-exact, = mpz_gcd)."""
import math
import random
import threading
import time

import pytest

pytestmark = pytest.mark.gpu


def _requests(wl, n, seed):
    rnd = random.Random(seed)
    sizes = [rnd.choice([2000, 2000, 2000, 8000, 50000]) for _ in range(n)]
    return [wl.serve_pair(seed, i, nb) for i, nb in enumerate(sizes)]


def test_results_alone_and_in_the_mix(ctx_for, mgcd, wl):
    ctx = mgcd.Context(32)  # the server owns its context
    reqs = _requests(wl, 120, 41)
    alone = {}
    srv = mgcd.Server(ctx, budget_bytes=1 << 30, slots=2)
    for i in (0, 7, 33):  # each alone: nothing else submitted
        alone[i] = srv.wait(srv.submit(*reqs[i]))[0]
    # The mix: submitted from two threads with small random gaps, so later
    # requests arrive while earlier launches are on the device.
    tickets = [None] * len(reqs)

    def feed(idx, seed):
        rnd = random.Random(seed)
        for i in idx:
            tickets[i] = srv.submit(*reqs[i])
            time.sleep(rnd.random() * 0.002)

    ths = [threading.Thread(target=feed, args=(range(k, len(reqs), 2), k)) for k in range(2)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    got = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    srv.close()
    ctx.close()
    for i, (u, v) in enumerate(reqs):
        assert got[i] == math.gcd(u, v), i
    for i, g in alone.items():
        assert got[i] == g
    assert st["completed"] == len(reqs) + len(alone)
    assert 0 < st["p50_s"] <= st["p99_s"]


def test_budget_admission(mgcd, wl):
    ctx = mgcd.Context(32)
    # A budget that holds about two 50k-bit requests: the rest wait for
    # memory and are admitted as launches finish.
    u, v = wl.serve_pair(42, 0, 50000)
    nq = mgcd.num_moduli_needed(max(u.bit_length(), v.bit_length()), 32)
    one = 24 * (nq + 64) + 4 * ((u.bit_length() + 31) // 32 + (v.bit_length() + 31) // 32) + 1024
    srv = mgcd.Server(ctx, budget_bytes=2 * one + one // 2, slots=2)
    reqs = [wl.serve_pair(42, i, 50000) for i in range(8)]
    tickets = [srv.submit(*r) for r in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    assert got == [math.gcd(*r) for r in reqs]
    assert st["peak_bytes"] <= 2 * one + one // 2
    # A request larger than the whole budget is refused.
    big = wl.serve_pair(42, 99, 200000)
    with pytest.raises(mgcd.MgcdError):
        srv.submit(*big)
    srv.close()
    ctx.close()


def test_large_requests_beyond_a_slots_share(mgcd, wl):
    # Eight slots leave each a sixteenth of the co-resident grid or less:
    # eight 500k-bit requests (N_Q 42,752) submitted together split over more
    # workgroups than the large-request slot's share when grouped.  Such a
    # launch cannot be planned; its requests go through the one-pair path,
    # which may use the whole device, and every result is still exact.
    ctx = mgcd.Context(32)
    srv = mgcd.Server(ctx, budget_bytes=1 << 30, slots=8)
    reqs = [wl.serve_pair(43, i, 500000) for i in range(8)]
    tickets = [srv.submit(*r) for r in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    srv.close()
    ctx.close()
    assert got == [math.gcd(*r) for r in reqs]
