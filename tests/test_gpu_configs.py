"""The batched and serving configurations at their real shapes, every result
against GMP mpz_gcd (the reference's discipline, testmodgcd.cpp:283-295):

* 4096 mixed-size pairs (1k-200k bits, 10% coprime, 1% v = 0, L = 32) in one
  batch with the built-in size-class tilings;
* a skewed batch at L = 27 whose 500k-bit pairs (N_Q 45,568) form a group of
  45 parts at tiling (512, 2, 16): 720 exchange slots per round, polled in
  three passes without the relay, the launch captured as a hipGraph and
  replayed, each replay against the eager run;
* the serving loop with 1M-bit requests among 2k- and 50k-bit ones (three
  slots, the last for requests above 16,384 moduli): each result equal to
  the request's result alone, and a wait that expires reports PENDING and
  keeps the ticket.
"""
import random
import threading
import time

import pytest

pytestmark = pytest.mark.gpu


def test_mixed_4096_batch(ctx_for, mgcd, oracle, wl):
    ctx = ctx_for(32)
    pairs = wl.mixed(4096, seed=3)
    b = mgcd.Batch.from_pairs(ctx, pairs)
    b.run()
    res = b.results()
    for i, (u, v) in enumerate(pairs):
        assert res[i] == oracle.mpz_gcd(u, v), i
    b.close()


def test_skewed_batch_three_pass_exchange_graph(ctx_for, mgcd, oracle, wl):
    ctx = ctx_for(27)
    rnd = random.Random(6)
    large = [wl.pair(wl.Stream(61, i), 500000, 100000) for i in range(4)]
    small = [wl.pair(wl.Stream(62, i), 4096, 819) for i in range(60)]
    pairs = small[:20] + large[:2] + small[20:] + large[2:]
    rnd.shuffle(pairs)
    nq = mgcd.num_moduli_needed(500000, 27)
    G = (nq + 1023) // 1024
    assert nq == 45568 and G == 45 and G * 16 > 256 and G < 48  # several passes, no relay
    want = [oracle.mpz_gcd(u, v) for u, v in pairs]
    eager = mgcd.Batch.from_pairs(ctx, pairs, workgroup=512, slots=2, group=16)
    eager.run()
    assert eager.results() == want
    big = [i for i, (u, v) in enumerate(pairs) if u.bit_length() > 400000]
    assert len(big) == 4 and all(eager.pair_stats(i)[0] == nq for i in big)
    ctx.set_option(mgcd.OPT_BATCH_GRAPH, 1)
    try:
        g = mgcd.Batch.from_pairs(ctx, pairs, workgroup=512, slots=2, group=16)
    finally:
        ctx.set_option(mgcd.OPT_BATCH_GRAPH, 0)
    for rep in range(3):
        g.run()
        assert g.results() == want, f"graph replay {rep}"
        eager.run()  # eager work between the replays
        assert eager.results() == want
    g.close()
    eager.close()


def test_serving_with_1M_requests(mgcd, oracle, wl):
    ctx = mgcd.Context(32)  # the server owns its context
    rnd = random.Random(51)
    sizes = [2000] * 24 + [50000] * 8 + [1000000] * 3
    rnd.shuffle(sizes)
    reqs = [wl.serve_pair(151, i, nb) for i, nb in enumerate(sizes)]
    want = [oracle.mpz_gcd(u, v) for u, v in reqs]
    big = [i for i, nb in enumerate(sizes) if nb == 1000000]
    srv = mgcd.Server(ctx, budget_bytes=64 << 30, slots=3)
    try:
        alone = {}
        for i in (big[0], sizes.index(2000)):  # each alone: nothing else submitted
            t = srv.submit(*reqs[i])
            if i == big[0]:
                with pytest.raises(mgcd.MgcdError) as e:  # far shorter than a 1M-bit gcd
                    srv.wait(t, timeout=0.0)
                assert e.value.status == mgcd.ERR_PENDING
            alone[i] = srv.wait(t)[0]
        tickets = [None] * len(reqs)

        def feed():
            for i in range(len(reqs)):
                tickets[i] = srv.submit(*reqs[i])
                time.sleep(rnd.random() * 0.003)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        got = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
        ctx.close()
    assert got == want
    for i, g in alone.items():
        assert got[i] == g == want[i]
    assert st["completed"] == len(reqs) + len(alone)
