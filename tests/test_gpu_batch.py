"""Batch planning on the GPU: size classes with their own launches, the
hipGraph-captured launch sequence, and the autotuner.  Every result is
checked against math.gcd (equal to mpz_gcd)."""
import math
import os

import pytest

pytestmark = pytest.mark.gpu


def _mix(wl, n, seed):
    # Sizes across all four N_Q classes (1k .. 200k bits), coprime and v = 0
    # pairs included.
    return wl.mixed(n=n, seed=seed, lo=1000, hi=200000)


def _check(b, pairs):
    for i, (u, v) in enumerate(pairs):
        assert b.result(i) == math.gcd(u, v), i


def test_size_classes_get_their_own_tilings(ctx_for, mgcd, wl):
    ctx = ctx_for(32)
    # Two classes of comparable elimination work (about N_Q^2 per pair), each
    # with its own built-in tiling and launch: 4,096 small pairs (N_Q <= 1024)
    # fill the device with (512, 2, 16) groups; 48 mid pairs (N_Q 3,328) do
    # not, and take the smaller parts of (256, 1, 16).
    small = wl.headline(4096, seed=31, nbits=5000, gbits=500)
    mid = wl.headline(48, seed=35, nbits=30000, gbits=3000)
    assert mgcd.num_moduli_needed(5000, 32) <= 1024 < mgcd.num_moduli_needed(30000, 32) <= 4096
    pairs = small + mid
    b = mgcd.Batch.from_pairs(ctx, pairs)
    b.run()
    launches = b.launches()
    assert len(launches) == 2
    assert [x["tiling"] for x in launches] == [(512, 2, 16), (256, 1, 16)]
    assert launches[0]["nq"][1] <= 1024 < launches[1]["nq"][0]
    assert [x["pairs"] for x in launches] == [4096, 48]
    for x in launches:
        assert x["device_s"] > 0
    _check(b, pairs)
    b.close()


def test_built_in_tiling_follows_the_population(ctx_for, mgcd, wl):
    # Small populations spread over the smaller parts of (256, 1, 16); one
    # that fills the co-resident grid of (512, 2, 16) takes it.
    ctx = ctx_for(32)
    few = wl.headline(256, seed=38, nbits=8192, gbits=1024)
    b = mgcd.Batch.from_pairs(ctx, few)
    assert b.tiling() == (256, 1, 16)
    b.run()
    _check(b, few)
    b.close()
    many = few + wl.headline(3840, seed=39, nbits=8192, gbits=1024)
    b = mgcd.Batch.from_pairs(ctx, many)
    assert b.tiling() == (512, 2, 16)
    b.run()
    _check(b, many)
    b.close()


def test_small_class_shares_the_dominant_launch(ctx_for, mgcd, wl):
    ctx = ctx_for(32)
    # Under a tenth of the largest class's work: the small pairs take the
    # large ones' tiling and run in their launch.
    pairs = wl.headline(8, seed=36, nbits=5000, gbits=500) + wl.headline(8, seed=37, nbits=30000, gbits=3000)
    b = mgcd.Batch.from_pairs(ctx, pairs)
    b.run()
    launches = b.launches()
    assert len(launches) == 1 and launches[0]["pairs"] == 16
    assert launches[0]["nq"][0] <= 1024 < launches[0]["nq"][1]
    _check(b, pairs)
    b.close()
    # Any mix across all four classes stays exact.
    pairs = _mix(wl, 96, 31)
    b = mgcd.Batch.from_pairs(ctx, pairs)
    b.run()
    assert sum(x["pairs"] for x in b.launches()) == sum(1 for u, v in pairs if u and v)
    _check(b, pairs)
    b.close()


def test_graph_replays_match_eager(ctx_for, mgcd, wl):
    ctx = ctx_for(32)
    pairs = _mix(wl, 64, 32)
    want = [math.gcd(u, v) for u, v in pairs]
    eager = mgcd.Batch.from_pairs(ctx, pairs)
    eager.run()
    assert eager.results() == want
    ctx.set_option(mgcd.OPT_BATCH_GRAPH, 1)
    try:
        g = mgcd.Batch.from_pairs(ctx, pairs)
    finally:
        ctx.set_option(mgcd.OPT_BATCH_GRAPH, 0)
    for rep in range(3):
        g.run()
        assert g.results() == want, f"graph replay {rep}"
        eager.run()  # eager work between the replays
        assert eager.results() == want
    g.close()
    eager.close()


def test_autotune_returns_a_working_tiling(ctx_for, mgcd, wl):
    ctx = ctx_for(32)
    pairs = wl.headline(48, seed=33, nbits=6000, gbits=600)
    T, K, B = mgcd.autotune(ctx, pairs)
    assert (T, K, B) in {(256, 2, 16), (512, 2, 16), (256, 1, 16), (512, 1, 16), (256, 2, 8), (256, 1, 8),
                         (256, 1, 32), (128, 1, 16)}
    b = mgcd.Batch.from_pairs(ctx, pairs, workgroup=T, slots=K, group=B)
    b.run()
    assert b.results() == [math.gcd(u, v) for u, v in pairs]
    b.close()


def test_autotuned_batch_and_cache_file(mgcd, wl, tmp_path):
    cache = tmp_path / "tune.txt"
    os.environ["MGCD_TUNE_CACHE"] = str(cache)
    try:
        ctx = mgcd.Context(32)
        ctx.set_option(mgcd.OPT_AUTOTUNE, 1)
        pairs = _mix(wl, 40, 34)
        b = mgcd.Batch.from_pairs(ctx, pairs)
        b.run()
        assert b.results() == [math.gcd(u, v) for u, v in pairs]
        lines = cache.read_text().split()
        assert len(lines) >= 4 and len(lines) % 4 == 0  # "L:class:population T K B" per tuned class
        b.close()
        ctx.close()
        # A new context reads the tuned tilings back instead of measuring.
        ctx2 = mgcd.Context(32)
        ctx2.set_option(mgcd.OPT_AUTOTUNE, 1)
        b2 = mgcd.Batch.from_pairs(ctx2, pairs)
        b2.run()
        assert b2.results() == [math.gcd(u, v) for u, v in pairs]
        b2.close()
        ctx2.close()
    finally:
        del os.environ["MGCD_TUNE_CACHE"]
