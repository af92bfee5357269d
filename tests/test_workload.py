"""The seeded workload generator (modgcd-onegpu_amd/python/workload.py),
without a GPU: SplitMix64 against its published output sequences, the shape
of each configuration's pairs, and the SHA-256 of every workload the
benchmark and the GPU tests use against the committed digests
(tests/golden/workload_digests.json, written by make_workload_digests.py)."""
import json
import math
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_workload_digests as mwd  # noqa: E402


def test_splitmix64_published_sequences(wl):
    # Vigna's splitmix64.c from state 0, and the sequence from state 1234567
    # that implementations are commonly checked against.
    assert [int(x) for x in wl.splitmix64(0, 4)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    assert [int(x) for x in wl.splitmix64(1234567, 5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
        16408922859458223821]
    # A stream continues the sequence across draws.
    st = wl.Stream(0, 0)
    assert [int(x) for x in st.words(2)] + [int(x) for x in st.words(2)] == [int(x) for x in wl.splitmix64(0, 4)]


def test_configuration_shapes(wl):
    u, v = wl.single_small(3)
    g = math.gcd(u, v)
    assert 8100 < u.bit_length() <= 8192 and g.bit_length() >= 1024 and g % 2 == 1
    coprime, equal, vzero = wl.single_small_edges(0)
    assert math.gcd(*coprime) < 2**64 and equal[0] == equal[1] and vzero[1] == 0
    u, v = wl.single_large(0)
    assert 999_000 < u.bit_length() <= 1_000_000 and math.gcd(u, v).bit_length() >= 100_000
    for u, v in wl.headline(4):
        assert 99_900 < max(u.bit_length(), v.bit_length()) <= 100_000 and math.gcd(u, v).bit_length() >= 10_000
    sizes = wl.serve_sizes(2000, seed=5)
    assert set(sizes) == {2000, 50000, 1000000}
    assert abs(sizes.count(2000) / 2000 - 0.70) < 0.04 and abs(sizes.count(1000000) / 2000 - 0.05) < 0.02


@pytest.mark.parametrize("name", list(mwd.workloads()))
def test_workload_digests(name):
    want = json.load(open(os.path.join(ROOT, "tests", "golden", "workload_digests.json")))[name]
    desc, make = mwd.workloads()[name]
    pairs = make()
    assert len(pairs) == want["pairs"]
    assert mwd.digest(pairs) == want["sha256"], f"{name}: the generator's {desc} changed"
