"""The C++ drop-in (GmpCuda::GmpCudaDevice over the C ABI) and the
command-line harness on the GPU, both checked against mpz_gcd."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "modgcd-onegpu_amd", "bin")


def test_dropin_program():
    r = subprocess.run([os.path.join(BIN, "test_dropin")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


@pytest.mark.parametrize("L", ["32", "22"])
def test_harness_random_pairs(tmp_path, L):
    env = dict(os.environ, MGCD_L=L)
    r = subprocess.run([os.path.join(BIN, "testmodgcd"), "r", "8192", "1024"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path, env=env)
    print(r.stdout)
    assert r.returncode == 0 and "Modular gcd correct" in r.stdout


def test_harness_cache_replay(tmp_path):
    # First run writes tests/<bits>-<gcd_bits> in base 62; the second replays it.
    exe = os.path.join(BIN, "testmodgcd")
    r1 = subprocess.run([exe, "3000", "500"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r1.returncode == 0 and "Modular gcd correct" in r1.stdout
    cache = tmp_path / "tests" / "3000-500"
    assert cache.exists() and len(cache.read_text().splitlines()) == 20
    r2 = subprocess.run([exe, "3000", "500"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r2.returncode == 0 and "Modular gcd correct" in r2.stdout
