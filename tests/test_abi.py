"""The boundary without a GPU: the C ABI library loads and exports every
entry point include/mgcd.h declares, the drop-in C++ library exports the
GmpCudaDevice interface of the reference header, the moduli tables are the
reference's, the messages are the reference's, and the product refuses to
run without a device instead of falling back to the CPU."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "modgcd-onegpu_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mgcd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgcd_\w+)\s*\(", text)))


def test_every_declared_entry_point_is_exported(mgcd):
    lib = mgcd.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(mgcd.EXPORTS) == names


def test_dropin_library_exports_reference_interface():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, "lib", "libgmpcuda.so")],
                         capture_output=True, text=True, check=True).stdout
    for sym in ("_ZN7GmpCuda13GmpCudaDevice9setDeviceEi",             # static void setDevice(int)
                "_ZN7GmpCuda13GmpCudaDeviceC1Ev",                     # GmpCudaDevice()
                "_ZN7GmpCuda13GmpCudaDeviceD1Ev",                     # ~GmpCudaDevice()
                "_ZN7GmpCuda13GmpCudaDevice3gcdEP12__mpz_structS2_S2_",  # gcd(mpz_t, mpz_t, mpz_t)
                "_ZN7GmpCuda6moduliE"):                               # extern const uint32_t moduli[]
        assert sym in out, sym


def test_header_is_host_only_and_compiles_reference_style_code(tmp_path):
    src = tmp_path / "user.cpp"
    src.write_text(
        '#include "GmpCuda.h"\n'
        "int main() {\n"
        "  GmpCuda::GmpCudaDevice::setDevice(0);\n"
        "  GmpCuda::GmpCudaDevice dev;\n"
        "  mpz_t g, u, v; mpz_inits(g, u, v, NULL);\n"
        "  dev.gcd(g, u, v);\n"
        "  int n = dev.getMaxGridSize() + GmpCuda::GmpCudaDevice::GCD_BLOCK_SZ + GmpCuda::L + GmpCuda::W;\n"
        "  GmpCuda::GmpCudaDevice::pair_t p{GmpCuda::moduli[0], 1};\n"
        "  return n + (int)p.modulus == 0;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        "-I/opt/conda/include", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    header = open(os.path.join(ROOT, "include", "GmpCuda.h")).read()
    assert not re.search(r"#\s*include\s*<hip/|\bhip[A-Z]\w*_t\b", header)


def test_messages_are_the_references(mgcd):
    L = mgcd.lib()
    want = {1: "Invalid device number.", 2: "Not enough moduli available for given input.",
            3: "Cannot allocate enough threads to support computation.",
            4: "Ran out of moduli in the reduction loop.", 5: "Ran out of moduli in the recovery loop.",
            6: "Unknown error in the gcd kernel."}
    for code, msg in want.items():
        assert L.mgcd_strerror(code).decode() == msg


def read_table(L, all_primes=False, path=None):
    path = path or os.path.join(PKG, "data", f"moduli_L{L}{'_all' if all_primes else ''}.bin")
    raw = open(path, "rb").read()
    assert raw[:8] == b"MGCDMODS"
    version, width, W, count = np.frombuffer(raw[8:24], np.uint32)
    checksum = int(np.frombuffer(raw[24:32], np.uint64)[0])
    mods = np.frombuffer(raw[48:], np.uint32)
    assert (version, width, W) == (1, L, 0 if all_primes else 64) and count == len(mods)
    return mods, checksum


def is_prime32(n):
    # deterministic Miller-Rabin below 2^32 (bases 2, 7, 61)
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 61):
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in (2, 7, 61):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def fnv1a64(arr):
    h = 0xcbf29ce484222325
    for b in np.asarray(arr, np.uint32).tobytes():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_table_L32_is_the_references():
    mods, _ = read_table(32)
    fx = json.load(open(os.path.join(GOLDEN, "ref_moduli32.json")))
    assert len(mods) >= fx["count"]
    assert mods[:64].tolist() == fx["head"]
    for i, m in fx["samples"]:
        assert int(mods[i]) == m
    assert fnv1a64(mods[: fx["count"]]) == int(fx["fnv1a64"], 16)


@pytest.mark.parametrize("L,count", [(22, 97358), (27, 262144), (32, 262144)])
def test_tables_match_oracle_and_checksums(oracle, L, count):
    mods, checksum = read_table(L)
    assert len(mods) == count
    assert fnv1a64(mods) == checksum
    head = 20000 if L != 22 else count
    assert np.array_equal(mods[:head], oracle.create_moduli(L, head))
    assert np.all(mods[:-1] > mods[1:])


@pytest.mark.parametrize("L,count", [(22, 140336), (27, 262144), (32, 262144)])
def test_all_primes_tables(L, count):
    # Every L-bit prime, descending (CW section 2.4): L = 22 has 140,336.
    mods, checksum = read_table(L, all_primes=True)
    assert len(mods) == count and fnv1a64(mods) == checksum
    assert np.all(mods[:-1] > mods[1:])
    assert all(is_prime32(int(m)) for m in mods[:: 97])
    # complete at the top: no prime above the 2000th entry is missing
    head = set(int(m) for m in mods[:2001])
    for x in range((1 << L) - 1, int(mods[2000]), -2):
        assert (x in head) == is_prime32(x), x
    # the filtered table is a subsequence of it
    cw, _ = read_table(L)
    assert np.isin(cw[:20000], mods).all()
    if L == 22:
        assert int(mods[-1]) > (1 << 21) and np.isin(cw, mods).all()


def test_all_primes_generator_exhaustive_small_width(tmp_path):
    # At L = 14 the generator must list exactly the primes in (2^13, 2^14), descending.
    exe = os.path.join(PKG, "bin", "createModuli")
    out = tmp_path / "m14.bin"
    r = subprocess.run([exe, "14", "100000", str(out), "--all-primes"], capture_output=True, text=True)
    assert r.returncode == 2  # fewer than asked for exist
    mods, _ = read_table(14, all_primes=True, path=str(out))
    want = [x for x in range((1 << 14) - 1, 1 << 13, -1) if is_prime32(x)]
    assert mods.tolist() == want


def test_reference_createModuli_matches_table_when_built():
    exe = os.path.join(ROOT, "oracle", "_ref", "createModuli")
    if not os.path.exists(exe):
        pytest.skip("reference createModuli not built here (oracle/_ref)")
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    ref = [int(x) for x in re.findall(r"^\s+(\d+),", out, re.M)]
    mods, _ = read_table(32)
    assert len(ref) == 65536 and mods[:65536].tolist() == ref


def test_cpu_only_harness(tmp_path):
    # bin/testmodgcd-nogpu times mpz_gcd alone (the reference's
    # testmodgcd-nogpu, Makefile:61-62); its pair cache is the file the GPU
    # harness replays.  When the reference's own CPU-only harness is built
    # from source (oracle/_ref), both must draw and write the same pairs.
    exe = os.path.join(PKG, "bin", "testmodgcd-nogpu")
    a = tmp_path / "ours"
    a.mkdir()
    r = subprocess.run([exe, "3000", "500"], capture_output=True, text=True, timeout=60, cwd=a)
    assert r.returncode == 0, r.stderr
    assert "Executing tests only on CPU." in r.stdout and "GMP time" in r.stdout
    assert "MOD time" not in r.stdout
    ours = (a / "tests" / "3000-500").read_text()
    assert len(ours.splitlines()) == 20
    ref = os.path.join(ROOT, "oracle", "_ref", "testmodgcd-nogpu")
    if not os.path.exists(ref):
        pytest.skip("reference testmodgcd-nogpu not built here (oracle/_ref)")
    b = tmp_path / "ref"
    b.mkdir()
    subprocess.run([ref, "3000", "500"], capture_output=True, text=True, timeout=60, cwd=b, check=True)
    assert (b / "tests" / "3000-500").read_text() == ours


def test_no_gpu_means_an_error_not_a_fallback(mgcd):
    n = ctypes.c_int(0)
    if mgcd.lib().mgcd_device_count(ctypes.byref(n)) == mgcd.OK and n.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(mgcd.MgcdError):
        mgcd.Context(32)


def test_sizing_matches_reference_formula_edges(mgcd):
    # 0- and 1-bit operands: the reference divides by ln(1); the library
    # sizes them as 2-bit operands.  Sizes beyond any table saturate at 2^30.
    assert mgcd.num_moduli_needed(0, 32) == mgcd.num_moduli_needed(1, 32) == mgcd.num_moduli_needed(2, 32) == 256
    assert mgcd.num_moduli_needed(2**64 - 1, 22) == mgcd.num_moduli_needed(2**40, 32) == 2**30
    assert mgcd.num_moduli_needed(1000000, 22) == 92160
