"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py to compute expected answers and to time GMP.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

OK, REDUX_ERROR, RECOVERY_ERROR, CAPACITY, ARG = 0, 1, 2, 3, 4


class Pair(ctypes.Structure):
    _fields_ = [("modulus", ctypes.c_uint32), ("value", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_num_moduli_needed.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.oracle_num_moduli_needed.restype = ctypes.c_int
        L.oracle_cw_usable.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.oracle_cw_usable.restype = ctypes.c_int
        L.oracle_create_moduli.argtypes = [ctypes.c_int, u32p, ctypes.c_size_t]
        L.oracle_create_moduli.restype = ctypes.c_size_t
        L.oracle_cw_inverse.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.oracle_cw_inverse.restype = ctypes.c_uint64
        L.oracle_mod.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int]
        L.oracle_mod.restype = ctypes.c_uint32
        L.oracle_mod_mul.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.oracle_mod_mul.restype = ctypes.c_uint32
        L.oracle_mod_inv.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_mod_inv.restype = ctypes.c_uint32
        L.oracle_mod_mp.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int]
        L.oracle_mod_mp.restype = ctypes.c_uint32
        pp = ctypes.POINTER(Pair)
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.oracle_gcd_kernel.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, u32p, ctypes.c_int,
                                        ctypes.c_int, pp, ctypes.c_size_t, szp, pp, ctypes.c_size_t, szp]
        L.oracle_gcd_kernel.restype = ctypes.c_int
        L.oracle_gcd_kernel_exact.argtypes = L.oracle_gcd_kernel.argtypes
        L.oracle_gcd_kernel_exact.restype = ctypes.c_int
        L.oracle_horner.argtypes = [pp, ctypes.c_size_t, u32p, ctypes.c_size_t]
        L.oracle_horner.restype = ctypes.c_size_t
        L.oracle_mpz_gcd.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, u32p, ctypes.c_size_t]
        L.oracle_mpz_gcd.restype = ctypes.c_size_t
        L.oracle_gmp_version.restype = ctypes.c_char_p
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_time_mpz_gcd.argtypes = [u32p, u64p, u32p, u64p, u32p, ctypes.c_size_t, ctypes.c_int, u64p]
        L.oracle_time_mpz_gcd.restype = ctypes.c_double
        _lib = L
    return _lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def int_to_limbs(x):
    """Non-negative Python int -> uint32 limbs, least significant first (no trailing zero limbs)."""
    x = abs(int(x))
    n = (x.bit_length() + 31) // 32
    return np.frombuffer(x.to_bytes(4 * n, "little"), dtype=np.uint32).copy() if n else np.zeros(0, np.uint32)


def limbs_to_int(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return int.from_bytes(a.tobytes(), "little")


def num_moduli_needed(bits, L):
    return lib().oracle_num_moduli_needed(bits, L)


def create_moduli(L, count):
    out = np.zeros(count, np.uint32)
    n = lib().oracle_create_moduli(L, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), count)
    return out[:n]


def mod_inv(v, m):
    return lib().oracle_mod_inv(v, m)


def mod_mp(x, m, L):
    a, p = _u32(int_to_limbs(x))
    return lib().oracle_mod_mp(p, len(a), m, L)


def gcd_kernel(u, v, moduli, L, dcap=None, tcap=None, exact=False):
    """Runs the kernel restatement.  u, v: Python ints (the limb arrays passed
    are |u|, |v| exactly as the reference's mpz_export would stage them).
    exact: reduce with x % m instead of the reference's division by multiply
    (required for moduli outside the Cavagnino-Werbrouck condition).
    Returns (status, digits[(m, g)], pivots[(p, b)])."""
    ua, up = _u32(int_to_limbs(u))
    va, vp = _u32(int_to_limbs(v))
    ma, mp = _u32(moduli)
    nq = len(ma)
    dcap = dcap or nq + 1
    tcap = tcap or nq + 1
    digits = (Pair * dcap)()
    trace = (Pair * tcap)()
    nd = ctypes.c_size_t()
    nt = ctypes.c_size_t()
    fn = lib().oracle_gcd_kernel_exact if exact else lib().oracle_gcd_kernel
    st = fn(up, len(ua), vp, len(va), mp, nq, L, digits, dcap, ctypes.byref(nd), trace, tcap, ctypes.byref(nt))
    d = [(digits[i].modulus, digits[i].value) for i in range(min(nd.value, dcap))]
    t = [(trace[i].modulus, trace[i].value) for i in range(min(nt.value, tcap))]
    return st, d, t


def horner(digits):
    """Host Horner of the signed mixed-radix digits (least significant first) -> |G|."""
    k = len(digits)
    arr = (Pair * max(k, 1))()
    for i, (m, g) in enumerate(digits):
        arr[i].modulus = m
        arr[i].value = g
    bits = sum(int(m).bit_length() for m, _ in digits) + 64
    cap = bits // 32 + 2
    out = np.zeros(cap, np.uint32)
    n = lib().oracle_horner(arr, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap)
    return limbs_to_int(out[:n])


def mpz_gcd(u, v):
    ua, up = _u32(int_to_limbs(u))
    va, vp = _u32(int_to_limbs(v))
    cap = max(len(ua), len(va)) + 1
    out = np.zeros(cap, np.uint32)
    n = lib().oracle_mpz_gcd(up, len(ua), vp, len(va), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap)
    return limbs_to_int(out[:n])


def gmp_version():
    return lib().oracle_gmp_version().decode()


def time_mpz_gcd(limbs, u_off, nu, v_off, nv, threads):
    """Times mpz_gcd over packed pairs with `threads` workers.  Returns (seconds, gcd bit lengths)."""
    limbs = np.ascontiguousarray(limbs, np.uint32)
    u_off = np.ascontiguousarray(u_off, np.uint64)
    v_off = np.ascontiguousarray(v_off, np.uint64)
    nu = np.ascontiguousarray(nu, np.uint32)
    nv = np.ascontiguousarray(nv, np.uint32)
    n = len(nu)
    gbits = np.zeros(n, np.uint64)
    P32 = ctypes.POINTER(ctypes.c_uint32)
    P64 = ctypes.POINTER(ctypes.c_uint64)
    sec = lib().oracle_time_mpz_gcd(limbs.ctypes.data_as(P32), u_off.ctypes.data_as(P64), nu.ctypes.data_as(P32),
                                    v_off.ctypes.data_as(P64), nv.ctypes.data_as(P32), n, threads,
                                    gbits.ctypes.data_as(P64))
    return sec, gbits
