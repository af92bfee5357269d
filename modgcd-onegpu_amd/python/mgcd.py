"""ctypes binding of libmgcd.so (the C ABI in include/mgcd.h).

This is plumbing for Python callers (tests, bench): every call goes through
the HIP library.  There is no fallback: if libmgcd.so is not built, or no
GPU is visible, the calls raise.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MGCD_LIB selects another build of the library (the instrumented libmgcd_phases.so).
LIB_PATH = os.environ.get("MGCD_LIB") or os.path.join(PKG_DIR, "lib", "libmgcd.so")
GMPCUDA_PATH = os.path.join(PKG_DIR, "lib", "libgmpcuda.so")
DATA_DIR = os.path.join(PKG_DIR, "data")

OK = 0
ERR_INVALID_DEVICE = 1
ERR_NOT_ENOUGH_MODULI = 2
ERR_TOO_MANY_THREADS = 3
ERR_REDUX = 4
ERR_RECOVERY = 5
ERR_KERNEL = 6
ERR_HIP = 7
ERR_ARG = 8
ERR_TABLE = 9
ERR_CAPACITY = 10
ERR_NOMEM = 11
ERR_TIMEOUT = 12
ERR_BUSY = 13
ERR_SCRATCH = 14
ERR_PENDING = 15

# Every entry point declared in include/mgcd.h.
EXPORTS = [
    "mgcd_strerror", "mgcd_device_count", "mgcd_set_device", "mgcd_num_moduli_needed",
    "mgcd_ctx_create", "mgcd_ctx_destroy", "mgcd_ctx_L", "mgcd_ctx_num_moduli", "mgcd_ctx_max_grid",
    "mgcd_ctx_moduli", "mgcd_gcd_limbs", "mgcd_gcd_limbs_trace", "mgcd_batch_create", "mgcd_batch_run",
    "mgcd_batch_status", "mgcd_batch_result", "mgcd_batch_timing", "mgcd_batch_tiling", "mgcd_batch_destroy",
    "mgcd_certify_quasi_quotient", "mgcd_last_error_detail", "mgcd_ctx_last_kernel_seconds",
    "mgcd_ctx_synchronize", "mgcd_batch_pair_stats", "mgcd_ctx_set_option", "mgcd_batch_launch_count",
    "mgcd_batch_launch_info", "mgcd_autotune", "mgcd_server_create", "mgcd_server_submit", "mgcd_server_wait",
    "mgcd_server_stats", "mgcd_server_destroy", "mgcd_mixed_radix_to_limbs", "mgcd_ctx_last_pair_stats",
    "mgcd_gcd_limbs_noalloc", "mgcd_batch_pivots", "mgcd_ctx_phase_clocks",
]

OPT_FORCE_MODULI = 1
OPT_RETRY = 2
OPT_SINGLE_WORKGROUP = 3
OPT_SINGLE_SLOTS = 4
OPT_AUTOTUNE = 5
OPT_BATCH_GRAPH = 6
OPT_RELAY_PARTS = 7
OPT_DEVICE_CONVERT = 8
OPT_SPIN_LIMIT = 9
OPT_COOPERATIVE = 10


class Pair(ctypes.Structure):
    _fields_ = [("modulus", ctypes.c_uint32), ("value", ctypes.c_int32)]


class MgcdError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = _lib.mgcd_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: {msg} (status {status})" if where else msg)


_lib = None
P32 = ctypes.POINTER(ctypes.c_uint32)
P64 = ctypes.POINTER(ctypes.c_uint64)
PSZ = ctypes.POINTER(ctypes.c_size_t)
PPair = ctypes.POINTER(Pair)


def lib():
    """Loads libmgcd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C modgcd-onegpu_amd`")
        L = ctypes.CDLL(LIB_PATH)
        L.mgcd_strerror.restype = ctypes.c_char_p
        L.mgcd_strerror.argtypes = [ctypes.c_int]
        L.mgcd_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.mgcd_set_device.argtypes = [ctypes.c_int]
        L.mgcd_num_moduli_needed.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.mgcd_ctx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
        L.mgcd_ctx_destroy.argtypes = [ctypes.c_void_p]
        for f in ("mgcd_ctx_L", "mgcd_ctx_num_moduli", "mgcd_ctx_max_grid"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.mgcd_ctx_moduli.argtypes = [ctypes.c_void_p, P32, ctypes.c_size_t]
        L.mgcd_gcd_limbs.argtypes = [ctypes.c_void_p, P32, ctypes.c_size_t, P32, ctypes.c_size_t, P32,
                                     ctypes.c_size_t, PSZ]
        L.mgcd_gcd_limbs_noalloc.argtypes = L.mgcd_gcd_limbs.argtypes
        L.mgcd_gcd_limbs_trace.argtypes = [ctypes.c_void_p, P32, ctypes.c_size_t, P32, ctypes.c_size_t, P32,
                                           ctypes.c_size_t, PSZ, PPair, ctypes.c_size_t, PSZ, PPair,
                                           ctypes.c_size_t, PSZ]
        L.mgcd_batch_create.argtypes = [ctypes.c_void_p, P32, P64, P32, P64, P32, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.mgcd_batch_run.argtypes = [ctypes.c_void_p]
        L.mgcd_batch_status.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.mgcd_batch_result.argtypes = [ctypes.c_void_p, ctypes.c_size_t, P32, ctypes.c_size_t, PSZ]
        L.mgcd_batch_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
        PI = ctypes.POINTER(ctypes.c_int)
        L.mgcd_batch_tiling.argtypes = [ctypes.c_void_p, PI, PI, PI]
        L.mgcd_batch_destroy.argtypes = [ctypes.c_void_p]
        L.mgcd_certify_quasi_quotient.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P64, P64]
        L.mgcd_last_error_detail.restype = ctypes.c_char_p
        L.mgcd_ctx_last_kernel_seconds.argtypes = [ctypes.c_void_p]
        L.mgcd_ctx_last_kernel_seconds.restype = ctypes.c_double
        L.mgcd_ctx_synchronize.argtypes = [ctypes.c_void_p]
        L.mgcd_ctx_last_pair_stats.argtypes = [ctypes.c_void_p, P32, P32, P32, ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.mgcd_ctx_set_option.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
        L.mgcd_ctx_phase_clocks.argtypes = [ctypes.c_void_p, P64, ctypes.c_size_t, ctypes.c_int]
        L.mgcd_batch_pivots.argtypes = [ctypes.c_void_p, ctypes.c_size_t, PPair, ctypes.c_size_t, PSZ]
        L.mgcd_batch_pair_stats.argtypes = [ctypes.c_void_p, ctypes.c_size_t, P32, P32, P32]
        L.mgcd_batch_launch_count.argtypes = [ctypes.c_void_p, PSZ]
        L.mgcd_batch_launch_info.argtypes = [ctypes.c_void_p, ctypes.c_size_t, PI, PI, PI, PSZ, P32, P32,
                                             ctypes.POINTER(ctypes.c_double)]
        L.mgcd_autotune.argtypes = [ctypes.c_void_p, P32, P64, P32, P64, P32, ctypes.c_size_t, PI, PI, PI]
        L.mgcd_server_create.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.mgcd_server_submit.argtypes = [ctypes.c_void_p, P32, ctypes.c_size_t, P32, ctypes.c_size_t, P64]
        L.mgcd_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, P32, ctypes.c_size_t, PSZ,
                                       ctypes.POINTER(ctypes.c_double)]
        L.mgcd_server_stats.argtypes = [ctypes.c_void_p, P64, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), P64]
        L.mgcd_server_destroy.argtypes = [ctypes.c_void_p]
        L.mgcd_mixed_radix_to_limbs.argtypes = [ctypes.c_void_p, PPair, ctypes.c_size_t, P32, ctypes.c_size_t, PSZ]
        _lib = L
    return _lib


def check(status, where=""):
    if status != OK:
        raise MgcdError(status, where + " [" + lib().mgcd_last_error_detail().decode() + "]")


def int_to_limbs(x):
    x = abs(int(x))
    n = (x.bit_length() + 31) // 32
    return np.frombuffer(x.to_bytes(4 * n, "little"), dtype=np.uint32).copy() if n else np.zeros(0, np.uint32)


def limbs_to_int(a):
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint32).tobytes(), "little")


def _p32(a):
    return a.ctypes.data_as(P32)


def num_moduli_needed(bits, L):
    return lib().mgcd_num_moduli_needed(bits, L)


def device_count():
    n = ctypes.c_int()
    check(lib().mgcd_device_count(ctypes.byref(n)), "mgcd_device_count")
    return n.value


class Context:
    """A device with one moduli table (L in 22, 27, 32) resident in HBM."""

    def __init__(self, L=32, device=0, table_path=None, all_primes=False):
        """all_primes: the table of every L-bit prime (moduli_L<L>_all.bin)
        instead of the Cavagnino-Werbrouck-filtered one the reference uses."""
        self.L = L
        h = ctypes.c_void_p()
        path = table_path.encode() if table_path else None
        if path is None and (all_primes or "MGCD_TABLE_DIR" not in os.environ):
            d = os.environ.get("MGCD_TABLE_DIR") or DATA_DIR
            path = os.path.join(d, f"moduli_L{L}{'_all' if all_primes else ''}.bin").encode()
        check(lib().mgcd_ctx_create(ctypes.byref(h), device, L, path), "mgcd_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().mgcd_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_moduli(self):
        return lib().mgcd_ctx_num_moduli(self.h)

    @property
    def max_grid(self):
        return lib().mgcd_ctx_max_grid(self.h)

    def moduli(self):
        n = self.num_moduli
        out = np.zeros(n, np.uint32)
        check(lib().mgcd_ctx_moduli(self.h, _p32(out), n), "mgcd_ctx_moduli")
        return out

    def gcd(self, u, v):
        """gcd(|u|, |v|) of Python ints through the one-pair path."""
        ua, va = int_to_limbs(u), int_to_limbs(v)
        cap = max(len(ua), len(va)) + 1
        g = np.zeros(cap, np.uint32)
        ng = ctypes.c_size_t()
        check(lib().mgcd_gcd_limbs(self.h, _p32(ua), len(ua), _p32(va), len(va), _p32(g), cap, ctypes.byref(ng)),
              "mgcd_gcd_limbs")
        return limbs_to_int(g[:ng.value])

    def gcd_noalloc(self, u, v):
        """gcd(|u|, |v|) through mgcd_gcd_limbs_noalloc (no device allocation)."""
        ua, va = int_to_limbs(u), int_to_limbs(v)
        cap = max(len(ua), len(va)) + 1
        g = np.zeros(cap, np.uint32)
        ng = ctypes.c_size_t()
        check(lib().mgcd_gcd_limbs_noalloc(self.h, _p32(ua), len(ua), _p32(va), len(va), _p32(g), cap,
                                           ctypes.byref(ng)), "mgcd_gcd_limbs_noalloc")
        return limbs_to_int(g[:ng.value])

    def gcd_trace(self, u, v, dcap=None, tcap=None):
        """Runs the kernel on (u, v) in the given order.  Returns
        (status, g, digits, pivots) without raising on kernel errors."""
        ua, va = int_to_limbs(u), int_to_limbs(v)
        cap = max(len(ua), len(va)) + 1
        g = np.zeros(cap, np.uint32)
        nq = max(num_moduli_needed(max(abs(int(u)).bit_length(), 2), self.L), 256)
        dcap = dcap or nq + 1
        tcap = tcap or nq + 1
        digits = (Pair * dcap)()
        pivots = (Pair * tcap)()
        ng, nd, nt = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        st = lib().mgcd_gcd_limbs_trace(self.h, _p32(ua), len(ua), _p32(va), len(va), _p32(g), cap,
                                        ctypes.byref(ng), digits, dcap, ctypes.byref(nd), pivots, tcap,
                                        ctypes.byref(nt))
        d = [(digits[i].modulus, digits[i].value) for i in range(min(nd.value, dcap))]
        t = [(pivots[i].modulus, pivots[i].value) for i in range(min(nt.value, tcap))]
        return st, (limbs_to_int(g[:ng.value]) if st == OK else None), d, t

    def mixed_radix(self, digits):
        """|G| of signed mixed-radix digits [(m_0, g_0), (m_1, g_1), ...] (least
        significant first), converted on the device by the product tree."""
        k = len(digits)
        arr = (Pair * max(k, 1))()
        for i, (m, g) in enumerate(digits):
            arr[i].modulus = m
            arr[i].value = g
        cap = k + 2
        out = np.zeros(cap, np.uint32)
        ng = ctypes.c_size_t()
        check(lib().mgcd_mixed_radix_to_limbs(self.h, arr, k, _p32(out), cap, ctypes.byref(ng)),
              "mgcd_mixed_radix_to_limbs")
        return limbs_to_int(out[:ng.value])

    def set_option(self, option, value):
        check(lib().mgcd_ctx_set_option(self.h, option, value), "mgcd_ctx_set_option")

    def synchronize(self):
        check(lib().mgcd_ctx_synchronize(self.h), "mgcd_ctx_synchronize")

    def last_kernel_seconds(self):
        return lib().mgcd_ctx_last_kernel_seconds(self.h)

    PHASES = ["residues", "elim_compute", "elim_barrier", "elim_exchange", "elim_bookkeeping", "elim_barrier2",
              "elim_tail", "replay", "recovery"]

    def phase_clocks(self, reset=True):
        """Phase sums of the instrumented build (MGCD_LIB=...libmgcd_phases.so): dict of shader
        cycles per phase plus rounds, items, realtime ticks (100 MHz) and memtime cycles, the
        one-pass collect counts, and each wave's (0-7) compute and barrier cycles."""
        out = np.zeros(32, np.uint64)
        check(lib().mgcd_ctx_phase_clocks(self.h, out.ctypes.data_as(P64), 32, 1 if reset else 0),
              "mgcd_ctx_phase_clocks")
        d = {name: int(out[i]) for i, name in enumerate(self.PHASES)}
        d.update(rounds=int(out[9]), items=int(out[10]), realtime_ticks=int(out[11]), memtime_cycles=int(out[12]),
                 first_load_cycles=int(out[13]), collects_ready=int(out[14]), collects=int(out[15]),
                 wave_compute=[int(x) for x in out[16:24]], wave_barrier=[int(x) for x in out[24:32]])
        return d

    def last_pair_stats(self):
        """dict(nq, iters, digits, tiling=(T, K), parts) of the last one-pair launch."""
        q, r, k = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        t, s, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mgcd_ctx_last_pair_stats(self.h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(k),
                                             ctypes.byref(t), ctypes.byref(s), ctypes.byref(g)))
        return {"nq": q.value, "iters": r.value, "digits": k.value, "tiling": (t.value, s.value), "parts": g.value}

    def certify(self, mode, L=0):
        fail, checked = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().mgcd_certify_quasi_quotient(self.h, mode, L, ctypes.byref(fail), ctypes.byref(checked)),
              "mgcd_certify_quasi_quotient")
        return fail.value, checked.value


def pack_pairs(pairs):
    """[(u, v), ...] of ints -> (limbs, u_off, nu, v_off, nv) arrays."""
    chunks, u_off, nu, v_off, nv = [], [], [], [], []
    pos = 0
    for u, v in pairs:
        ua, va = int_to_limbs(u), int_to_limbs(v)
        u_off.append(pos); nu.append(len(ua)); chunks.append(ua); pos += len(ua)
        v_off.append(pos); nv.append(len(va)); chunks.append(va); pos += len(va)
    limbs = np.concatenate(chunks) if chunks else np.zeros(1, np.uint32)
    if len(limbs) == 0:
        limbs = np.zeros(1, np.uint32)
    return (limbs.astype(np.uint32), np.array(u_off, np.uint64), np.array(nu, np.uint32),
            np.array(v_off, np.uint64), np.array(nv, np.uint32))


def autotune(ctx, pairs):
    """The fastest batch tiling (threads, slots, pairs per group) measured on these pairs."""
    limbs, u_off, nu, v_off, nv = pack_pairs(pairs)
    t, k, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().mgcd_autotune(ctx.h, _p32(limbs), u_off.ctypes.data_as(P64), _p32(nu), v_off.ctypes.data_as(P64),
                              _p32(nv), len(nu), ctypes.byref(t), ctypes.byref(k), ctypes.byref(g)),
          "mgcd_autotune")
    return t.value, k.value, g.value


class Batch:
    """Pairs resident in HBM, computed by the persistent kernel."""

    def __init__(self, ctx, limbs, u_off, nu, v_off, nv, workgroup=0, slots=0, group=0):
        self.ctx = ctx
        self.n = len(nu)
        self._keep = [np.ascontiguousarray(x) for x in (limbs, u_off, nu, v_off, nv)]
        limbs, u_off, nu, v_off, nv = self._keep
        self.nv = nv
        self.nu = nu
        h = ctypes.c_void_p()
        check(lib().mgcd_batch_create(ctx.h, _p32(limbs), u_off.ctypes.data_as(P64), _p32(nu),
                                      v_off.ctypes.data_as(P64), _p32(nv), self.n, workgroup, slots, group,
                                      ctypes.byref(h)), "mgcd_batch_create")
        self.h = h

    @classmethod
    def from_pairs(cls, ctx, pairs, **kw):
        return cls(ctx, *pack_pairs(pairs), **kw)

    def run(self):
        check(lib().mgcd_batch_run(self.h), "mgcd_batch_run")

    def status(self, i):
        return lib().mgcd_batch_status(self.h, i)

    def result(self, i):
        cap = int(max(self.nu[i], self.nv[i])) + 1
        g = np.zeros(cap, np.uint32)
        ng = ctypes.c_size_t()
        check(lib().mgcd_batch_result(self.h, i, _p32(g), cap, ctypes.byref(ng)), f"mgcd_batch_result[{i}]")
        return limbs_to_int(g[:ng.value])

    def results(self):
        return [self.result(i) for i in range(self.n)]

    def timing(self):
        d, h = ctypes.c_double(), ctypes.c_double()
        check(lib().mgcd_batch_timing(self.h, ctypes.byref(d), ctypes.byref(h)))
        return d.value, h.value

    def pair_stats(self, i):
        q, r, k = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().mgcd_batch_pair_stats(self.h, i, ctypes.byref(q), ctypes.byref(r), ctypes.byref(k)))
        return q.value, r.value, k.value

    def pivots(self, i):
        """Pair i's elimination pivots [(p, b), ...] from the last run."""
        n = ctypes.c_size_t()
        check(lib().mgcd_batch_pivots(self.h, i, None, 0, ctypes.byref(n)), f"mgcd_batch_pivots[{i}]")
        arr = (Pair * max(n.value, 1))()
        check(lib().mgcd_batch_pivots(self.h, i, arr, n.value, ctypes.byref(n)), f"mgcd_batch_pivots[{i}]")
        return [(arr[j].modulus, arr[j].value) for j in range(n.value)]

    def tiling(self):
        t, k, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mgcd_batch_tiling(self.h, ctypes.byref(t), ctypes.byref(k), ctypes.byref(g)))
        return t.value, k.value, g.value

    def launches(self):
        """Per launch: dict(tiling, pairs, nq range, device seconds of the last run)."""
        n = ctypes.c_size_t()
        check(lib().mgcd_batch_launch_count(self.h, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            t, k, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            npairs, lo, hi, sec = ctypes.c_size_t(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
            check(lib().mgcd_batch_launch_info(self.h, i, ctypes.byref(t), ctypes.byref(k), ctypes.byref(g),
                                               ctypes.byref(npairs), ctypes.byref(lo), ctypes.byref(hi),
                                               ctypes.byref(sec)))
            out.append({"tiling": (t.value, k.value, g.value), "pairs": npairs.value, "nq": (lo.value, hi.value),
                        "device_s": sec.value})
        return out

    def close(self):
        if self.h:
            lib().mgcd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (mgcd_server_*): submit() returns a
    ticket at once; wait(ticket) returns (gcd, latency seconds)."""

    def __init__(self, ctx, budget_bytes=64 << 30, slots=2):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().mgcd_server_create(ctx.h, budget_bytes, slots, ctypes.byref(h)), "mgcd_server_create")
        self.h = h
        self._cap = {}

    def submit(self, u, v):
        ua, va = int_to_limbs(abs(u)), int_to_limbs(abs(v))
        t = ctypes.c_uint64()
        check(lib().mgcd_server_submit(self.h, _p32(ua), len(ua), _p32(va), len(va), ctypes.byref(t)),
              "mgcd_server_submit")
        self._cap[t.value] = max(len(ua), len(va)) + 1
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """(gcd, latency seconds); raises MgcdError with status ERR_PENDING when
        the timeout expires first (the ticket stays valid: wait again)."""
        cap = self._cap[ticket]
        g = np.zeros(cap, np.uint32)
        ng, lat = ctypes.c_size_t(), ctypes.c_double()
        st = lib().mgcd_server_wait(self.h, ticket, timeout, _p32(g), cap, ctypes.byref(ng), ctypes.byref(lat))
        if st not in (ERR_PENDING, ERR_CAPACITY):
            del self._cap[ticket]
        check(st, "mgcd_server_wait")
        return limbs_to_int(g[:ng.value]), lat.value

    def stats(self):
        n, p50, p99, rate, peak = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), \
            ctypes.c_uint64()
        check(lib().mgcd_server_stats(self.h, ctypes.byref(n), ctypes.byref(p50), ctypes.byref(p99),
                                      ctypes.byref(rate), ctypes.byref(peak)))
        return {"completed": n.value, "p50_s": p50.value, "p99_s": p99.value, "gcds_per_s": rate.value,
                "peak_bytes": peak.value}

    def close(self):
        if self.h:
            lib().mgcd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
