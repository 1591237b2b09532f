"""ctypes binding of the CPU oracle (oracle/libsidon_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libsidon_oracle.so"
REF_PROBE = HERE / "_ref" / "probe"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ldp = np.ctypeslib.ndpointer(dtype=np.longdouble, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile the oracle library (and the reference probe when its source exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if Path("/root/reference/tools/probe.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    L.so_F_hostorder.argtypes = [_dp, C.c_int]
    L.so_F_hostorder.restype = C.c_double
    L.so_F_gather.argtypes = [_dp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.so_F_gather.restype = C.c_double
    L.so_verify_ld.argtypes = [_ldp, C.c_int] + [C.POINTER(C.c_longdouble)] * 3 + \
        [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_longdouble)]
    L.so_verify_ld.restype = C.c_int
    L.so_verify_file.argtypes = [C.c_char_p] + [C.POINTER(C.c_longdouble)] * 3 + \
        [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_longdouble), C.POINTER(C.c_int)]
    L.so_verify_file.restype = C.c_int
    L.so_verify_file_str.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.so_verify_file_str.restype = C.c_int
    L.so_F_exact.argtypes = [_dp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                             C.POINTER(C.c_int), C.c_void_p]
    L.so_F_exact.restype = C.c_double
    L.so_fsum.argtypes = [_dp, C.c_int]
    L.so_fsum.restype = C.c_double
    L.so_autoconv.argtypes = [_dp, C.c_int, _dp]
    L.so_corr.argtypes = [_dp, _dp, C.c_int, _dp]
    L.so_lp_weights.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp]
    L.so_lp_weights.restype = C.c_double
    L.so_Jp_grad.argtypes = [_dp, C.c_int, C.c_double, C.c_void_p]
    L.so_Jp_grad.restype = C.c_double
    L.so_Jp_grad_ld.argtypes = [_ldp, C.c_int, C.c_longdouble, C.c_void_p]
    L.so_Jp_grad_ld.restype = C.c_longdouble
    L.so_local_search.argtypes = [_dp, C.c_int, C.c_int] + [C.c_double] * 6 + \
        [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    L.so_local_search.restype = C.c_double
    L.so_local_search_gmin.argtypes = [_dp, C.c_int, C.c_int] + [C.c_double] * 6 + \
        [C.c_int, C.c_void_p, C.c_void_p, _dp]
    L.so_local_search_gmin.restype = C.c_double
    L.so_local_search_batch.argtypes = [_dp, C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + \
        [C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.so_frank_wolfe.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.POINTER(C.c_double)]
    L.so_frank_wolfe.restype = C.c_int
    L.so_frank_wolfe_trace.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.POINTER(C.c_double), _dp, _dp]
    L.so_frank_wolfe_trace.restype = C.c_int
    L.so_pdhg.argtypes = [_dp, C.c_int, C.c_double, C.c_int, _dp, _dp]
    L.so_refine.argtypes = [_dp, C.c_int, _dp]
    L.so_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.so_variation.argtypes = [_fp, C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_uint64,
                               C.c_int, _fp]
    L.so_variation2.argtypes = [_fp, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                C.c_uint64, C.c_int, _fp]
    L.so_op_params.argtypes = [C.c_uint64, C.c_uint64, C.c_int, _fp]
    L.so_op_uses_mean.argtypes = [C.c_int]
    L.so_op_uses_mean.restype = C.c_int
    L.so_support_sig.argtypes = [_fp, C.c_int, C.c_float, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_int)]
    L.so_sig_mix.argtypes = [C.c_int]
    L.so_sig_mix.restype = C.c_uint64
    L.so_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int]
    L.so_normal.restype = C.c_float
    L.so_select_plan.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                 C.c_int, C.c_float, _ip, _ip, _ip, _fp, _ip, _ip]
    L.so_max_threads.restype = C.c_int
    _lib = L
    return L


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def F_hostorder(a) -> float:
    a = _d(a)
    return lib().so_F_hostorder(a, a.size)


def F_gather(a):
    a = _d(a)
    M, k = C.c_double(), C.c_int()
    F = lib().so_F_gather(a, a.size, C.byref(M), C.byref(k))
    return F, M.value, k.value


def verify_ld(a):
    """Long-double F (probe.cpp --verify semantics) -> dict."""
    a = np.ascontiguousarray(a, dtype=np.longdouble)
    F, M, S, rb = (C.c_longdouble() for _ in range(4))
    km, na = C.c_int(), C.c_int()
    rc = lib().so_verify_ld(a, a.size, C.byref(F), C.byref(M), C.byref(S), C.byref(km),
                            C.byref(na), C.byref(rb))
    return dict(rc=rc, F=np.longdouble(F.value), M=np.longdouble(M.value),
                S=np.longdouble(S.value), kmax=km.value, nActive=na.value,
                relBound=np.longdouble(rb.value))


def verify_file(path):
    F, M, S, rb = (C.c_longdouble() for _ in range(4))
    km, na, n = C.c_int(), C.c_int(), C.c_int()
    rc = lib().so_verify_file(str(path).encode(), C.byref(F), C.byref(M), C.byref(S),
                              C.byref(km), C.byref(na), C.byref(rb), C.byref(n))
    return dict(rc=rc, F=F.value, M=M.value, S=S.value, kmax=km.value, nActive=na.value,
                relBound=rb.value, n=n.value)


def verify_file_str(path):
    """(rc, [S, M, kmax, nActive, F, relBound, F_upper] as probe prints them)."""
    buf = C.create_string_buffer(512)
    rc = lib().so_verify_file_str(str(path).encode(), buf, 512)
    return rc, buf.value.decode().split()


def write_vec(path, a, F) -> None:
    """Saved-vector text format (src/main.cu:329-338)."""
    with open(path, "w") as f:
        f.write("# AlphaEvolve problem 2 -- autocorrelation / Sidon\n")
        f.write(f"n {len(a)}\n")
        f.write(f"F {F:.17e}\n")
        for x in a:
            f.write(f"{float(x)!r}\n")


def F_exact(a, want_b=False):
    a = _d(a)
    M, S, k = C.c_double(), C.c_double(), C.c_int()
    b = np.zeros(2 * a.size - 1) if want_b else None
    F = lib().so_F_exact(a, a.size, C.byref(M), C.byref(S), C.byref(k),
                         b.ctypes.data if want_b else None)
    out = dict(F=F, M=M.value, S=S.value, kmax=k.value)
    if want_b:
        out["b"] = b
    return out


def fsum(x) -> float:
    x = _d(x)
    return lib().so_fsum(x, x.size)


def autoconv(a):
    a = _d(a)
    b = np.zeros(2 * a.size - 1)
    lib().so_autoconv(a, a.size, b)
    return b


def corr(c, a):
    a = _d(a)
    c = _d(c)
    assert c.size >= 2 * a.size - 1
    G = np.zeros(a.size)
    lib().so_corr(c, a, a.size, G)
    return G


def lp_weights(b, M, p):
    b = _d(b)
    c = np.zeros_like(b)
    R = lib().so_lp_weights(b, b.size, M, p, c)
    return c, R


def Jp_grad(a, p):
    a = _d(a)
    g = np.zeros_like(a)
    J = lib().so_Jp_grad(a, a.size, p, g.ctypes.data)
    return J, g


def Jp(a, p) -> float:
    a = _d(a)
    return lib().so_Jp_grad(a, a.size, p, None)


def Jp_grad_ld(a, p):
    a = np.ascontiguousarray(a, dtype=np.longdouble)
    g = np.zeros_like(a)
    J = lib().so_Jp_grad_ld(a, a.size, p, g.ctypes.data)
    return np.longdouble(J), g


def local_search(a, iters, pA, pB, lrA, lrB, b1=0.9, b2=0.999, mode=0, probe_semantics=False):
    """Returns (a_out, F, m, v, iters_done); a is not modified."""
    a = _d(a).copy()
    m = np.zeros_like(a)
    v = np.zeros_like(a)
    done = C.c_int()
    F = lib().so_local_search(a, a.size, iters, pA, pB, lrA, lrB, b1, b2, mode,
                              int(probe_semantics), m.ctypes.data, v.ctypes.data,
                              C.byref(done))
    return a, F, m, v, done.value


def local_search_gmin(a, iters, pA, pB, lrA, lrB, b1=0.9, b2=0.999, mode=0):
    """Kernel-semantics local search; returns (a_out, F, m, v, gmin) with
    gmin[i] = min over the steps of |g_i| / M."""
    a = _d(a).copy()
    m = np.zeros_like(a)
    v = np.zeros_like(a)
    gmin = np.zeros_like(a)
    F = lib().so_local_search_gmin(a, a.size, iters, pA, pB, lrA, lrB, b1, b2, mode,
                                   m.ctypes.data, v.ctypes.data, gmin)
    return a, F, m, v, gmin


def local_search_batch(pop, iters, pA, pB, lrA, lrB, b1=0.9, b2=0.999, mode=0,
                       probe_semantics=True, nthreads=0):
    pop = _d(pop).copy()
    P, n = pop.shape
    F = np.zeros(P)
    lib().so_local_search_batch(pop, P, n, iters, pA, pB, lrA, lrB, b1, b2, mode,
                                int(probe_semantics), nthreads, F.ctypes.data)
    return pop, F


def frank_wolfe(a, rounds, cp_iters, nT=12, tMax=0.30):
    a = _d(a).copy()
    M = C.c_double()
    acc = lib().so_frank_wolfe(a, a.size, rounds, cp_iters, nT, tMax, C.byref(M))
    return a, acc, M.value


def frank_wolfe_trace(a, rounds, cp_iters, nT=12, tMax=0.30):
    """(a_out, accepted, M, g, lsM): g the last round's normalised LP
    direction, lsM the line-search maxima of t = tMax s/nT, s = 1..nT."""
    a = _d(a).copy()
    M = C.c_double()
    g = np.zeros_like(a)
    lsM = np.zeros(nT)
    acc = lib().so_frank_wolfe_trace(a, a.size, rounds, cp_iters, nT, tMax, C.byref(M), g, lsM)
    return a, acc, M.value, g, lsM


def pdhg(a, M, cp_iters):
    a = _d(a)
    g = np.zeros_like(a)
    lam = np.zeros(2 * a.size - 1)
    lib().so_pdhg(a, a.size, M, cp_iters, g, lam)
    return g, lam


def refine(a):
    a = _d(a)
    out = np.zeros(2 * a.size)
    lib().so_refine(a, a.size, out)
    return out


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().so_philox4x32_10(c, k, o)
    return o


def variation(A, op, sigma, seed, gen, slot):
    A = np.ascontiguousarray(A, dtype=np.float32)
    out = np.zeros_like(A)
    lib().so_variation(A, A.size, op, sigma, seed, gen, slot, out)
    return out


def variation2(A, B, op, sigma, seed, gen, slot):
    """All fourteen operators (B = None: B = A)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    out = np.zeros_like(A)
    if B is None:
        lib().so_variation2(A, None, A.size, op, sigma, seed, gen, slot, out)
    else:
        B = np.ascontiguousarray(B, dtype=np.float32)
        assert B.size == A.size
        lib().so_variation2(A, B.ctypes.data, A.size, op, sigma, seed, gen, slot, out)
    return out


def op_params(seed, gen, slot):
    prm = np.zeros(64, np.float32)
    lib().so_op_params(seed, gen, slot, prm)
    return prm


def op_uses_mean(op) -> bool:
    return bool(lib().so_op_uses_mean(op))


def support_sig(a, tol=0.02):
    """(signature, support size) as kSupportSig computes them."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    sig, nnz = C.c_uint64(), C.c_int()
    lib().so_support_sig(a, a.size, tol, C.byref(sig), C.byref(nnz))
    return sig.value, nnz.value


def sig_mix(i) -> int:
    return lib().so_sig_mix(i)


def normal(seed, gen, slot, i) -> float:
    return lib().so_normal(seed, gen, slot, i)


def select_plan(fit, NI, elite, seed, gen, migrate_every=12, sigma=0.05):
    fit = _d(fit).copy()
    P = fit.size
    p1 = np.zeros(P, np.int32)
    p2 = np.zeros(P, np.int32)
    op = np.zeros(P, np.int32)
    st = np.zeros(P, np.float32)
    ms = np.zeros(NI, np.int32)
    md = np.zeros(NI, np.int32)
    lib().so_select_plan(fit, P, NI, elite, seed, gen, migrate_every, sigma, p1, p2, op, st,
                         ms, md)
    return dict(fit=fit, p1=p1, p2=p2, op=op, strength=st, mig_src=ms, mig_dst=md)


def max_threads() -> int:
    return lib().so_max_threads()


def ref_probe_available() -> bool:
    return REF_PROBE.exists() and os.access(REF_PROBE, os.X_OK)
