"""Host-side mirror of the reference's search-path interface over the gfx950 C ABI.

The reference keeps its device state in ``struct Gpu`` (src/main.cu:107-144)
and drives the kernels through ``runLocalSearch`` / ``runEval`` /
``runFrankWolfe`` (src/main.cu:146-184) plus an inline ``kVariation`` launch
(src/main.cu:737-742).  This module offers the same names and argument meaning
over ``libsidon_amd.so`` (include/sidon.h): torch tensors provide device memory
and the current HIP stream, every compute call goes through the C ABI, and a
failing status raises :class:`SidonError`.  There is no CPU fallback: if the
HIP library is missing or no GPU is visible, calls raise.

Import it with ``importlib.import_module("sidon-autocorrelation-bound_amd")``
(the directory name is not a Python identifier).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import torch

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libsidon_amd.so"
INCLUDE = HERE.parent / "include" / "sidon.h"

SIDON_OK, SIDON_ERR_ARG, SIDON_ERR_LDS, SIDON_ERR_HIP, SIDON_ERR_UNSUPPORTED = range(5)

# every entry point include/sidon.h declares
EXPORTS = (
    "sidon_init", "sidon_last_error", "sidon_status_string", "sidon_lds_bytes",
    "sidon_lds_limit", "sidon_eval_f64", "sidon_eval_f64_ragged", "sidon_eval_f64_d",
    "sidon_eval_exact", "sidon_local_search", "sidon_local_search_f64",
    "sidon_local_search_ragged", "sidon_local_search_ragged_workspace",
    "sidon_polish_f64_workspace", "sidon_polish_f64",
    "sidon_polish_f64_noalloc",
    "sidon_set_ls_block", "sidon_get_ls_block", "sidon_set_ls_tile", "sidon_get_ls_tile",
    "sidon_ls_tile_config", "sidon_lp_grad",
    "sidon_lp_grad_f64", "sidon_frank_wolfe", "sidon_refine", "sidon_variation",
    "sidon_variation_ragged", "sidon_support_sig", "sidon_select", "sidon_advance_gen", "sidon_order_by_length",
    "sidon_generation_create", "sidon_generation_run", "sidon_generation_current",
    "sidon_generation_enqueue", "sidon_generation_destroy", "sidon_tune_local_search",
    "sidon_measure_fp32_ceilings",
    "sidon_ls_block_steps", "sidon_ls_phase_clocks", "sidon_fw_block_steps",
    "sidon_policy_create", "sidon_policy_destroy", "sidon_policy_adopt", "sidon_policy_new_level",
    "sidon_policy_observe", "sidon_policy_plan", "sidon_policy_credit", "sidon_policy_stats",
    "sidon_policy_tiers",
)


class GenerationDesc(C.Structure):
    """sidon_generation_desc (include/sidon.h)."""
    _fields_ = [(name, C.c_void_p) for name in (
        "pop_a", "pop_b", "ns_a", "ns_b", "fit", "nact", "p1", "p2", "op", "strength", "order",
        "gen", "ls_workspace")] + [("ls_workspace_bytes", C.c_uint64)] + [
            (name, C.c_int) for name in (
            "P", "stride", "nmax", "islands", "elite", "migrate_every", "iters", "split")] + [
        (name, C.c_float) for name in ("pStart", "pEnd", "lrStart", "lrEnd", "sigma")] + [
        ("activeTol", C.c_double), ("seed", C.c_uint64)]


class SidonError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        names = {1: "SIDON_ERR_ARG", 2: "SIDON_ERR_LDS", 3: "SIDON_ERR_HIP",
                 4: "SIDON_ERR_UNSUPPORTED"}
        super().__init__(f"{where}: {names.get(status, status)} {detail}".strip())


def build(verbose: bool = False) -> Path:
    """Compile libsidon_amd.so for gfx950 (always recompiles out-of-date objects)."""
    out = subprocess.run(["make", "-C", str(HERE), "-j8"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("HIP build failed:\n" + out.stdout[-4000:] + out.stderr[-4000:])
    if verbose:
        print(out.stdout)
    return LIB_PATH


_lib = None

_vp = C.c_void_p
_i, _f, _d, _u64 = C.c_int, C.c_float, C.c_double, C.c_uint64


def _declare(L):
    sig = {
        "sidon_init": ([_i], _i),
        "sidon_last_error": ([], C.c_char_p),
        "sidon_status_string": ([_i], C.c_char_p),
        "sidon_lds_bytes": ([_i, _i, _i], C.c_longlong),
        "sidon_lds_limit": ([], _i),
        "sidon_eval_f64": ([_vp, _i, _i, _vp, _vp, _d, _vp], _i),
        "sidon_eval_f64_ragged": ([_vp, _i, _vp, _i, _i, _vp, _vp, _d, _vp], _i),
        "sidon_eval_f64_d": ([_vp, _i, _i, _vp, _vp, _d, _vp], _i),
        "sidon_eval_exact": ([_vp, _i, _i, _vp, _vp, _vp, _vp, _vp], _i),
        "sidon_local_search": ([_vp, _vp, _vp, _vp, _vp, _i, _i, _i] + [_f] * 6 + [_i, _vp], _i),
        "sidon_local_search_f64": ([_vp, _vp, _vp, _vp, _vp, _i, _i, _i] + [_d] * 6 + [_i, _vp],
                                   _i),
        "sidon_local_search_ragged": ([_vp, _i, _vp, _i, _i, _vp, _vp, _vp, _i] + [_f] * 6 +
                                      [_i, _i, _vp, C.c_size_t, _vp], _i),
        "sidon_local_search_ragged_workspace": ([_i, _i], C.c_size_t),
        "sidon_set_ls_block": ([_i, _i], _i),
        "sidon_get_ls_block": ([_i], _i),
        "sidon_set_ls_tile": ([_i, _i, _i, _i], _i),
        "sidon_get_ls_tile": ([_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)], _i),
        "sidon_ls_tile_config": ([_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)], _i),
        "sidon_lp_grad": ([_vp, _i, _i, _f, _vp, _vp, _vp], _i),
        "sidon_lp_grad_f64": ([_vp, _i, _i, _d, _vp, _vp, _vp], _i),
        "sidon_frank_wolfe": ([_vp, _vp, _i, _i, _i, _i, _i, _f, _vp], _i),
        "sidon_refine": ([_vp, _i, _i, _vp, _vp], _i),
        "sidon_variation": ([_vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _u64, _u64, _vp], _i),
        "sidon_support_sig": ([_vp, _i, _i, C.c_float, _vp, _vp, _vp], _i),
        "sidon_polish_f64_workspace": ([_i, _i], C.c_size_t),
        "sidon_polish_f64": ([_vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _vp,
                              C.c_size_t, _vp], _i),
        "sidon_polish_f64_noalloc": ([_vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i,
                                      _vp], _i),
        "sidon_policy_create": ([_i, _i, _i, _u64, _vp], _i),
        "sidon_policy_destroy": ([_vp], None),
        "sidon_policy_adopt": ([_vp, _d], _i),
        "sidon_policy_new_level": ([_vp], _i),
        "sidon_policy_observe": ([_vp, _vp, _vp, _vp, _vp], _i),
        "sidon_policy_plan": ([_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
        "sidon_policy_credit": ([_vp, _vp], _i),
        "sidon_policy_stats": ([_vp, _vp, _vp, _vp, _vp, _vp], _i),
        "sidon_policy_tiers": ([_vp, _i, _vp, _vp, _vp], _i),
        "sidon_variation_ragged": ([_vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _u64, _vp,
                                    _vp], _i),
        "sidon_select": ([_vp, _vp, _vp, _i, _i, _i, _i, _i, _f, _u64, _vp, _vp, _vp, _vp, _vp,
                          _vp], _i),
        "sidon_advance_gen": ([_vp, _u64, _vp], _i),
        "sidon_order_by_length": ([_vp, _i, _vp, _vp], _i),
        "sidon_generation_create": ([C.POINTER(GenerationDesc), C.POINTER(_vp), _vp], _i),
        "sidon_generation_run": ([_vp, _i, _vp], _i),
        "sidon_generation_current": ([_vp, C.POINTER(_i)], _i),
        "sidon_generation_enqueue": ([C.POINTER(GenerationDesc), _i, _vp], _i),
        "sidon_generation_destroy": ([_vp], _i),
        "sidon_tune_local_search": ([_vp, _vp, _vp, _i, _i, _i, _i, C.POINTER(_i),
                                     C.POINTER(_i), C.POINTER(_i), C.POINTER(_f),
                                     C.POINTER(_f), _vp], _i),
        "sidon_measure_fp32_ceilings": ([_vp, _i, C.POINTER(_d), C.POINTER(_d), _vp], _i),
        "sidon_ls_block_steps": ([C.POINTER(C.c_ulonglong), _i], _i),
        "sidon_ls_phase_clocks": ([C.POINTER(C.c_ulonglong), _i], _i),
        "sidon_fw_block_steps": ([C.POINTER(C.c_ulonglong), _i], _i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def use_phase_clock_library() -> None:
    """Load the profiling build of the library (libsidon_amd_clocks.so, `make
    clocks`: the local search records per-phase shader-clock sums) instead of
    the product one.  Must come before the first call into the library."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the library is already loaded")
    LIB_PATH = HERE / "libsidon_amd_clocks.so"


def ls_phase_clocks(reset: bool = False) -> list:
    """The profiling build's phase sums (sidon_ls_phase_clocks): cycles of the
    autoconvolution, weights, correlation and update phases, idle wave-cycles
    at the end of the two product phases, steps, candidates."""
    out = (C.c_ulonglong * 8)()
    _check(lib().sidon_ls_phase_clocks(out, int(reset)), "sidon_ls_phase_clocks")
    return [int(x) for x in out]


def lib():
    """The loaded C ABI.  Raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build the HIP library first "
                               "(make -C sidon-autocorrelation-bound_amd)")
        L = C.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def _ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device tensor required")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor required")
    return t.data_ptr()


def _stream(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _check(rc: int, where: str) -> None:
    if rc != SIDON_OK:
        detail = lib().sidon_last_error().decode(errors="replace")
        raise SidonError(rc, where, detail)


def _need_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("the Sidon HIP path needs a visible GPU; there is no CPU fallback")


def init(device: int = -1) -> None:
    _need_gpu()
    _check(lib().sidon_init(device), "sidon_init")


def lds_limit() -> int:
    return lib().sidon_lds_limit()


def lds_bytes(kernel: int, n: int, block: int) -> int:
    return lib().sidon_lds_bytes(kernel, n, block)


# ------------------------------------------------------------- raw calls ---

def eval_f64(pop: torch.Tensor, active_tol: float = 1e-4, with_active: bool = True,
             stream=None):
    """F (fp64) and near-active peak counts of each row: kEval / runEval."""
    _need_gpu()
    P, n = pop.shape
    F = torch.empty(P, dtype=torch.float64, device=pop.device)
    na = torch.empty(P, dtype=torch.int32, device=pop.device) if with_active else None
    if pop.dtype == torch.float32:
        rc = lib().sidon_eval_f64(_ptr(pop), n, P, _ptr(F), _ptr(na), active_tol, _stream(stream))
    elif pop.dtype == torch.float64:
        rc = lib().sidon_eval_f64_d(_ptr(pop), n, P, _ptr(F), _ptr(na), active_tol,
                                    _stream(stream))
    else:
        raise TypeError("pop must be float32 or float64")
    _check(rc, "sidon_eval_f64")
    return F, na


def eval_f64_ragged(pop: torch.Tensor, ns: torch.Tensor, nmax: int, active_tol: float = 1e-4,
                    stream=None):
    _need_gpu()
    P, stride = pop.shape
    F = torch.empty(P, dtype=torch.float64, device=pop.device)
    na = torch.empty(P, dtype=torch.int32, device=pop.device)
    _check(lib().sidon_eval_f64_ragged(_ptr(pop), stride, _ptr(ns), P, nmax, _ptr(F), _ptr(na),
                                       active_tol, _stream(stream)), "sidon_eval_f64_ragged")
    return F, na


def eval_exact(pop: torch.Tensor, stream=None):
    """Correctly rounded F, M, S and the ambiguity count of each fp32 row."""
    _need_gpu()
    P, n = pop.shape
    dev = pop.device
    F = torch.empty(P, dtype=torch.float64, device=dev)
    M = torch.empty(P, dtype=torch.float64, device=dev)
    S = torch.empty(P, dtype=torch.float64, device=dev)
    amb = torch.empty(P, dtype=torch.int32, device=dev)
    _check(lib().sidon_eval_exact(_ptr(pop), n, P, _ptr(F), _ptr(M), _ptr(S), _ptr(amb),
                                  _stream(stream)), "sidon_eval_exact")
    return F, M, S, amb


def local_search(pop: torch.Tensor, iters: int, p_start: float, p_end: float, lr_start: float,
                 lr_end: float, beta1: float = 0.9, beta2: float = 0.999, mode: int = 0,
                 mode_arr: torch.Tensor | None = None, skip: torch.Tensor | None = None,
                 mom: torch.Tensor | None = None, vel: torch.Tensor | None = None,
                 stream=None) -> None:
    """In-place fused local search on every row of pop (float32 or float64).
    float32 rows longer than the on-chip form (lds_bytes(0, n, 1024) grows past
    the G-in-LDS layout at n = 9296) keep the correlation output in global
    memory and need `mom` or `vel`: the launch uses it as that scratch (both,
    when both are given: the two-row form, with no atomic adds in memory)."""
    _need_gpu()
    P, n = pop.shape
    if pop.dtype == torch.float32:
        rc = lib().sidon_local_search(_ptr(pop), _ptr(mom), _ptr(vel), _ptr(skip), _ptr(mode_arr),
                                      n, P, iters, p_start, p_end, lr_start, lr_end, beta1, beta2,
                                      mode, _stream(stream))
    elif pop.dtype == torch.float64:
        rc = lib().sidon_local_search_f64(_ptr(pop), _ptr(mom), _ptr(vel), _ptr(skip),
                                          _ptr(mode_arr), n, P, iters, p_start, p_end, lr_start,
                                          lr_end, beta1, beta2, mode, _stream(stream))
    else:
        raise TypeError("pop must be float32 or float64")
    _check(rc, "sidon_local_search")


def local_search_ragged(pop: torch.Tensor, ns: torch.Tensor, nmax: int, iters: int,
                        p_start: float, p_end: float, lr_start: float, lr_end: float,
                        beta1: float = 0.9, beta2: float = 0.999, mode: int = 0,
                        order: torch.Tensor | None = None, skip: torch.Tensor | None = None,
                        mode_arr: torch.Tensor | None = None, split: int = 1,
                        workspace: torch.Tensor | None = None, stream=None) -> None:
    """Persistent uneven-length local search.  `workspace` (uint8 device
    tensor of ragged_workspace_bytes(P, stride)) holds the work queue and the
    buffers of candidates split over two workgroups (split = 1); a fresh one
    when not given.  Launches that may overlap need distinct workspaces."""
    _need_gpu()
    P, stride = pop.shape
    if workspace is None:
        workspace = torch.empty(ragged_workspace_bytes(P, stride), dtype=torch.uint8,
                                device=pop.device)
    _check(lib().sidon_local_search_ragged(_ptr(pop), stride, _ptr(ns), P, nmax, _ptr(order),
                                           _ptr(skip), _ptr(mode_arr), iters, p_start, p_end,
                                           lr_start, lr_end, beta1, beta2, mode, split,
                                           _ptr(workspace), workspace.numel(), _stream(stream)),
           "sidon_local_search_ragged")


def ragged_workspace_bytes(P: int, stride: int) -> int:
    return lib().sidon_local_search_ragged_workspace(P, stride)


def pair_timeouts(workspace: torch.Tensor) -> int:
    """Pair waits of the last ragged launch on `workspace` that gave up (0 in
    a normal run): the int at byte 4."""
    return int(workspace[4:8].cpu().view(torch.int32).item())


def split_count(workspace: torch.Tensor) -> int:
    """Candidates the last ragged launch on `workspace` split over two
    workgroups: the int at byte 8."""
    return int(workspace[8:12].cpu().view(torch.int32).item())


def set_ls_block(n: int, block: int) -> None:
    _check(lib().sidon_set_ls_block(n, block), "sidon_set_ls_block")


def get_ls_block(n: int) -> int:
    return lib().sidon_get_ls_block(n)


LS_TILES = 10


def set_ls_tile(n: int, block: int, depth: int = 4, per_cu: int = 1) -> None:
    """Launch configuration of the fp32 local search at length n: workgroup
    size, operand steps in flight (2 or 4), candidates a CU (2 only with 512
    threads and n <= 4096, 4 only with 256 threads and n <= 2304).  block 0
    restores the built-in rule."""
    _check(lib().sidon_set_ls_tile(n, block, depth, per_cu), "sidon_set_ls_tile")


def get_ls_tile(n: int) -> tuple[int, int, int]:
    b, d, c = C.c_int(), C.c_int(), C.c_int()
    _check(lib().sidon_get_ls_tile(n, C.byref(b), C.byref(d), C.byref(c)), "sidon_get_ls_tile")
    return b.value, d.value, c.value


def ls_tile_config(i: int) -> tuple[int, int, int]:
    """The autotuner's i-th configuration (block, depth, per_cu)."""
    b, d, c = C.c_int(), C.c_int(), C.c_int()
    _check(lib().sidon_ls_tile_config(i, C.byref(b), C.byref(d), C.byref(c)),
           "sidon_ls_tile_config")
    return b.value, d.value, c.value


def lp_grad(a: torch.Tensor, p: float, want_J: bool = True, want_grad: bool = True,
            stream=None):
    """J_p and dJ_p/da for every row (b recomputed on the device, never stored)."""
    _need_gpu()
    B, n = a.shape
    J = torch.empty(B, dtype=a.dtype, device=a.device) if want_J else None
    g = torch.empty_like(a) if want_grad else None
    fn = lib().sidon_lp_grad if a.dtype == torch.float32 else lib().sidon_lp_grad_f64
    _check(fn(_ptr(a), n, B, p, _ptr(J), _ptr(g), _stream(stream)), "sidon_lp_grad")
    return J, g


def frank_wolfe(pop: torch.Tensor, rounds: int, cp_iters: int, nT: int = 12,
                t_max: float = 0.30, skip: torch.Tensor | None = None, stream=None) -> None:
    _need_gpu()
    P, n = pop.shape
    _check(lib().sidon_frank_wolfe(_ptr(pop), _ptr(skip), n, P, rounds, cp_iters, nT, t_max,
                                   _stream(stream)), "sidon_frank_wolfe")


def refine(src: torch.Tensor, stream=None) -> torch.Tensor:
    """n -> 2n cell split of every row."""
    _need_gpu()
    P, n = src.shape
    dst = torch.empty(P, 2 * n, dtype=torch.float32, device=src.device)
    _check(lib().sidon_refine(_ptr(src), n, P, _ptr(dst), _stream(stream)), "sidon_refine")
    return dst


def variation(pop: torch.Tensor, out: torch.Tensor, p1, p2, op, strength, seed: int, gen: int,
              stream=None) -> None:
    _need_gpu()
    P, n = pop.shape
    _check(lib().sidon_variation(_ptr(pop), _ptr(out), n, out.shape[0], _ptr(p1), _ptr(p2),
                                 _ptr(op), _ptr(strength), seed, gen, _stream(stream)),
           "sidon_variation")


def variation_ragged(pop, ns, out, ns_out, p1, p2, op, strength, seed: int, gen_dev,
                     stream=None) -> None:
    _need_gpu()
    stride = pop.shape[1]
    _check(lib().sidon_variation_ragged(_ptr(pop), _ptr(ns), _ptr(out), _ptr(ns_out), stride,
                                        out.shape[0], _ptr(p1), _ptr(p2), _ptr(op),
                                        _ptr(strength), seed, _ptr(gen_dev), _stream(stream)),
           "sidon_variation_ragged")


def polish_f64(pop: torch.Tensor, iters: int, pStart: float, pEnd: float, lrStart: float,
               lrEnd: float, beta1: float = 0.9, beta2: float = 0.999, mode: int = 0,
               mom=None, vel=None, workspace=None, stream=None) -> None:
    """fp64 local search of the rows of `pop` (float64, K x n) spread over the
    whole device: sidon_polish_f64.  `workspace` (uint8 device tensor) is
    allocated here when not given."""
    _need_gpu()
    if pop.dtype != torch.float64:
        raise ValueError("polish_f64 needs float64 rows")
    K, n = pop.shape
    need = lib().sidon_polish_f64_workspace(n, K)
    if workspace is None:
        workspace = torch.empty(need, dtype=torch.uint8, device=pop.device)
    _check(lib().sidon_polish_f64(_ptr(pop), _ptr(mom), _ptr(vel), n, K, iters, pStart, pEnd,
                                  lrStart, lrEnd, beta1, beta2, mode, _ptr(workspace),
                                  workspace.numel(), _stream(stream)), "sidon_polish_f64")


def polish_f64_noalloc(pop: torch.Tensor, iters: int, pStart: float, pEnd: float,
                       lrStart: float, lrEnd: float, beta1: float = 0.9, beta2: float = 0.999,
                       mode: int = 0, mom=None, vel=None, stream=None) -> None:
    """sidon_polish_f64_noalloc: the fp64 polish on the library's static
    scratch array (no workspace, no allocation; K x n must fit 32 MiB of
    workspace)."""
    _need_gpu()
    if pop.dtype != torch.float64:
        raise ValueError("polish_f64_noalloc needs float64 rows")
    K, n = pop.shape
    _check(lib().sidon_polish_f64_noalloc(_ptr(pop), _ptr(mom), _ptr(vel), n, K, iters, pStart,
                                          pEnd, lrStart, lrEnd, beta1, beta2, mode,
                                          _stream(stream)), "sidon_polish_f64_noalloc")


def support_sig(pop: torch.Tensor, tol: float = 0.02, stream=None):
    """(sig uint64 as int64, nnz int32) of each row: kSupportSig."""
    _need_gpu()
    P, n = pop.shape
    sig = torch.empty(P, dtype=torch.int64, device=pop.device)
    nnz = torch.empty(P, dtype=torch.int32, device=pop.device)
    _check(lib().sidon_support_sig(_ptr(pop), n, P, tol, _ptr(sig), _ptr(nnz), _stream(stream)),
           "sidon_support_sig")
    return sig, nnz


class Policy:
    """The island search policy (sidon_policy_*): host-only, no device needed.
    Arrays are numpy (host) arrays of length P."""

    OPS = 14

    def __init__(self, P: int, islands: int, elite: int, seed: int):
        import numpy as np
        self._np = np
        self.P = P
        h = C.c_void_p()
        _check(lib().sidon_policy_create(P, islands, elite, seed, C.byref(h)),
               "sidon_policy_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sidon_policy_destroy(h)
            self._h = None

    def adopt(self, f: float) -> None:
        _check(lib().sidon_policy_adopt(self._h, f), "sidon_policy_adopt")

    def new_level(self) -> None:
        _check(lib().sidon_policy_new_level(self._h), "sidon_policy_new_level")

    def observe(self, fit, sig):
        np = self._np
        fit = np.ascontiguousarray(fit, np.float64)
        sig = np.ascontiguousarray(sig, np.uint64)
        bi, imp = C.c_int(), C.c_int()
        _check(lib().sidon_policy_observe(self._h, fit.ctypes.data, sig.ctypes.data,
                                          C.byref(bi), C.byref(imp)), "sidon_policy_observe")
        return bi.value, bool(imp.value)

    def plan(self, gen: int) -> dict:
        np = self._np
        P = self.P
        p1, p2, op = (np.zeros(P, np.int32) for _ in range(3))
        st = np.zeros(P, np.float32)
        ms, md = np.zeros(P, np.int32), np.zeros(P, np.int32)
        nm, rs = C.c_int(), C.c_int()
        _check(lib().sidon_policy_plan(self._h, gen, p1.ctypes.data, p2.ctypes.data,
                                       op.ctypes.data, st.ctypes.data, ms.ctypes.data,
                                       md.ctypes.data, C.byref(nm), C.byref(rs)),
               "sidon_policy_plan")
        return dict(p1=p1, p2=p2, op=op, strength=st, mig_src=ms[:nm.value],
                    mig_dst=md[:nm.value], reseed=rs.value)

    def credit(self, child_fit) -> None:
        np = self._np
        cf = np.ascontiguousarray(child_fit, np.float64)
        _check(lib().sidon_policy_credit(self._h, cf.ctypes.data), "sidon_policy_credit")

    def stats(self) -> dict:
        np = self._np
        w = np.zeros(self.OPS)
        bf, lb = C.c_double(), C.c_double()
        cells, since = C.c_int(), C.c_int()
        _check(lib().sidon_policy_stats(self._h, w.ctypes.data, C.byref(bf), C.byref(lb),
                                        C.byref(cells), C.byref(since)), "sidon_policy_stats")
        return dict(weights=w, best=bf.value, level_best=lb.value, cells=cells.value,
                    since_improved=since.value)

    @staticmethod
    def tiers(cur_op):
        import numpy as np
        cur = np.ascontiguousarray(cur_op, np.int32)
        tier = np.zeros(cur.size, np.int32)
        nc, nm = C.c_int(), C.c_int()
        _check(lib().sidon_policy_tiers(cur.ctypes.data, cur.size, tier.ctypes.data,
                                        C.byref(nc), C.byref(nm)), "sidon_policy_tiers")
        return tier, nc.value, nm.value


def select(fit, pop, ns, NI: int, elite: int, migrate_every: int, sigma: float, seed: int,
           gen_dev, p1, p2, op, strength, stream=None) -> None:
    _need_gpu()
    P = fit.shape[0]
    stride = pop.shape[1] if pop is not None else 0
    _check(lib().sidon_select(_ptr(fit), _ptr(pop), _ptr(ns), stride, P, NI, elite,
                              migrate_every, sigma, seed, _ptr(gen_dev), _ptr(p1), _ptr(p2),
                              _ptr(op), _ptr(strength), _stream(stream)), "sidon_select")


def advance_gen(gen_dev, delta: int = 1, stream=None) -> None:
    _check(lib().sidon_advance_gen(_ptr(gen_dev), delta, _stream(stream)), "sidon_advance_gen")


def order_by_length(ns: torch.Tensor, order: torch.Tensor, stream=None) -> None:
    _check(lib().sidon_order_by_length(_ptr(ns), ns.shape[0], _ptr(order), _stream(stream)),
           "sidon_order_by_length")


def measure_fp32_ceilings(stream=None):
    """(VALU, MFMA) FP32 TFLOP/s measured with register-only kernels on every CU."""
    _need_gpu()
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    scratch = torch.empty(cus * 8 * 256, dtype=torch.float32, device="cuda")
    v, m = C.c_double(), C.c_double()
    _check(lib().sidon_measure_fp32_ceilings(_ptr(scratch), scratch.numel(), C.byref(v),
                                             C.byref(m), _stream(stream)),
           "sidon_measure_fp32_ceilings")
    return v.value, m.value


def ls_block_steps(reset: bool = False) -> int:
    """Matrix block steps (4 x 2048 FLOP each) the local-search and gradient
    kernels have executed on the current device since the last reset
    (synchronous)."""
    _need_gpu()
    out = C.c_ulonglong()
    _check(lib().sidon_ls_block_steps(C.byref(out), int(reset)), "sidon_ls_block_steps")
    return out.value


def fw_block_steps(reset: bool = False) -> int:
    """Matrix block steps (4 x 2048 FLOP each) the Frank-Wolfe kernel has
    executed on the current device since the last reset (synchronous)."""
    _need_gpu()
    out = C.c_ulonglong()
    _check(lib().sidon_fw_block_steps(C.byref(out), int(reset)), "sidon_fw_block_steps")
    return out.value


def tune_local_search(n: int, P: int = 64, iters: int = 20, repeats: int = 2, seed: int = 1,
                      with_rows: bool = True, stream=None):
    """Autotune the launch configuration of the fused local search at length n
    (workgroup size, operand steps in flight, candidates a CU) on a
    structured-prior sample; the choice is kept by the library for n.  With
    `with_rows` the trials run with both moments' buffers (the correlation
    output in their rows), as launches that pass them do.  Returns ((block, depth,
    per_cu), ms of the choice, {configuration: ms} of every trial that fits
    n)."""
    _need_gpu()
    from .workloads import structured_prior
    ref = torch.as_tensor(structured_prior(P, n, seed=seed), device="cuda")
    scratch = torch.empty_like(ref)
    rows = torch.empty((2,) + tuple(ref.shape), dtype=ref.dtype, device=ref.device) if with_rows else None
    blk, dep, per, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
    trials = (C.c_float * LS_TILES)()
    _check(lib().sidon_tune_local_search(_ptr(scratch), _ptr(ref), _ptr(rows), n, P, iters, repeats,
                                         C.byref(blk), C.byref(dep), C.byref(per), C.byref(ms),
                                         trials, _stream(stream)),
           "sidon_tune_local_search")
    table = {ls_tile_config(i): float(trials[i]) for i in range(LS_TILES) if trials[i] >= 0}
    return (blk.value, dep.value, per.value), ms.value, table


class Generation:
    """Island-search generations captured as HIP graphs (sidon_generation_*).

    Buffers are torch tensors owned here; populations are [P x stride] with
    per-slot lengths.  `run(k)` replays k generations; `population()` returns
    the current (pop, ns) pair."""

    def __init__(self, pop: torch.Tensor, ns: torch.Tensor, islands: int, iters: int,
                 p_start: float = 4.0, p_end: float = 1e4, lr_start: float = 0.085,
                 lr_end: float = 0.0015, sigma: float = 0.05, elite: int = 2,
                 migrate_every: int = 12, seed: int = 20260905, active_tol: float = 1e-4,
                 graph: bool = True, split: int = 1, stream=None):
        _need_gpu()
        P, stride = pop.shape
        dev = pop.device
        i32 = dict(dtype=torch.int32, device=dev)
        self.pop_a = pop.contiguous()
        self.pop_b = torch.zeros_like(self.pop_a)
        self.ns_a = ns.to(torch.int32).contiguous()
        self.ns_b = self.ns_a.clone()
        self.fit = torch.zeros(P, dtype=torch.float64, device=dev)
        self.nact = torch.zeros(P, **i32)
        self.p1 = torch.zeros(P, **i32)
        self.p2 = torch.zeros(P, **i32)
        self.op = torch.zeros(P, **i32)
        self.strength = torch.zeros(P, dtype=torch.float32, device=dev)
        self.order = torch.zeros(P, **i32)
        self.gen = torch.zeros(1, dtype=torch.int64, device=dev)
        self.ls_workspace = torch.empty(ragged_workspace_bytes(P, stride), dtype=torch.uint8,
                                        device=dev)
        nmax = int(self.ns_a.max().item())
        d = GenerationDesc()
        for name in ("pop_a", "pop_b", "ns_a", "ns_b", "fit", "nact", "p1", "p2", "op",
                     "strength", "order", "gen", "ls_workspace"):
            setattr(d, name, _ptr(getattr(self, name)))
        d.ls_workspace_bytes = self.ls_workspace.numel()
        d.P, d.stride, d.nmax, d.islands, d.elite = P, stride, nmax, islands, elite
        d.migrate_every, d.iters, d.split = migrate_every, iters, split
        d.pStart, d.pEnd, d.lrStart, d.lrEnd, d.sigma = p_start, p_end, lr_start, lr_end, sigma
        d.activeTol, d.seed = active_tol, seed
        self.desc = d
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.handle = None
        self.eager_parity = 0
        if graph:
            h = C.c_void_p()
            _check(lib().sidon_generation_create(C.byref(d), C.byref(h), self.stream.cuda_stream),
                   "sidon_generation_create")
            self.handle = h

    def run(self, generations: int = 1) -> None:
        if self.handle is not None:
            _check(lib().sidon_generation_run(self.handle, generations, self.stream.cuda_stream),
                   "sidon_generation_run")
        else:
            for _ in range(generations):
                _check(lib().sidon_generation_enqueue(C.byref(self.desc), self.eager_parity,
                                                      self.stream.cuda_stream),
                       "sidon_generation_enqueue")
                self.eager_parity ^= 1

    def parity(self) -> int:
        if self.handle is None:
            return self.eager_parity
        p = C.c_int()
        _check(lib().sidon_generation_current(self.handle, C.byref(p)), "sidon_generation_current")
        return p.value

    def population(self):
        return (self.pop_b, self.ns_b) if self.parity() else (self.pop_a, self.ns_a)

    def close(self) -> None:
        if self.handle is not None:
            lib().sidon_generation_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------- reference-shaped device state --

class Gpu:
    """Device state of one ladder level: the SoA buffers of src/main.cu:107-144."""

    def __init__(self, n: int = 0, P: int = 0, device: str | torch.device = "cuda"):
        self.device = torch.device(device)
        self.n = self.P = 0
        if n and P:
            self.alloc(n, P)

    def alloc(self, n: int, P: int) -> None:
        _need_gpu()
        dev = self.device
        self.n, self.P = n, P
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.pop = torch.zeros(P, n, **f32)
        self.nxt = torch.zeros(P, n, **f32)
        self.mom = torch.zeros(P, n, **f32)
        self.vel = torch.zeros(P, n, **f32)
        self.fit = torch.zeros(P, dtype=torch.float64, device=dev)
        self.nact = torch.zeros(P, **i32)
        self.modeArr = torch.zeros(P, **i32)
        self.skip = torch.zeros(P, **i32)
        self.p1 = torch.zeros(P, **i32)
        self.p2 = torch.zeros(P, **i32)
        self.op = torch.zeros(P, **i32)
        self.strength = torch.zeros(P, **f32)

    def free(self) -> None:
        for k in ("pop", "nxt", "mom", "vel", "fit", "nact", "modeArr", "skip", "p1", "p2", "op",
                  "strength"):
            setattr(self, k, None)


def runLocalSearch(g: Gpu, iters: int, pA: float, pB: float, lrA: float, lrB: float,
                   modeArr: torch.Tensor | None = None, skip: torch.Tensor | None = None,
                   stream=None) -> None:
    """src/main.cu:152-159: beta = (0.9, 0.999), default mode 0, moments into g.mom/g.vel."""
    local_search(g.pop, iters, pA, pB, lrA, lrB, 0.9, 0.999, 0, mode_arr=modeArr, skip=skip,
                 mom=g.mom, vel=g.vel, stream=stream)


def runEval(g: Gpu, activeTol: float = 1e-4, stream=None) -> None:
    """src/main.cu:161-166: fills g.fit (fp64 F) and g.nact."""
    _need_gpu()
    _check(lib().sidon_eval_f64(_ptr(g.pop), g.n, g.P, _ptr(g.fit), _ptr(g.nact), activeTol,
                                _stream(stream)), "runEval")


def runFrankWolfe(g: Gpu, rounds: int, pdhgIters: int, nT: int, tMax: float,
                  skip: torch.Tensor | None = None, stream=None) -> bool:
    """src/main.cu:170-178: False when n does not fit the workgroup's LDS."""
    _need_gpu()
    rc = lib().sidon_frank_wolfe(_ptr(g.pop), _ptr(skip), g.n, g.P, rounds, pdhgIters, nT, tMax,
                                 _stream(stream))
    if rc == SIDON_ERR_LDS:
        return False
    _check(rc, "runFrankWolfe")
    return True


def runVariation(g: Gpu, seed: int, gen: int, stream=None) -> None:
    """src/main.cu:737-742: offspring of the (p1, op, strength) plan into g.nxt, then swap."""
    variation(g.pop, g.nxt, g.p1, g.p2, g.op, g.strength, seed, gen, stream=stream)
    g.pop, g.nxt = g.nxt, g.pop


from .lp import LpObjective, lp_objective  # noqa: E402  (autograd harness)
