"""Python handle on the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.  The
product path (gpumetropolis_amd) never imports it.

The C functions restate the reference's f64 CPU path
(src/rust/src/cpu_native.rs); see oracle.c for the line-by-line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

STREAM_T32 = 0
STREAM_PHILOX = 1

LOGLIK_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double),
                        C.POINTER(C.c_double), C.c_uint32, C.c_uint32)


class RunArgs(C.Structure):
    _fields_ = [
        ("code", C.POINTER(C.c_uint32)), ("code_len", C.c_uint32),
        ("consts", C.POINTER(C.c_double)),
        ("prior_code", C.POINTER(C.c_uint32)), ("prior_len", C.c_uint32),
        ("prior_consts", C.POINTER(C.c_double)),
        ("n_params", C.c_uint32), ("data", C.POINTER(C.c_double)),
        ("n_cols", C.c_uint32), ("n_obs", C.c_uint32),
        ("init", C.POINTER(C.c_double)), ("n_chains", C.c_uint32),
        ("proposal_sd", C.POINTER(C.c_double)),
        ("temperatures", C.POINTER(C.c_double)),
        ("n_iter", C.c_uint32), ("seed", C.c_uint32),
        ("proposal_mode", C.c_uint32), ("gamma", C.c_double),
        ("de_noise", C.c_double), ("proposal_l", C.POINTER(C.c_double)),
        ("stream", C.c_uint32), ("stream_id", C.c_uint32),
        ("n_threads", C.c_uint32), ("loglik_fn", C.c_void_p),
        ("iter_offset", C.c_uint32),
    ]


class RunOut(C.Structure):
    _fields_ = [
        ("draws", C.POINTER(C.c_float)), ("accept_rate", C.POINTER(C.c_float)),
        ("last_log_post", C.POINTER(C.c_double)),
        ("final_state", C.POINTER(C.c_double)),
        ("trace_plp", C.POINTER(C.c_double)),
        ("trace_prop", C.POINTER(C.c_double)),
        ("trace_accept", C.POINTER(C.c_uint8)),
    ]


_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    u32p, f64p = C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    L.orc_triple32.argtypes = [C.c_uint32]
    L.orc_triple32.restype = C.c_uint32
    L.orc_uniform_t32.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_uniform_t32.restype = C.c_double
    L.orc_seedmix.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_seedmix.restype = C.c_uint32
    L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
    L.orc_philox4x32_10.restype = None
    L.orc_uniform_philox.argtypes = [C.c_uint32] * 6
    L.orc_uniform_philox.restype = C.c_double
    L.orc_lgamma.argtypes = [C.c_double]
    L.orc_lgamma.restype = C.c_double
    L.orc_digamma.argtypes = [C.c_double]
    L.orc_digamma.restype = C.c_double
    L.orc_vm_eval.argtypes = [u32p, C.c_uint32, f64p, f64p, f64p]
    L.orc_vm_eval.restype = C.c_double
    L.orc_loglik.argtypes = [u32p, C.c_uint32, f64p, f64p, f64p, C.c_uint32,
                             C.c_uint32]
    L.orc_loglik.restype = C.c_double
    for fn in ("orc_loglik_batch", "orc_loglik_pointwise", "orc_grad_batch"):
        f = getattr(L, fn)
        f.argtypes = [u32p, C.c_uint32, f64p, C.c_uint32, f64p, C.c_uint32,
                      C.c_uint32, f64p, C.c_uint32, f64p]
        f.restype = None
    L.orc_run.argtypes = [C.POINTER(RunArgs), C.POINTER(RunOut)]
    L.orc_run.restype = C.c_int
    L.orc_run_de_sync.argtypes = [C.POINTER(RunArgs), C.POINTER(RunOut)]
    L.orc_run_de_sync.restype = C.c_int
    _lib = L
    return L


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).ravel())


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _p(a, t):
    if a is None or a.size == 0:
        return C.cast(None, C.POINTER(t))
    return a.ctypes.data_as(C.POINTER(t))


def triple32(h: int) -> int:
    return lib().orc_triple32(h & 0xFFFFFFFF)


def philox(ctr, key):
    c, k = _u32(ctr), _u32(key)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32),
                            _p(out, C.c_uint32))
    return out


def uniform_philox(seed, stream, tag, t, chain, slot) -> float:
    return lib().orc_uniform_philox(seed, stream, tag, t, chain, slot)


def lgamma(x: float) -> float:
    return lib().orc_lgamma(float(x))


def digamma(x: float) -> float:
    return lib().orc_digamma(float(x))


def vm_eval(code, consts, params, row) -> float:
    code, consts, params, row = _u32(code), _f64(consts), _f64(params), _f64(row)
    return lib().orc_vm_eval(_p(code, C.c_uint32), code.size,
                             _p(consts, C.c_double), _p(params, C.c_double),
                             _p(row, C.c_double))


def loglik_batch(code, consts, n_params, data, n_cols, n_obs, points):
    code, consts, data, points = _u32(code), _f64(consts), _f64(data), _f64(points)
    n = points.size // n_params
    out = np.zeros(max(n, 1))
    lib().orc_loglik_batch(_p(code, C.c_uint32), code.size,
                           _p(consts, C.c_double), n_params,
                           _p(data, C.c_double), n_cols, n_obs,
                           _p(points, C.c_double), n, _p(out, C.c_double))
    return out[:n]


def loglik_pointwise(code, consts, n_params, data, n_cols, n_obs, points):
    code, consts, data, points = _u32(code), _f64(consts), _f64(data), _f64(points)
    n = points.size // n_params
    out = np.zeros(max(n * n_obs, 1))
    lib().orc_loglik_pointwise(_p(code, C.c_uint32), code.size,
                               _p(consts, C.c_double), n_params,
                               _p(data, C.c_double), n_cols, n_obs,
                               _p(points, C.c_double), n, _p(out, C.c_double))
    return out[:n * n_obs].reshape(n, n_obs)


def grad_batch(code, consts, n_params, data, n_cols, n_obs, points):
    code, consts, data, points = _u32(code), _f64(consts), _f64(data), _f64(points)
    n = points.size // n_params
    out = np.zeros(max(n * n_params, 1))
    lib().orc_grad_batch(_p(code, C.c_uint32), code.size,
                         _p(consts, C.c_double), n_params,
                         _p(data, C.c_double), n_cols, n_obs,
                         _p(points, C.c_double), n, _p(out, C.c_double))
    return out[:n * n_params].reshape(n, n_params)


@dataclass
class OracleResult:
    draws: np.ndarray          # (n_iter, n_chains, n_params) float32
    accept_rate: np.ndarray
    last_log_post: np.ndarray
    final_state: np.ndarray
    trace_plp: Optional[np.ndarray] = None
    trace_prop: Optional[np.ndarray] = None
    trace_accept: Optional[np.ndarray] = None


def run(code, consts, prior_code, prior_consts, n_params, data, n_cols, n_obs,
        init, proposal_sd, temperatures, n_iter, seed, proposal_mode=0,
        gamma=0.0, de_noise=0.0, proposal_l=None, stream=STREAM_PHILOX,
        stream_id=0, n_threads=None, trace=False, loglik_fn=None,
        de_sync=False, iter_offset=0) -> OracleResult:
    """cpu_native::run (cpu_native.rs:510-825) or run_de_sync (:400-506)."""
    L = lib()
    keep = dict(code=_u32(code), consts=_f64(consts), pc=_u32(prior_code),
                pk=_f64(prior_consts), data=_f64(data), init=_f64(init),
                sd=_f64(proposal_sd), temps=_f64(temperatures),
                pl=None if proposal_l is None else _f64(proposal_l))
    n_chains = keep["init"].size // n_params
    a = RunArgs()
    a.code = _p(keep["code"], C.c_uint32)
    a.code_len = keep["code"].size
    a.consts = _p(keep["consts"], C.c_double)
    a.prior_code = _p(keep["pc"], C.c_uint32)
    a.prior_len = keep["pc"].size
    a.prior_consts = _p(keep["pk"], C.c_double)
    a.n_params = n_params
    a.data = _p(keep["data"], C.c_double)
    a.n_cols = n_cols
    a.n_obs = n_obs
    a.init = _p(keep["init"], C.c_double)
    a.n_chains = n_chains
    a.proposal_sd = _p(keep["sd"], C.c_double)
    a.temperatures = _p(keep["temps"], C.c_double)
    a.n_iter = n_iter
    a.seed = seed & 0xFFFFFFFF
    a.proposal_mode = proposal_mode
    a.gamma = gamma
    a.de_noise = de_noise
    a.proposal_l = _p(keep["pl"], C.c_double)
    a.stream = stream
    a.stream_id = stream_id
    a.n_threads = n_threads or os.cpu_count() or 1
    a.loglik_fn = int(loglik_fn) if loglik_fn else None
    a.iter_offset = iter_offset
    draws = np.zeros(max(n_iter * n_chains * n_params, 1), np.float32)
    acc = np.zeros(n_chains, np.float32)
    lp = np.zeros(n_chains, np.float64)
    fin = np.zeros(n_chains * n_params, np.float64)
    o = RunOut()
    o.draws = _p(draws, C.c_float)
    o.accept_rate = _p(acc, C.c_float)
    o.last_log_post = _p(lp, C.c_double)
    o.final_state = _p(fin, C.c_double)
    if trace:
        tplp = np.zeros(n_iter * n_chains)
        tprop = np.zeros(n_iter * n_chains * n_params)
        tacc = np.zeros(n_iter * n_chains, np.uint8)
        o.trace_plp = _p(tplp, C.c_double)
        o.trace_prop = _p(tprop, C.c_double)
        o.trace_accept = _p(tacc, C.c_uint8)
    fn = L.orc_run_de_sync if de_sync else L.orc_run
    rc = fn(C.byref(a), C.byref(o))
    if rc != 0:
        raise ValueError(f"oracle run rejected its arguments (code {rc})")
    n = n_iter * n_chains * n_params
    d = draws[:n].reshape(n_params, n_chains, n_iter).transpose(2, 1, 0).copy()
    res = OracleResult(draws=d, accept_rate=acc, last_log_post=lp,
                       final_state=fin.reshape(n_chains, n_params))
    if trace:
        res.trace_plp = tplp.reshape(n_iter, n_chains)
        res.trace_prop = tprop.reshape(n_iter, n_chains, n_params)
        res.trace_accept = tacc.reshape(n_iter, n_chains).astype(bool)
    return res
