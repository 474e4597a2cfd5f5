"""ctypes binding of the C ABI in include/gpum_hip.h.

The product path goes through this module only.  Loading fails loudly when
the in-tree library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from .build import LIB

_lib = None


class BackendError(RuntimeError):
    """A failed call into the HIP backend (message from gpum_hip_last_error)."""


class ChainResult(C.Structure):
    _fields_ = [("n_iter", C.c_uint32), ("n_chains", C.c_uint32),
                ("n_params", C.c_uint32), ("draws", C.POINTER(C.c_float)),
                ("accept_rate", C.POINTER(C.c_float)),
                ("last_log_post", C.POINTER(C.c_double))]


class RunExt(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("stream_id", C.c_uint32),
                ("trace_plp", C.POINTER(C.c_double)),
                ("trace_prop", C.POINTER(C.c_double)),
                ("trace_accept", C.POINTER(C.c_uint8)),
                ("final_state", C.POINTER(C.c_double)),
                ("init_log_post", C.POINTER(C.c_double)),
                ("chains_per_block", C.c_int32),
                ("waves_per_chain", C.c_int32),
                ("precompute_data", C.c_int32), ("fuse", C.c_int32),
                ("kernel_ms", C.c_double),
                ("iter_offset", C.c_uint32), ("swap_every", C.c_uint32),
                ("swap_group", C.c_uint32), ("adapt_batches", C.c_uint32),
                ("batch0", C.c_uint32),
                ("init_scale", C.POINTER(C.c_double)),
                ("swap_hist", C.POINTER(C.c_uint8)),
                ("acc_hist", C.POINTER(C.c_float)),
                ("final_sd", C.POINTER(C.c_double)),
                ("final_scale", C.POINTER(C.c_double)),
                ("trace_ll", C.POINTER(C.c_double)),
                ("chains_per_wave", C.c_int32)]


class DaExt(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("stream_id", C.c_uint32),
                ("iter_offset", C.c_uint32), ("work_items", C.c_uint32),
                ("work_groups", C.c_uint32),
                ("final_state", C.POINTER(C.c_double)),
                ("stage1_rate", C.POINTER(C.c_double)),
                ("trace_s", C.POINTER(C.c_double)),
                ("trace_plp", C.POINTER(C.c_double)),
                ("trace_stage", C.POINTER(C.c_uint8)),
                ("kernel_ms", C.c_double), ("exact_evals", C.c_double),
                ("trace_prop", C.POINTER(C.c_double)),
                ("trace_ll", C.POINTER(C.c_double)),
                ("chains_per_wave", C.c_uint32),
                ("gamma", C.c_double), ("de_noise", C.c_double)]


class BatchModel(C.Structure):
    _fields_ = [("model", C.c_void_p), ("n_chains", C.c_uint32),
                ("init", C.POINTER(C.c_double)),
                ("proposal_sd", C.POINTER(C.c_double)),
                ("seed", C.c_uint32), ("stream_id", C.c_uint32),
                ("gamma", C.c_double), ("de_noise", C.c_double),
                ("draws", C.POINTER(C.c_float)),
                ("accept_hist", C.POINTER(C.c_float)),
                ("final_sd", C.POINTER(C.c_double)),
                ("last_log_post", C.POINTER(C.c_double)),
                ("final_state", C.POINTER(C.c_double))]


class BatchOpts(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("proposal_mode", C.c_uint32),
                ("n_iter", C.c_uint32), ("batch_len", C.c_uint32),
                ("warmup_batches", C.c_uint32), ("adapt", C.c_int32),
                ("chains_per_block", C.c_int32),
                ("waves_per_chain", C.c_int32), ("use_graph", C.c_int32),
                ("kernel_ms", C.c_double), ("n_workgroups", C.c_uint32),
                ("chains_per_wave", C.c_int32)]


class WarmupOpts(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("warmup", C.c_uint32),
                ("seed", C.c_uint32), ("auto_stop", C.c_int32),
                ("use_graph", C.c_int32),
                ("final_init", C.POINTER(C.c_double)),
                ("final_sd", C.POINTER(C.c_double)),
                ("final_scale", C.POINTER(C.c_double)),
                ("final_l", C.POINTER(C.c_double)),
                ("accept_hist", C.POINTER(C.c_float)),
                ("n_batches_planned", C.c_uint32), ("n_batches", C.c_uint32),
                ("warmup_used", C.c_uint32), ("has_l", C.c_uint32),
                ("kernel_ms", C.c_double)]


class TrainOpts(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("steps", C.c_uint32),
                ("batch", C.c_uint32), ("seed", C.c_uint32),
                ("lr", C.c_double), ("lr_final", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double)]


class ModelInfo(C.Structure):
    _fields_ = [("n_vm_instr", C.c_uint32), ("vm_depth", C.c_uint32),
                ("ref_depth", C.c_uint32), ("n_uniform", C.c_uint32),
                ("n_cols_total", C.c_uint32), ("n_hoist_words", C.c_uint32),
                ("w_exec", C.c_double), ("w_bytecode", C.c_double),
                ("sum_mode", C.c_uint32), ("n_sum_tail", C.c_uint32)]


_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)


def lib():
    """The loaded backend library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise BackendError(
            f"HIP backend library not built: {LIB} is missing "
            "(run `python -m gpumetropolis_amd.build`)")
    L = C.CDLL(LIB)
    sz = C.c_size_t
    L.gpum_hip_available.restype = C.c_int
    L.gpum_hip_last_error.restype = C.c_char_p
    run_args = [_u32p, sz, _f64p, sz, _u32p, sz, _f64p, sz, C.c_uint32, _f64p,
                C.c_uint32, C.c_uint32, _f64p, C.c_uint32, _f64p, _f64p,
                C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                _f64p]
    L.gpum_hip_run_chains.argtypes = run_args + [C.POINTER(ChainResult)]
    L.gpum_hip_run_chains.restype = C.c_int
    L.gpum_hip_run_chains_ex.argtypes = run_args + [C.POINTER(RunExt),
                                                    C.POINTER(ChainResult)]
    L.gpum_hip_run_chains_ex.restype = C.c_int
    L.gpum_hip_result_free.argtypes = [C.POINTER(ChainResult)]
    L.gpum_hip_result_free.restype = None
    L.gpum_hip_loglik_batch.argtypes = [_u32p, sz, _f64p, sz, C.c_uint32,
                                        _f64p, C.c_uint32, C.c_uint32, _f64p,
                                        C.c_uint32, _f64p]
    L.gpum_hip_loglik_batch.restype = C.c_int
    for fn in (L.gpum_hip_loglik_pointwise, L.gpum_hip_grad_batch):
        fn.argtypes = [_u32p, sz, _f64p, sz, C.c_uint32, _f64p, C.c_uint32,
                       C.c_uint32, _f64p, C.c_uint32, _f64p]
        fn.restype = C.c_int
    L.gpum_hip_last_eval_ms.argtypes = []
    L.gpum_hip_last_eval_ms.restype = C.c_double
    L.gpum_hip_last_eval_compiled.argtypes = []
    L.gpum_hip_last_eval_compiled.restype = C.c_int
    L.gpum_hip_set_eval_jit.argtypes = [C.c_int32]
    L.gpum_hip_set_eval_jit.restype = None
    L.gpum_hip_jit_check.argtypes = [_u32p, sz, _f64p, sz, _u32p, sz, _f64p,
                                     sz, C.c_uint32, C.c_uint32, C.c_int32,
                                     C.c_uint32, C.POINTER(C.c_double)]
    L.gpum_hip_jit_check.restype = C.c_int
    L.gpum_hip_model_create.argtypes = [_u32p, sz, _f64p, sz, _u32p, sz,
                                        _f64p, sz, C.c_uint32, _f64p,
                                        C.c_uint32, C.c_uint32, C.c_int32,
                                        C.c_int32]
    L.gpum_hip_model_create.restype = C.c_void_p
    L.gpum_hip_model_info.argtypes = [C.c_void_p, C.POINTER(ModelInfo)]
    L.gpum_hip_model_info.restype = C.c_int
    L.gpum_hip_lower_info.argtypes = [_u32p, C.c_size_t, _f64p, C.c_size_t,
                                      C.c_uint32, C.c_uint32, C.c_int32,
                                      C.c_int32, C.POINTER(ModelInfo)]
    L.gpum_hip_lower_info.restype = C.c_int
    L.gpum_hip_model_run.argtypes = [C.c_void_p, _f64p, C.c_uint32, _f64p,
                                     _f64p, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_double, C.c_double,
                                     _f64p, C.POINTER(RunExt),
                                     C.POINTER(ChainResult)]
    L.gpum_hip_model_run.restype = C.c_int
    L.gpum_hip_model_loglik.argtypes = [C.c_void_p, _f64p, C.c_uint32, _f64p]
    L.gpum_hip_model_loglik.restype = C.c_int
    L.gpum_hip_model_free.argtypes = [C.c_void_p]
    L.gpum_hip_model_free.restype = None
    L.gpum_hip_model_set_jit.argtypes = [C.c_void_p, C.c_int32,
                                         C.POINTER(C.c_double)]
    L.gpum_hip_model_set_jit.restype = C.c_int
    _f32p = C.POINTER(C.c_float)
    L.gpum_hip_surrogate_n_weights.restype = C.c_uint32
    L.gpum_hip_surrogate_create.argtypes = [C.c_uint32, C.c_uint32]
    L.gpum_hip_surrogate_create.restype = C.c_void_p
    L.gpum_hip_surrogate_free.argtypes = [C.c_void_p]
    L.gpum_hip_surrogate_free.restype = None
    L.gpum_hip_surrogate_set_scaling.argtypes = [C.c_void_p, _f64p, _f64p,
                                                 C.c_double, C.c_double]
    L.gpum_hip_surrogate_set_scaling.restype = C.c_int
    L.gpum_hip_surrogate_get_weights.argtypes = [C.c_void_p, _f32p]
    L.gpum_hip_surrogate_get_weights.restype = C.c_int
    L.gpum_hip_surrogate_set_weights.argtypes = [C.c_void_p, _f32p]
    L.gpum_hip_surrogate_set_weights.restype = C.c_int
    L.gpum_hip_surrogate_gradient.argtypes = [C.c_void_p, _f64p, _f64p,
                                              C.c_uint32, _f32p, _f64p]
    L.gpum_hip_surrogate_gradient.restype = C.c_int
    L.gpum_hip_surrogate_train.argtypes = [C.c_void_p, _f64p, _f64p,
                                           C.c_uint32, C.POINTER(TrainOpts),
                                           _f64p, _f64p]
    L.gpum_hip_surrogate_train.restype = C.c_int
    L.gpum_hip_surrogate_predict.argtypes = [C.c_void_p, _f64p, C.c_uint32,
                                             _f64p]
    L.gpum_hip_surrogate_predict.restype = C.c_int
    L.gpum_hip_run_de_sync.argtypes = [_u32p, sz, _f64p, sz, _u32p, sz, _f64p,
                                       sz, C.c_uint32, _f64p, C.c_uint32,
                                       C.c_uint32, _f64p, C.c_uint32, _f64p,
                                       C.c_uint32, C.c_uint32, C.c_double,
                                       C.c_double, C.POINTER(ChainResult)]
    L.gpum_hip_run_de_sync.restype = C.c_int
    L.gpum_hip_model_run_da.argtypes = [C.c_void_p, C.c_void_p, _f64p,
                                        C.c_uint32, _f64p, _f64p, C.c_uint32,
                                        C.c_uint32, C.c_uint32, _f64p,
                                        C.POINTER(DaExt),
                                        C.POINTER(ChainResult)]
    L.gpum_hip_model_run_da.restype = C.c_int
    L.gpum_hip_model_warmup.argtypes = [C.c_void_p, _f64p, C.c_uint32, _f64p,
                                        C.POINTER(WarmupOpts)]
    L.gpum_hip_model_warmup.restype = C.c_int
    L.gpum_hip_batch_run.argtypes = [C.POINTER(BatchModel), C.c_uint32,
                                     C.POINTER(BatchOpts)]
    L.gpum_hip_batch_run.restype = C.c_int
    _lib = L
    return L


def last_error() -> str:
    return lib().gpum_hip_last_error().decode()


def available() -> bool:
    return bool(lib().gpum_hip_available())


def _check(rc: int):
    if rc != 0:
        raise BackendError(last_error())


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).ravel())


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _p(a: Optional[np.ndarray], t):
    if a is None or a.size == 0:
        return C.cast(None, C.POINTER(t))
    return a.ctypes.data_as(C.POINTER(t))


@dataclass
class RunOutput:
    """The reference's ChainResult (chain_kernel.rs:360-369) as numpy."""
    draws: Optional[np.ndarray]  # (n_iter, n_chains, n_params) float32
    # With raw_draws: the library's layout, (n_params, n_chains, n_iter),
    # and draws is None (the caller rearranges when it needs to).
    draws_pct: Optional[np.ndarray] = None
    accept_rate: np.ndarray = None    # (n_chains,) float32
    last_log_post: np.ndarray = None  # (n_chains,) float64
    final_state: Optional[np.ndarray] = None
    trace_plp: Optional[np.ndarray] = None
    trace_prop: Optional[np.ndarray] = None
    trace_accept: Optional[np.ndarray] = None
    init_log_post: Optional[np.ndarray] = None
    kernel_ms: float = 0.0
    swap_hist: Optional[np.ndarray] = None    # (n_batches, n_chains) uint8
    acc_hist: Optional[np.ndarray] = None     # (n_batches, n_chains)
    final_sd: Optional[np.ndarray] = None
    final_scale: Optional[np.ndarray] = None
    trace_ll: Optional[np.ndarray] = None     # (n_iter, n_chains) data term
    # stepwise / delayed-acceptance runs
    stage1_rate: Optional[np.ndarray] = None  # (n_chains,)
    exact_evals: float = 0.0
    trace_s: Optional[np.ndarray] = None      # (n_iter, n_chains)
    trace_stage: Optional[np.ndarray] = None  # (n_iter, n_chains) uint8


def _collect(res: ChainResult, raw: bool = False) -> RunOutput:
    T, Cc, P = res.n_iter, res.n_chains, res.n_params
    n = T * Cc * P
    flat = np.ctypeslib.as_array(res.draws, shape=(max(n, 1),))[:n].copy()
    acc = np.ctypeslib.as_array(res.accept_rate, shape=(Cc,)).copy()
    lp = np.ctypeslib.as_array(res.last_log_post, shape=(Cc,)).copy()
    lib().gpum_hip_result_free(C.byref(res))
    if raw:
        return RunOutput(draws=None, draws_pct=flat.reshape((P, Cc, T)),
                         accept_rate=acc, last_log_post=lp)
    draws = flat.reshape((P, Cc, T)).transpose(2, 1, 0).copy()
    return RunOutput(draws=draws, accept_rate=acc, last_log_post=lp)


def _ext(n_iter, n_chains, n_params, trace, final_state, stream_id, tiling,
         precompute_data, fuse, batches=None):
    ext = RunExt()
    ext.struct_size = C.sizeof(RunExt)
    ext.stream_id = stream_id
    keep = {}
    if batches:
        ext.iter_offset = batches.get("iter_offset", 0)
        ext.swap_every = batches.get("swap_every", 0)
        ext.swap_group = batches.get("swap_group", 0)
        ext.adapt_batches = batches.get("adapt_batches", 0)
        ext.batch0 = batches.get("batch0", 0)
        if batches.get("init_scale") is not None:
            keep["iscale"] = _f64(batches["init_scale"])
            ext.init_scale = _p(keep["iscale"], C.c_double)
        if ext.swap_every:
            nb = max(1, n_iter // ext.swap_every)
            keep["nb"] = nb
            keep["swaph"] = np.zeros(nb * n_chains, np.uint8)
            keep["acch"] = np.zeros(nb * n_chains, np.float32)
            ext.swap_hist = _p(keep["swaph"], C.c_uint8)
            ext.acc_hist = _p(keep["acch"], C.c_float)
        keep["fsd"] = np.zeros(n_chains * n_params, np.float64)
        keep["fscale"] = np.zeros(n_chains, np.float64)
        ext.final_sd = _p(keep["fsd"], C.c_double)
        ext.final_scale = _p(keep["fscale"], C.c_double)
    if trace:
        keep["plp"] = np.zeros(n_iter * n_chains, np.float64)
        keep["prop"] = np.zeros(n_iter * n_chains * n_params, np.float64)
        keep["acc"] = np.zeros(n_iter * n_chains, np.uint8)
        keep["ll"] = np.zeros(n_iter * n_chains, np.float64)
        ext.trace_ll = _p(keep["ll"], C.c_double)
        ext.trace_plp = _p(keep["plp"], C.c_double)
        ext.trace_prop = _p(keep["prop"], C.c_double)
        ext.trace_accept = _p(keep["acc"], C.c_uint8)
        keep["init"] = np.zeros(n_chains, np.float64)
        ext.init_log_post = _p(keep["init"], C.c_double)
    if final_state:
        keep["final"] = np.zeros(n_chains * n_params, np.float64)
        ext.final_state = _p(keep["final"], C.c_double)
    if tiling:
        # (chains per block, waves per chain[, chains per wave])
        ext.chains_per_block, ext.waves_per_chain = tiling[:2]
        if len(tiling) > 2:
            ext.chains_per_wave = tiling[2]
    ext.precompute_data = -1 if precompute_data is None else int(precompute_data)
    ext.fuse = -1 if fuse is None else int(fuse)
    return ext, keep


def _finish(out: RunOutput, ext, keep, n_iter, n_chains, n_params):
    out.kernel_ms = ext.kernel_ms
    if "plp" in keep:
        out.trace_plp = keep["plp"].reshape(n_iter, n_chains)
        out.trace_prop = keep["prop"].reshape(n_iter, n_chains, n_params)
        out.trace_accept = keep["acc"].reshape(n_iter, n_chains).astype(bool)
        out.trace_ll = keep["ll"].reshape(n_iter, n_chains)
        out.init_log_post = keep["init"]
    if "final" in keep:
        out.final_state = keep["final"].reshape(n_chains, n_params)
    if "swaph" in keep:
        out.swap_hist = keep["swaph"].reshape(keep["nb"], n_chains)
        out.acc_hist = keep["acch"].reshape(keep["nb"], n_chains)
    if "fsd" in keep:
        out.final_sd = keep["fsd"].reshape(n_chains, n_params)
        out.final_scale = keep["fscale"]
    return out


def run_chains(code, consts, prior_code, prior_consts, n_params, data, n_cols,
               n_obs, init, proposal_sd, temperatures, n_iter, seed,
               proposal_mode=0, gamma=0.0, de_noise=0.0, proposal_l=None,
               trace=False, final_state=False, stream_id=0, tiling=None,
               precompute_data=None, fuse=None, batches=None) -> RunOutput:
    """gpu_metropolis_chain (chain_kernel.rs:468-488) on the HIP backend.
    `batches` (dict: swap_every, swap_group, adapt_batches, iter_offset,
    batch0, init_scale) fuses the PT / adaptation batches into the launch."""
    L = lib()
    code, prior_code = _u32(code), _u32(prior_code)
    consts, prior_consts = _f64(consts), _f64(prior_consts)
    data, init = _f64(data), _f64(init)
    sd, temps = _f64(proposal_sd), _f64(temperatures)
    pl = None if proposal_l is None else _f64(proposal_l)
    n_chains = init.size // n_params
    ext, keep = _ext(n_iter, n_chains, n_params, trace, final_state, stream_id,
                     tiling, precompute_data, fuse, batches)
    res = ChainResult()
    rc = L.gpum_hip_run_chains_ex(
        _p(code, C.c_uint32), code.size, _p(consts, C.c_double), consts.size,
        _p(prior_code, C.c_uint32), prior_code.size,
        _p(prior_consts, C.c_double), prior_consts.size, n_params,
        _p(data, C.c_double), n_cols, n_obs, _p(init, C.c_double), n_chains,
        _p(sd, C.c_double), _p(temps, C.c_double), n_iter, seed & 0xFFFFFFFF,
        proposal_mode, gamma, de_noise, _p(pl, C.c_double), C.byref(ext),
        C.byref(res))
    _check(rc)
    return _finish(_collect(res), ext, keep, n_iter, n_chains, n_params)


def run_de_sync(code, consts, prior_code, prior_consts, n_params, data,
                n_cols, n_obs, init, proposal_sd, n_iter, seed, gamma,
                de_noise) -> RunOutput:
    """rust_gpu_metropolis_de_sync (lib.rs:215-260): synchronous
    Differential Evolution (path B) on the HIP backend."""
    L = lib()
    code, prior_code = _u32(code), _u32(prior_code)
    consts, prior_consts = _f64(consts), _f64(prior_consts)
    data, init, sd = _f64(data), _f64(init), _f64(proposal_sd)
    n_chains = init.size // n_params
    res = ChainResult()
    _check(L.gpum_hip_run_de_sync(
        _p(code, C.c_uint32), code.size, _p(consts, C.c_double), consts.size,
        _p(prior_code, C.c_uint32), prior_code.size,
        _p(prior_consts, C.c_double), prior_consts.size, n_params,
        _p(data, C.c_double), n_cols, n_obs, _p(init, C.c_double), n_chains,
        _p(sd, C.c_double), n_iter, seed & 0xFFFFFFFF, gamma, de_noise,
        C.byref(res)))
    return _collect(res)


def lower_info(code, consts, n_params, n_cols, precompute_data=None,
               fuse=None) -> dict:
    """What the kernels would run for a log-likelihood program (host-only
    lowering, no device)."""
    L = lib()
    code, consts = _u32(code), _f64(consts)
    inf = ModelInfo()
    _check(L.gpum_hip_lower_info(
        _p(code, C.c_uint32), code.size, _p(consts, C.c_double), consts.size,
        n_params, n_cols, -1 if precompute_data is None else int(precompute_data),
        -1 if fuse is None else int(fuse), C.byref(inf)))
    return {k: getattr(inf, k) for k, _ in ModelInfo._fields_}


def loglik_batch(code, consts, n_params, data, n_cols, n_obs, points):
    """rust_loglik_batch (lib.rs:131-151) on the HIP backend."""
    L = lib()
    code, consts, data, points = _u32(code), _f64(consts), _f64(data), _f64(points)
    n_points = points.size // n_params if n_params else 0
    out = np.zeros(max(n_points, 1), np.float64)
    rc = L.gpum_hip_loglik_batch(_p(code, C.c_uint32), code.size,
                                 _p(consts, C.c_double), consts.size, n_params,
                                 _p(data, C.c_double), n_cols, n_obs,
                                 _p(points, C.c_double), n_points,
                                 _p(out, C.c_double))
    _check(rc)
    return out[:n_points]


def loglik_pointwise(code, consts, n_params, data, n_cols, n_obs, points):
    """rust_loglik_pointwise (lib.rs:187-207) on the HIP backend: the FP64
    log-likelihood of each observation at each point, (n_points, n_obs)."""
    L = lib()
    code, consts, data, points = _u32(code), _f64(consts), _f64(data), _f64(points)
    n_points = points.size // n_params if n_params else 0
    out = np.zeros(max(n_points * n_obs, 1), np.float64)
    rc = L.gpum_hip_loglik_pointwise(
        _p(code, C.c_uint32), code.size, _p(consts, C.c_double), consts.size,
        n_params, _p(data, C.c_double), n_cols, n_obs,
        _p(points, C.c_double), n_points, _p(out, C.c_double))
    _check(rc)
    return out[:n_points * n_obs].reshape(n_points, n_obs)


def grad_batch(code, consts, n_params, data, n_cols, n_obs, points):
    """rust_grad_batch (lib.rs:158-179) on the HIP backend: the FP64 gradient
    of the summed data term at each point, (n_points, n_params)."""
    L = lib()
    code, consts, data, points = _u32(code), _f64(consts), _f64(data), _f64(points)
    n_points = points.size // n_params if n_params else 0
    out = np.zeros(max(n_points * n_params, 1), np.float64)
    rc = L.gpum_hip_grad_batch(
        _p(code, C.c_uint32), code.size, _p(consts, C.c_double), consts.size,
        n_params, _p(data, C.c_double), n_cols, n_obs,
        _p(points, C.c_double), n_points, _p(out, C.c_double))
    _check(rc)
    return out[:n_points * n_params].reshape(n_points, n_params)


JIT_KINDS = ("chain_stream", "chain_resident", "points_stream",
             "points_resident", "batch", "pointwise64", "grad64")


def jit_check(code, consts, prior_code, prior_consts, n_params, n_cols,
              chains_per_wave=1, kinds=JIT_KINDS) -> float:
    """Compile a model's per-model kernels with hiprtc without a device
    (gpum_hip_jit_check); returns the compile time in ms, raises on a
    compile error."""
    L = lib()
    code, prior_code = _u32(code), _u32(prior_code)
    consts, prior_consts = _f64(consts), _f64(prior_consts)
    mask = sum(1 << JIT_KINDS.index(k) for k in kinds)
    ms = C.c_double(0.0)
    _check(L.gpum_hip_jit_check(
        _p(code, C.c_uint32), code.size, _p(consts, C.c_double), consts.size,
        _p(prior_code, C.c_uint32), prior_code.size,
        _p(prior_consts, C.c_double), prior_consts.size, n_params, n_cols,
        chains_per_wave, mask, C.byref(ms)))
    return ms.value


def last_eval_ms() -> float:
    """Kernel time of this thread's last loglik_pointwise / grad_batch."""
    return float(lib().gpum_hip_last_eval_ms())


def last_eval_compiled() -> bool:
    """Whether that call ran kernels compiled for its program."""
    return bool(lib().gpum_hip_last_eval_compiled())


def set_eval_jit(enable: bool) -> None:
    """Compiled kernels for loglik_pointwise / grad_batch (default) or
    always the interpreter; process-wide."""
    lib().gpum_hip_set_eval_jit(1 if enable else 0)


class DeviceModel:
    """A model whose lowered program and data stay on the device."""

    def __init__(self, code, consts, prior_code, prior_consts, n_params, data,
                 n_cols, n_obs, precompute_data=None, fuse=None, jit=False):
        L = lib()
        self._L = L
        code, prior_code = _u32(code), _u32(prior_code)
        consts, prior_consts = _f64(consts), _f64(prior_consts)
        data = _f64(data)
        self.n_params = int(n_params)
        self.n_obs = int(n_obs)
        self.n_cols = int(n_cols)
        h = L.gpum_hip_model_create(
            _p(code, C.c_uint32), code.size, _p(consts, C.c_double),
            consts.size, _p(prior_code, C.c_uint32), prior_code.size,
            _p(prior_consts, C.c_double), prior_consts.size, n_params,
            _p(data, C.c_double), n_cols, n_obs,
            -1 if precompute_data is None else int(precompute_data),
            -1 if fuse is None else int(fuse))
        if not h:
            raise BackendError(last_error())
        self._h = h
        self.jit = False
        self.jit_compile_ms = 0.0
        self.jit_error = ""
        if jit == "auto":
            # Compiled kernels where run-time compilation works in this
            # process, else the interpreter (the reason is kept).
            try:
                self.set_jit(True)
            except BackendError as e:
                self.jit_error = str(e)
        elif jit:
            self.set_jit(True)

    def set_jit(self, enable: bool) -> float:
        """Compiled per-model kernels (True) or the interpreter (False);
        returns the compile time in ms (0 when cached)."""
        ms = C.c_double(0.0)
        _check(self._L.gpum_hip_model_set_jit(self._h, 1 if enable else 0,
                                              C.byref(ms)))
        self.jit = bool(enable)
        self.jit_compile_ms = ms.value
        return ms.value

    def info(self) -> dict:
        inf = ModelInfo()
        _check(self._L.gpum_hip_model_info(self._h, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in ModelInfo._fields_}

    def run(self, init, proposal_sd, temperatures, n_iter, seed,
            proposal_mode=0, gamma=0.0, de_noise=0.0, proposal_l=None,
            trace=False, final_state=False, stream_id=0,
            tiling=None, batches=None, raw_draws=False) -> RunOutput:
        np_ = self.n_params
        init, sd, temps = _f64(init), _f64(proposal_sd), _f64(temperatures)
        pl = None if proposal_l is None else _f64(proposal_l)
        n_chains = init.size // np_
        ext, keep = _ext(n_iter, n_chains, np_, trace, final_state, stream_id,
                         tiling, None, None, batches)
        res = ChainResult()
        rc = self._L.gpum_hip_model_run(
            self._h, _p(init, C.c_double), n_chains, _p(sd, C.c_double),
            _p(temps, C.c_double), n_iter, seed & 0xFFFFFFFF, proposal_mode,
            gamma, de_noise, _p(pl, C.c_double), C.byref(ext), C.byref(res))
        _check(rc)
        return _finish(_collect(res, raw_draws), ext, keep, n_iter, n_chains,
                       np_)

    def warmup(self, init, proposal_sd, warmup, seed, auto_stop=False,
               use_graph=True) -> Dict:
        """The adaptive warmup of R/adaptive.R:101-254 on the device
        (gpum_hip_model_warmup): returns the final states, proposal sd,
        Robbins-Monro scalars, proposal factors L (n_chains x np x np, or
        None), the acceptance history (n_chains x batches run), the batch
        plan and the device time."""
        np_ = self.n_params
        init, sd = _f64(init), _f64(proposal_sd)
        n_chains = init.size // np_
        nb_max = min(40, int(warmup))
        out_init = np.empty(n_chains * np_)
        out_sd = np.empty(n_chains * np_)
        out_scale = np.empty(n_chains)
        out_l = np.empty(n_chains * np_ * np_)
        hist = np.full(nb_max * n_chains, np.nan, np.float32)
        o = WarmupOpts()
        o.struct_size = C.sizeof(WarmupOpts)
        o.warmup = int(warmup)
        o.seed = int(seed) & 0xFFFFFFFF
        o.auto_stop = 1 if auto_stop else 0
        o.use_graph = 1 if use_graph else 0
        o.final_init = _p(out_init, C.c_double)
        o.final_sd = _p(out_sd, C.c_double)
        o.final_scale = _p(out_scale, C.c_double)
        o.final_l = _p(out_l, C.c_double)
        o.accept_hist = _p(hist, C.c_float)
        _check(self._L.gpum_hip_model_warmup(self._h, _p(init, C.c_double),
                                             n_chains, _p(sd, C.c_double),
                                             C.byref(o)))
        nb = int(o.n_batches)
        return {"init": out_init.reshape(n_chains, np_),
                "sd": out_sd.reshape(n_chains, np_),
                "scale": out_scale,
                "L": (out_l.reshape(n_chains, np_, np_) if o.has_l else None),
                "accept_hist": hist.reshape(nb_max, n_chains)[:nb].T.copy(),
                "n_batches_planned": int(o.n_batches_planned),
                "n_batches": nb, "warmup_used": int(o.warmup_used),
                "kernel_ms": float(o.kernel_ms)}

    def run_da(self, surrogate, init, temperatures, n_iter, seed,
               proposal_sd=None, proposal_l=None, trace=False,
               final_state=False, stream_id=0, iter_offset=0,
               work_items=0, work_groups=0, chains_per_wave=0,
               de_sync=None, langevin=False, raw_draws=False) -> RunOutput:
        """Stepwise random-walk Metropolis; with `surrogate` (a Surrogate)
        two-stage delayed acceptance, without it plain Metropolis with the
        uniforms of `run`.  Mode 2 when proposal_l is given, else mode 0;
        `de_sync=(gamma, de_noise)` runs synchronous Differential Evolution
        (mode 1) with proposal_sd as the jitter scale; `langevin` runs MALA
        (mode 3) preconditioned by proposal_l when given, else proposal_sd.
        `raw_draws`: the draws in the library's layout only (draws_pct,
        parameter x chain x iteration), as `run` gives them."""
        np_ = self.n_params
        init, temps = _f64(init), _f64(temperatures)
        n_chains = init.size // np_
        if langevin:
            mode = 3
        else:
            mode = 2 if proposal_l is not None else (1 if de_sync else 0)
        sd = None if proposal_sd is None else _f64(proposal_sd)
        pl = None if proposal_l is None else _f64(proposal_l)
        ext = DaExt()
        ext.struct_size = C.sizeof(DaExt)
        ext.stream_id = stream_id
        ext.iter_offset = iter_offset
        ext.work_items = work_items
        ext.work_groups = work_groups
        ext.chains_per_wave = chains_per_wave
        if de_sync:
            ext.gamma, ext.de_noise = float(de_sync[0]), float(de_sync[1])
        keep = {"rate": np.zeros(n_chains, np.float64)}
        ext.stage1_rate = _p(keep["rate"], C.c_double)
        if final_state:
            keep["final"] = np.zeros(n_chains * np_, np.float64)
            ext.final_state = _p(keep["final"], C.c_double)
        if trace:
            keep["s"] = np.zeros(n_iter * n_chains, np.float64)
            keep["plp"] = np.zeros(n_iter * n_chains, np.float64)
            keep["stage"] = np.zeros(n_iter * n_chains, np.uint8)
            ext.trace_s = _p(keep["s"], C.c_double)
            ext.trace_plp = _p(keep["plp"], C.c_double)
            ext.trace_stage = _p(keep["stage"], C.c_uint8)
            keep["prop"] = np.zeros(n_iter * n_chains * np_, np.float64)
            keep["ll"] = np.zeros(n_iter * n_chains, np.float64)
            ext.trace_prop = _p(keep["prop"], C.c_double)
            ext.trace_ll = _p(keep["ll"], C.c_double)
        res = ChainResult()
        sur = surrogate._h if surrogate is not None else None
        rc = self._L.gpum_hip_model_run_da(
            self._h, sur, _p(init, C.c_double), n_chains, _p(sd, C.c_double),
            _p(temps, C.c_double), n_iter, seed & 0xFFFFFFFF, mode,
            _p(pl, C.c_double), C.byref(ext), C.byref(res))
        _check(rc)
        out = _collect(res, raw=raw_draws)
        out.kernel_ms = ext.kernel_ms
        out.exact_evals = ext.exact_evals
        out.stage1_rate = keep["rate"]
        if final_state:
            out.final_state = keep["final"].reshape(n_chains, np_)
        if trace:
            out.trace_s = keep["s"].reshape(n_iter, n_chains)
            out.trace_plp = keep["plp"].reshape(n_iter, n_chains)
            out.trace_stage = keep["stage"].reshape(n_iter, n_chains)
            out.trace_prop = keep["prop"].reshape(n_iter, n_chains, np_)
            out.trace_ll = keep["ll"].reshape(n_iter, n_chains)
        return out

    def loglik(self, points) -> np.ndarray:
        points = _f64(points)
        n = points.size // self.n_params
        out = np.zeros(max(n, 1), np.float64)
        _check(self._L.gpum_hip_model_loglik(self._h, _p(points, C.c_double),
                                             n, _p(out, C.c_double)))
        return out[:n]

    def close(self):
        if getattr(self, "_h", None):
            self._L.gpum_hip_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """The device-resident surrogate network (include/gpum_hip.h)."""

    def __init__(self, n_in: int, seed: int = 0):
        L = lib()
        self._L = L
        self.n_in = int(n_in)
        h = L.gpum_hip_surrogate_create(self.n_in, seed & 0xFFFFFFFF)
        if not h:
            raise BackendError(last_error())
        self._h = h
        self.n_weights = int(L.gpum_hip_surrogate_n_weights())

    def set_scaling(self, mean, whiten, target_mean, target_sd):
        mean, whiten = _f64(mean), _f64(whiten)
        _check(self._L.gpum_hip_surrogate_set_scaling(
            self._h, _p(mean, C.c_double), _p(whiten, C.c_double),
            float(target_mean), float(target_sd)))

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.n_weights, np.float32)
        _check(self._L.gpum_hip_surrogate_get_weights(self._h, _p(w, C.c_float)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(np.asarray(w, np.float32).ravel())
        if w.size != self.n_weights:
            raise ValueError(f"expected {self.n_weights} weights, got {w.size}")
        _check(self._L.gpum_hip_surrogate_set_weights(self._h, _p(w, C.c_float)))

    def gradient(self, theta, loglik):
        theta, loglik = _f64(theta), _f64(loglik)
        n = loglik.size
        g = np.zeros(self.n_weights, np.float32)
        loss = C.c_double(0.0)
        _check(self._L.gpum_hip_surrogate_gradient(
            self._h, _p(theta, C.c_double), _p(loglik, C.c_double), n,
            _p(g, C.c_float), C.byref(loss)))
        return g, loss.value

    def train(self, theta, loglik, steps=2000, batch=1024, seed=0, lr=2e-3,
              lr_final=0.0):
        theta, loglik = _f64(theta), _f64(loglik)
        o = TrainOpts()
        o.struct_size = C.sizeof(TrainOpts)
        o.steps, o.batch, o.seed = steps, batch, seed & 0xFFFFFFFF
        o.lr, o.lr_final = lr, lr_final
        trace = np.zeros(max(steps, 1), np.float64)
        ms = C.c_double(0.0)
        _check(self._L.gpum_hip_surrogate_train(
            self._h, _p(theta, C.c_double), _p(loglik, C.c_double),
            loglik.size, C.byref(o), _p(trace, C.c_double), C.byref(ms)))
        return trace[:steps], ms.value

    def predict(self, theta) -> np.ndarray:
        theta = _f64(theta)
        n = theta.size // self.n_in
        out = np.zeros(max(n, 1), np.float64)
        _check(self._L.gpum_hip_surrogate_predict(
            self._h, _p(theta, C.c_double), n, _p(out, C.c_double)))
        return out[:n]

    def close(self):
        if getattr(self, "_h", None):
            self._L.gpum_hip_surrogate_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BatchOutput:
    draws: np.ndarray          # (n_iter, n_chains, n_params), every iteration
    accept_hist: np.ndarray    # (n_batches, n_chains)
    final_sd: np.ndarray
    last_log_post: np.ndarray
    final_state: np.ndarray


def batch_run(models, n_chains, inits, sds, seeds, stream_ids, gammas,
              de_noises, n_iter, batch_len, warmup_batches, adapt,
              proposal_mode=1, tiling=None, use_graph=True):
    """gpum_hip_batch_run over DeviceModels; returns (outputs, info)."""
    L = lib()
    k = len(models)
    arr = (BatchModel * k)()
    keep = []
    nb = max(1, n_iter // batch_len)
    outs = []
    for i, dm in enumerate(models):
        C_, np_ = int(n_chains[i]), dm.n_params
        init, sd = _f64(inits[i]), _f64(sds[i])
        if init.size != C_ * np_ or sd.size != C_ * np_:
            raise ValueError(f"model {i}: init and sd must be n_chains x n_params")
        o = BatchOutput(np.zeros(n_iter * C_ * np_, np.float32),
                        np.zeros(nb * C_, np.float32), np.zeros(C_ * np_),
                        np.zeros(C_), np.zeros(C_ * np_))
        keep += [init, sd]
        b = arr[i]
        b.model = dm._h
        b.n_chains = C_
        b.init, b.proposal_sd = _p(init, C.c_double), _p(sd, C.c_double)
        b.seed = int(seeds[i]) & 0xFFFFFFFF
        b.stream_id = int(stream_ids[i])
        b.gamma, b.de_noise = float(gammas[i]), float(de_noises[i])
        b.draws = _p(o.draws, C.c_float)
        b.accept_hist = _p(o.accept_hist, C.c_float)
        b.final_sd = _p(o.final_sd, C.c_double)
        b.last_log_post = _p(o.last_log_post, C.c_double)
        b.final_state = _p(o.final_state, C.c_double)
        outs.append(o)
    opts = BatchOpts()
    opts.struct_size = C.sizeof(BatchOpts)
    opts.proposal_mode = proposal_mode
    opts.n_iter = n_iter
    opts.batch_len = batch_len
    opts.warmup_batches = warmup_batches
    opts.adapt = 1 if adapt else 0
    if tiling:
        # (chains per block, waves per chain[, chains per wave])
        opts.chains_per_block, opts.waves_per_chain = tiling[:2]
        if len(tiling) > 2:
            opts.chains_per_wave = tiling[2]
    opts.use_graph = 1 if use_graph else 0
    _check(L.gpum_hip_batch_run(arr, k, C.byref(opts)))
    for i, o in enumerate(outs):
        C_, np_ = int(n_chains[i]), models[i].n_params
        o.draws = o.draws.reshape(np_, C_, n_iter).transpose(2, 1, 0).copy()
        o.accept_hist = o.accept_hist.reshape(nb, C_)
        o.final_sd = o.final_sd.reshape(C_, np_)
        o.final_state = o.final_state.reshape(C_, np_)
    return outs, {"kernel_ms": opts.kernel_ms,
                  "n_workgroups": opts.n_workgroups}
