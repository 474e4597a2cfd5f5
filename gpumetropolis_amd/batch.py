"""Many independent models sampled together on one GPU.

``gpu_metropolis_batch`` runs the reference's per-model orchestration
(``gpu_metropolis`` with method "de", path A, R/api.R:442-527 and
R/adaptive.R:427-524; or "rwm" with the same between-batch rule) for a list
of models at once: one persistent launch runs every batch of every model,
each model's next batch starting as soon as its own previous batch and
between-batch step are done.  Models may
differ in program, data size, parameter count and chain count.  Each model
draws its initial values and uniforms from its own Philox stream (stream
id = its index), so a model's draws do not depend on the rest of the batch.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import adaptive as ad
from . import autotune
from ._native import DeviceModel, batch_run
from .api import Fit, _data_matrix, _sd_matrix
from .philox_host import normals


@dataclass
class BatchFit:
    fits: List[Fit]
    kernel_ms: float
    tiling: tuple
    n_workgroups: int
    tile_key: str
    tile_timings: Dict = field(default_factory=dict)
    batch_sizes: list = field(default_factory=list)


def gpu_metropolis_batch(models: Sequence, datas: Sequence, n_chains=64,
                         n_iter: int = 4000, warmup=None, method: str = "de",
                         proposal_sd=0.1, init=None, gamma=None,
                         de_noise: float = 1e-3, de_every=None,
                         adapt: bool = True, seed: int = 1, tiling="auto",
                         use_graph: bool = True,
                         tile_cache: Optional[str] = None,
                         jit="auto") -> BatchFit:
    if method not in ("de", "rwm"):
        raise ValueError('batched runs take method "de" or "rwm"')
    if len(models) != len(datas) or not models:
        raise ValueError("one data set per model, at least one model")
    K = len(models)
    chains = ([int(n_chains)] * K if np.isscalar(n_chains)
              else [int(c) for c in n_chains])
    if len(chains) != K:
        raise ValueError("`n_chains` must be a number or one per model")
    if method == "de" and min(chains) < 4:
        raise ValueError('`method = "de"` needs at least 4 chains per model.')
    n_iter = int(n_iter)
    warmup = n_iter // 2 if warmup is None else int(warmup)
    if warmup < 0 or warmup >= n_iter:
        raise ValueError("`warmup` must be in [0, n_iter).")
    if de_every is None:
        de_every = max(max(chains), 10)
    de_every = int(de_every)
    if de_every < 1:
        raise ValueError("`de_every` must be a positive integer.")
    sizes, nb = ad.split_batches(n_iter, de_every)
    wb = max(0, min(nb, warmup // de_every))
    warmup_used = sum(sizes[:wb])

    dms, inits, sds, gammas = [], [], [], []
    try:
        for k, (m, data) in enumerate(zip(models, datas)):
            X = _data_matrix(m, data)
            np_ = m.n_params
            dms.append(DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                                   m.prior.consts, np_, X, m.n_cols, X.shape[0],
                                   jit=jit))
            C = chains[k]
            if init is None:
                im = normals(seed, C * np_, stream=k).reshape(C, np_)
            else:
                im = np.asarray(init[k], dtype=np.float64).reshape(C, np_)
            inits.append(im)
            psd = proposal_sd[k] if isinstance(proposal_sd, (list, tuple)) else proposal_sd
            sds.append(_sd_matrix(psd, C, np_))
            g = 2.38 / math.sqrt(2 * np_) if gamma is None else float(gamma)
            if not math.isfinite(g) or g <= 0:
                raise ValueError("`gamma` must be positive and finite.")
            gammas.append(g)
        mode = 1 if method == "de" else 0
        infos = [dm.info() for dm in dms]
        compiled = all(d.jit for d in dms)
        key = autotune.tile_key([i["n_vm_instr"] for i in infos],
                                [d.n_obs for d in dms]) + ("-jit" if compiled else "")
        timings = {}
        if tiling == "auto":
            def time_candidate(t):
                _, inf = batch_run(dms, chains, inits, sds, [seed] * K,
                                   list(range(K)), gammas, [de_noise] * K,
                                   8, 8, 0, False, mode, tiling=t,
                                   use_graph=False)
                return inf["kernel_ms"]
            cands = autotune.CANDIDATES
            if compiled and all(i["vm_depth"] <= 1 for i in infos):
                cands = cands + autotune.PAIR_CANDIDATES
                # affine or square sums: every model the same form
                modes = {i["sum_mode"] for i in infos}
                if modes == {2} or modes == {1}:
                    cands = cands + autotune.QUAD_CANDIDATES
            tiling, timings = autotune.choose(key, time_candidate,
                                              candidates=cands,
                                              cache_path=tile_cache)
        outs, info = batch_run(dms, chains, inits, sds, [seed] * K,
                               list(range(K)), gammas, [de_noise] * K, n_iter,
                               de_every, wb, adapt, mode, tiling=tiling,
                               use_graph=use_graph)
    finally:
        for dm in dms:
            dm.close()
    fits = []
    w = np.asarray(sizes, dtype=np.float64)
    for k, (m, o) in enumerate(zip(models, outs)):
        d = o.draws[warmup_used:].astype(np.float64)
        samp = o.accept_hist[wb:]
        rate = ((w[wb:, None] * samp).sum(axis=0) / w[wb:].sum() if samp.size
                else o.accept_hist[-1])
        fits.append(Fit(d, rate.astype(np.float64), method, d.shape[0], n_iter,
                        warmup_used, chains[k], m.n_params, m.params, seed,
                        adaptation={"final_proposal_sd": o.final_sd,
                                    "n_batches": nb, "batch_sizes": sizes,
                                    "accept_history": o.accept_hist.T,
                                    "final_state": o.final_state,
                                    "last_log_post": o.last_log_post,
                                    "stream_id": k}))
    return BatchFit(fits, info["kernel_ms"], tuple(tiling) if tiling else (16, 1),
                    info["n_workgroups"], key, timings, sizes)
