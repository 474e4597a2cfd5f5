"""Parallel-tempering orchestration on the oracle (test infrastructure only).

Restates the reference's host-side PT loop (R/adaptive.R:272-413) on top of
the C oracle's batch sampler, with the stream and state conventions the HIP
backend fuses into one launch, so the two can be compared draw for draw:

  * one Philox stream for the whole run: batch b's iterations are addressed
    from the run's first iteration (iteration offset), not re-seeded per
    batch;
  * chains continue from their exact FP64 states between batches;
  * per-chain adaptation during the first `adapt_batches` batches: Welford
    on the recorded f32 draws (diagonal), Robbins-Monro on the batch's f32
    acceptance with batch index b (R/adaptive.R:342-356);
  * after every batch, each ladder of `group` consecutive chains attempts
    the alternating adjacent sweep (R/adaptive.R:362-391) with the uniform
    of pair position k of group g from Philox block (b, k, g, TAG_SWAP).
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle

TAG_SWAP = 1
_LOW_D = [0.44, 0.35, 0.32, 0.30, 0.28, 0.27, 0.26, 0.25]


def _target(d):
    return _LOW_D[d - 1] if 1 <= d <= 8 else 0.234


def run_pt(code, consts, prior_code, prior_consts, n_params, data, n_cols,
           n_obs, init, proposal_sd, temperatures, n_iter, seed, swap_every,
           group, adapt_batches=0, stream_id=0, trace=False):
    init = np.asarray(init, float).reshape(-1, n_params).copy()
    sd = np.asarray(proposal_sd, float).reshape(-1, n_params).copy()
    sd0 = sd.copy()
    temps = np.asarray(temperatures, float)
    C = init.shape[0]
    nb = max(1, n_iter // swap_every)
    sizes = [swap_every] * nb
    sizes[-1] += n_iter - nb * swap_every
    scale = np.ones(C)
    wn = np.zeros(C)
    wmean = np.zeros((C, n_params))
    wm2 = np.zeros((C, n_params))
    draws, tplp, tprop, tacc = [], [], [], []
    swap_hist = np.full((nb, C), 255, np.uint8)
    acc_hist = np.zeros((nb, C), np.float32)
    state = init
    t0 = 0
    lp = None
    for b in range(1, nb + 1):
        r = oracle.run(code, consts, prior_code, prior_consts, n_params, data,
                       n_cols, n_obs, state, sd, temps, sizes[b - 1], seed,
                       stream=oracle.STREAM_PHILOX, stream_id=stream_id,
                       trace=trace, iter_offset=t0)
        draws.append(r.draws)
        if trace:
            tplp.append(r.trace_plp)
            tprop.append(r.trace_prop)
            tacc.append(r.trace_accept)
        state = r.final_state.copy()
        lp = r.last_log_post.copy()
        acc = r.accept_rate.astype(np.float64)  # f32 rate, as over the FFI
        acc_hist[b - 1] = r.accept_rate
        if b <= adapt_batches:
            X = r.draws.astype(np.float64)
            for t in range(X.shape[0]):
                wn += 1.0
                d1 = X[t] - wmean
                wmean += d1 / wn[:, None]
                wm2 += d1 * (X[t] - wmean)
            g = float(b) ** (-2.0 / 3.0)
            scale = np.exp(np.log(scale) + g * (acc - _target(n_params)))
            var = np.where(wn[:, None] >= 2.0, wm2 / np.maximum(wn - 1.0, 1.0)[:, None],
                           sd0 * sd0)
            sd = scale[:, None] * np.sqrt(np.maximum(var, 1e-12))
        if group > 1:
            for g0 in range(0, C, group):
                pairs = group - 1
                start = 0 if b % 2 == 1 else (1 if pairs >= 2 else 0)
                for k, i in enumerate(range(start, pairs, 2)):
                    w = oracle.philox([b, k, g0 // group, TAG_SWAP],
                                      [seed, stream_id])
                    u = (float(w[0]) + 0.5) / 4294967296.0
                    lo, hi = g0 + i, g0 + i + 1
                    delta = (lp[hi] - lp[lo]) * (1.0 / temps[lo] - 1.0 / temps[hi])
                    ok = math.isfinite(delta) and math.log(u) < delta
                    if ok:
                        state[[lo, hi]] = state[[hi, lo]]
                        lp[[lo, hi]] = lp[[hi, lo]]
                    swap_hist[b - 1, lo] = 1 if ok else 0
        t0 += sizes[b - 1]
    out = {"draws": np.concatenate(draws), "final_state": state,
           "last_log_post": lp, "swap_hist": swap_hist, "acc_hist": acc_hist,
           "final_sd": sd, "final_scale": scale, "batch_sizes": sizes}
    if trace:
        out["trace_plp"] = np.concatenate(tplp)
        out["trace_prop"] = np.concatenate(tprop)
        out["trace_accept"] = np.concatenate(tacc)
    return out
