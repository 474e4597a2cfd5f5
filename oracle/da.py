"""Replay of a two-stage delayed-acceptance run (test infrastructure only).

Imported by tests/ as the checker of the stepwise HIP sampler; the product
path never imports it.  The algorithm restated here is two-stage delayed
acceptance (Christen and Fox, "Markov chain Monte Carlo using an
approximation", J. Comput. Graph. Stat. 14(4), 2005) on top of the
reference's random-walk step (src/rust/src/cpu_native.rs:726-799):

  proposal    modes 0 and 2 with the Philox slot layout of the chain kernel
              (pairs (u1, u2) per parameter, acceptance uniform in slot 2 np);
  stage 1     log u1 < (s(prop) - s(theta)) / T with u1 = word 0 of Philox
              block (t, chain, 0, TAG_DA), s the surrogate log-posterior;
  stage 2     log u < ((lp(prop) - lp(theta)) - (s(prop) - s(theta))) / T
              with lp the exact log-posterior (FP64 oracle), u the
              acceptance slot.

The surrogate values are taken from the device run's own trace (the network
is checked separately), so every decision of the replay is fixed by FP64
arithmetic on the same uniforms; a decision may differ only where the
device's FP32-term evaluation error of lp straddles the threshold, which the
replay certifies before letting the two chains part.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import oracle

TAG_STEP, TAG_DA = 0, 3


@dataclass
class DaReplay:
    steps: int = 0
    stage1_checked: int = 0
    exact_checked: int = 0
    max_plp_rel: float = 0.0
    certified_flips: int = 0
    uncertified_flips: int = 0
    stage1_mismatch: int = 0


def _proposal(state, seed, stream, t, c, mode, sd, L):
    np_ = state.size
    z = np.empty(np_)
    for j in range(np_):
        u1 = oracle.uniform_philox(seed, stream, TAG_STEP, t, c, 2 * j)
        u2 = oracle.uniform_philox(seed, stream, TAG_STEP, t, c, 2 * j + 1)
        z[j] = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
    if mode == 2:
        return state + np.tril(L) @ z
    return state + sd * z


def replay_da(gpu, model, data, n_cols, n_obs, init, temperatures, seed,
              init_s=None, proposal_sd=None, proposal_l=None, stream=0,
              surrogate=True):
    """Walk every chain of a traced stepwise run (gpu: RunOutput with
    trace_s / trace_plp / trace_stage) through the restated algorithm.
    `model` has .loglik.code/.consts and .prior.code/.consts; init_s are the
    surrogate log-posteriors of the initial states."""
    rep = DaReplay()
    T, C = gpu.trace_stage.shape
    np_ = init.shape[1]
    mode = 2 if proposal_l is not None else 0

    def lp(theta):
        ll = oracle.loglik_batch(model.loglik.code, model.loglik.consts, np_,
                                 data, n_cols, n_obs, theta)[0]
        pr = (oracle.vm_eval(model.prior.code, model.prior.consts, theta, [0.0])
              if len(model.prior.code) else 0.0)
        return ll + pr

    for c in range(C):
        state = np.array(init[c], dtype=np.float64)
        cur = lp(state)
        cur_s = float(init_s[c]) if surrogate else 0.0
        Tc = float(temperatures[c])
        sd = None if proposal_sd is None else np.asarray(proposal_sd[c], float)
        L = None if proposal_l is None else np.asarray(proposal_l[c], float)
        for t in range(T):
            prop = _proposal(state, seed, stream, t, c, mode, sd, L)
            stage = int(gpu.trace_stage[t, c])
            rep.steps += 1
            if surrogate:
                sp = float(gpu.trace_s[t, c])
                u1 = oracle.uniform_philox(seed, stream, TAG_DA, t, c, 0)
                pass1 = math.isfinite(sp) and math.log(u1) < (sp - cur_s) / Tc
                rep.stage1_checked += 1
                if pass1 != (stage > 0):
                    # only a last-bit difference of log u1 may separate them
                    if abs(math.log(u1) - (sp - cur_s) / Tc) <= 1e-12:
                        rep.certified_flips += 1
                    else:
                        rep.stage1_mismatch += 1
                    break
                if not pass1:
                    continue
            else:
                sp = 0.0
            lo = lp(prop)
            lg = float(gpu.trace_plp[t, c])
            rep.exact_checked += 1
            if math.isfinite(lo):
                rep.max_plp_rel = max(rep.max_plp_rel,
                                      abs(lg - lo) / max(abs(lo), 1e-300))
            u = oracle.uniform_philox(seed, stream, TAG_STEP, t, c, 2 * np_)
            la_o = (lo - cur) - ((sp - cur_s) if surrogate else 0.0)
            acc_o = math.log(u) < la_o / Tc
            if acc_o != (stage == 2):
                la_g = (lg - cur) - ((sp - cur_s) if surrogate else 0.0)
                margin = abs(math.log(u) * Tc - la_o)
                if margin <= abs(la_g - la_o) * (1 + 1e-9) + 1e-12:
                    rep.certified_flips += 1
                else:
                    rep.uncertified_flips += 1
                break
            if acc_o:
                state, cur = prop, lo
                if surrogate:
                    cur_s = sp
    return rep
