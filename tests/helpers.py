"""Shared checks for the parity tests (HIP backend vs the CPU oracle)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def slots_per_iter(mode: int, n_params: int) -> int:
    return 2 * n_params + 4 if mode == 1 else 2 * n_params + 1


@dataclass
class ReplayReport:
    steps_compared: int = 0
    max_plp_rel: float = 0.0
    max_prop_rel: float = 0.0
    certified_flips: int = 0
    uncertified_flips: int = 0
    chains_diverged: int = 0


def replay_compare(gpu, orc_res, orc_mod, seed, temperatures, mode, n_params,
                   init_lp_gpu=None, init_lp_orc=None, stream_id=0,
                   prop_rtol=1e-12):
    """Step-by-step comparison of two traced runs driven by the same Philox
    uniforms.  Per chain, while the two chains share their state: the
    proposals must agree to `prop_rtol`, the proposal log-posterior to 1e-6
    relative, and the accept decisions must be identical unless the FP64
    margin |log u - dlp/T| is within the GPU's own evaluation error of the
    difference (a certified ambiguous decision, after which the chains
    legitimately part)."""
    rep = ReplayReport()
    T, C = gpu.trace_plp.shape
    spi = slots_per_iter(mode, n_params)
    for c in range(C):
        cur_g = init_lp_gpu[c] if init_lp_gpu is not None else None
        cur_o = init_lp_orc[c] if init_lp_orc is not None else None
        for t in range(T):
            pg, po = gpu.trace_prop[t, c], orc_res.trace_prop[t, c]
            scale = np.maximum(np.abs(po), 1.0)
            prel = float(np.max(np.abs(pg - po) / scale))
            rep.max_prop_rel = max(rep.max_prop_rel, prel)
            assert prel <= prop_rtol, (
                f"chain {c} iter {t}: proposals differ {pg} vs {po}")
            lg, lo = gpu.trace_plp[t, c], orc_res.trace_plp[t, c]
            if math.isfinite(lo):
                rel = abs(lg - lo) / max(abs(lo), 1e-300)
                rep.max_plp_rel = max(rep.max_plp_rel, rel)
            else:
                assert not math.isfinite(lg) or not (lg > lo), (lg, lo)
            rep.steps_compared += 1
            ag, ao = bool(gpu.trace_accept[t, c]), bool(orc_res.trace_accept[t, c])
            if ag != ao:
                u = orc_mod.uniform_philox(seed, stream_id, 0, t, c, spi - 1)
                d64 = (lo - cur_o) / temperatures[c]
                dg = (lg - cur_g) / temperatures[c]
                margin = abs(math.log(u) - d64)
                if margin <= abs(dg - d64) * (1 + 1e-9) + 1e-12:
                    rep.certified_flips += 1
                else:
                    rep.uncertified_flips += 1
                rep.chains_diverged += 1
                break
            if ao:
                cur_o, cur_g = lo, lg
    return rep
