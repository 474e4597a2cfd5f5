"""``gpu_metropolis`` on the HIP backend.

Host-side mirror of the reference's user entry point (R/api.R:247-599) and
its batch orchestrators (R/adaptive.R:101-524), driving the C ABI one batch
at a time exactly where the reference calls ``rust_gpu_metropolis``:

  method="rwm", adapt=True   .am_orchestrate_warmup, then one sampling batch
                             with the frozen per-chain Cholesky factor
                             (R/api.R:529-572, R/adaptive.R:101-254)
  method="rwm", adapt=False  one run, warmup rows trimmed (R/api.R:574-598)
  method="rwm", surrogate=..  the same warmup, recording every proposal and
                             its exact log-likelihood; a surrogate network
                             trained on those pairs then screens proposals
                             for two-stage delayed acceptance (no reference
                             counterpart), or plain sampling if its held-out
                             error exceeds the bound
  method="pt"                .pt_orchestrate (R/adaptive.R:272-413)
  method="de"                .de_orchestrate path A (R/adaptive.R:427-524)

The model's data are uploaded once (a device-resident model) instead of
once per batch.  Initial values and PT swap uniforms come from the Philox
stream keyed on ``seed`` (the reference draws them from R's generator), so
runs are reproducible from ``seed`` alone.  ``method="mala"`` and
``de_sync=True`` (synchronous Differential Evolution), which the reference
runs on its CPU path only (R/api.R:373-377, :464-472), run here on the
stepwise sampler: MALA with its log-posterior and gradient in FP64, path B
one launch sequence per generation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import adaptive as ad
from ._native import DeviceModel
from .dsl import Model
from .philox_host import normals, uniforms


@dataclass
class Fit:
    draws: np.ndarray  # (n_iter, n_chains, n_params) float64 of the f32 draws
    accept_rate: np.ndarray
    method: str
    n_iter: int
    n_iter_total: int
    warmup: int
    n_chains: int
    n_params: int
    params: list
    seed: int
    backend: str = "hip"
    temperatures: Optional[np.ndarray] = None
    adaptation: Dict = field(default_factory=dict)
    kernel_ms: float = 0.0


def _data_matrix(model: Model, data) -> np.ndarray:
    if model.n_cols == 0:
        return np.zeros((0, 0))
    if isinstance(data, dict):
        missing = [c for c in model.data if c not in data]
        if missing:
            raise ValueError("missing data columns: " + ", ".join(missing))
        cols = [np.asarray(data[c], dtype=np.float64).ravel() for c in model.data]
        return np.stack(cols, axis=1)
    a = np.asarray(data, dtype=np.float64)
    return a.reshape(-1, model.n_cols)


def _sd_matrix(proposal_sd, n_chains, np_) -> np.ndarray:
    """R/api.R:326-346: scalar, per-parameter vector or chain matrix."""
    sd = np.asarray(proposal_sd, dtype=np.float64)
    if sd.ndim == 2:
        if sd.shape != (n_chains, np_):
            raise ValueError(f"`proposal_sd` matrix must be `n_chains` by "
                             f"`n_params` ({n_chains} by {np_}).")
        m = sd.copy()
    else:
        v = np.resize(sd.ravel(), np_)
        m = np.tile(v, (n_chains, 1))
    if not np.all(np.isfinite(m)) or np.any(m <= 0):
        raise ValueError("`proposal_sd` must be positive and finite.")
    return m


def gpu_metropolis(model: Model, data=None, init=None, proposal_sd=0.1,
                   n_iter: int = 2000, n_chains: int = 4, warmup=None,
                   adapt: bool = True, method: str = "rwm",
                   temperatures=None, swap_every=None, gamma=None,
                   de_noise: float = 1e-3, de_every=None, de_sync=False,
                   seed: int = 1, backend: str = "hip",
                   tiling=None, ladder_size=None, surrogate=None,
                   engine: str = "auto", jit="auto",
                   device_warmup: bool = True) -> Fit:
    if backend not in ("hip", "auto"):
        raise ValueError(f"backend '{backend}' is not available here; this "
                         "package is the hip backend")
    if method not in ("rwm", "pt", "de", "mala"):
        raise ValueError(f"unknown method '{method}'")
    X = _data_matrix(model, data)
    n_obs = X.shape[0]
    np_ = model.n_params
    if init is None:
        n_chains = int(n_chains)
        init_mat = normals(seed, n_chains * np_).reshape(n_chains, np_)
    else:
        init_mat = np.asarray(init, dtype=np.float64)
        if init_mat.ndim == 1:
            init_mat = init_mat.reshape(-1, np_)
        if init_mat.shape[1] != np_:
            raise ValueError(f"`init` must have one column per parameter ({np_}).")
        n_chains = init_mat.shape[0]
    n_iter = int(n_iter)
    warmup_auto = warmup == "auto"
    if warmup is None or warmup_auto:
        warmup = n_iter // 2
    warmup = int(warmup)
    if warmup < 0 or warmup >= n_iter:
        raise ValueError("`warmup` must be a non-negative integer strictly "
                         "less than `n_iter`.")
    sd_mat = _sd_matrix(proposal_sd, n_chains, np_)

    dm = DeviceModel(model.loglik.code, model.loglik.consts, model.prior.code,
                     model.prior.consts, np_, X, model.n_cols, n_obs, jit=jit)
    if engine not in ("auto", "fused", "stepwise"):
        raise ValueError(f"unknown engine '{engine}'")
    if engine == "auto":
        engine = pick_engine(dm.info(), n_chains, n_obs) if method == "rwm" else "fused"
    if engine == "stepwise" and method not in ("rwm", "mala"):
        raise ValueError('`engine = "stepwise"` runs methods "rwm" and "mala".')
    call = (_StepCaller(dm, n_chains) if engine == "stepwise"
            else _Caller(dm, n_chains, tiling))

    try:
        if method == "pt":
            fit = _pt(call, model, init_mat, sd_mat, temperatures, n_iter,
                      warmup, swap_every, adapt, seed, n_chains,
                      ladder_size=ladder_size, dm=dm, tiling=tiling)
        elif method == "de":
            if n_chains < 4:
                raise ValueError('`method = "de"` needs at least 4 chains.')
            if de_sync:
                fit = _de_sync(dm, model, init_mat, sd_mat, gamma, de_noise,
                               n_iter, warmup, seed, n_chains)
            else:
                fit = _de(call, model, init_mat, sd_mat, gamma, de_noise,
                          n_iter, warmup, de_every, adapt, seed, n_chains)
        elif surrogate and method == "mala":
            raise ValueError("`surrogate` runs with the random-walk proposal "
                             '(method = "rwm").')
        elif adapt and warmup > 0:
            fit = _rwm_adaptive(call, model, init_mat, sd_mat, n_iter, warmup,
                                seed, warmup_auto, n_chains,
                                surrogate=surrogate, dm=dm,
                                langevin=method == "mala",
                                on_device=device_warmup)
        elif surrogate:
            raise ValueError("`surrogate` needs the adaptive warmup "
                             "(adapt = TRUE and warmup > 0) for its training "
                             "pairs.")
        else:
            r = call(init_mat, sd_mat, n_iter, seed,
                     mode=3 if method == "mala" else 0)
            d = r.draws[warmup:].astype(np.float64)
            fit = Fit(d, r.accept_rate.astype(np.float64), method, d.shape[0],
                      n_iter, warmup, n_chains, np_, model.params, seed)
    finally:
        dm.close()
    fit.kernel_ms = call.kernel_ms
    fit.adaptation["engine"] = engine
    fit.adaptation["compiled"] = dm.jit
    if dm.jit_error:
        fit.adaptation["compile_error"] = dm.jit_error
    return fit


# Geometry of the fused chain kernel (csrc/chain_kernel.h): one static LDS
# tile buffer of 4608 float4s; 16 waves per workgroup; 256 CUs.
_TILE_BYTES = 4608 * 16
_FULL_WAVES = 16 * 256


def pick_engine(info: dict, n_chains: int, n_obs: int) -> str:
    """Fused chain kernel or stepwise sampler for a random-walk run.

    The fused kernel keeps every chain's whole iteration loop in one
    workgroup, so it fills the GPU only with enough chains (times the waves
    each chain's rows allow in one LDS tile).  With fewer, and enough work
    per iteration to pay for a handful of launches, the stepwise sampler
    splits the observations over all CUs instead."""
    need = max(0, int(info["vm_depth"]) - 1)
    opt = 16 if need <= 3 else (8 if need <= 7 else 4)
    row_bytes = 64 * opt * max(1, int(info["n_cols_total"])) * 4
    wpc = 1
    while wpc < 16 and 2 * wpc * row_bytes <= _TILE_BYTES:
        wpc *= 2
    if n_chains * wpc < _FULL_WAVES and n_chains * n_obs >= (1 << 27):
        return "stepwise"
    return "fused"


class _Caller:
    """One batch on the device-resident model: the reference's `rust_call`
    closure (R/api.R:379-393)."""

    def __init__(self, dm, n_chains, tiling):
        self.dm, self.n_chains, self.tiling = dm, n_chains, tiling
        self.kernel_ms = 0.0

    def __call__(self, init_m, sd_m, iters, s, temps=None, mode=0, g=0.0,
                 noise=0.0, L=None, iter_offset=0, trace=False):
        t = np.ones(self.n_chains) if temps is None else temps
        batches = {"iter_offset": iter_offset} if iter_offset else None
        r = self.dm.run(init_m, sd_m, t, iters, int(s) & 0xFFFFFFFF,
                        proposal_mode=mode, gamma=g, de_noise=noise,
                        proposal_l=L, final_state=True, tiling=self.tiling,
                        batches=batches, trace=trace)
        self.kernel_ms += r.kernel_ms
        return r


class _StepCaller:
    """The `_Caller` interface on the stepwise sampler (proposal modes 0, 2
    and 3): one launch sequence per iteration, observations split over every
    CU."""

    def __init__(self, dm, n_chains):
        self.dm, self.n_chains = dm, n_chains
        self.kernel_ms = 0.0

    def __call__(self, init_m, sd_m, iters, s, temps=None, mode=0, g=0.0,
                 noise=0.0, L=None, iter_offset=0, trace=False):
        if mode not in (0, 2, 3):
            raise ValueError("the stepwise engine runs proposal modes 0, 2 "
                             "and 3")
        t = np.ones(self.n_chains) if temps is None else temps
        r = self.dm.run_da(None, init_m, t, iters, int(s) & 0xFFFFFFFF,
                           proposal_sd=sd_m if mode != 2 else None,
                           proposal_l=L if mode != 0 else None,
                           final_state=True, iter_offset=iter_offset,
                           trace=trace, langevin=mode == 3)
        self.kernel_ms += r.kernel_ms
        return r


@dataclass
class WarmupResult:
    init: np.ndarray        # (C, d) chain states to start sampling from
    sd: np.ndarray          # (C, d) per-chain diagonal proposal scale
    L: Optional[np.ndarray]  # (C * d * d,) per-chain Cholesky factor or None
    scales: np.ndarray
    used: int
    batches_run: int
    batch_sizes: list
    accept_history: np.ndarray
    pairs_theta: Optional[np.ndarray] = None   # (n, d) proposals
    pairs_loglik: Optional[np.ndarray] = None  # (n,) their data terms
    kernel_ms: float = 0.0                     # device time (device warmup)


def adaptive_warmup(call, n_params, n_chains, init_mat, sd_mat, sizes, seed,
                    auto_stop=False, collect=False,
                    langevin=False) -> WarmupResult:
    """.am_orchestrate_warmup (R/adaptive.R:101-254): per-chain and pooled
    Welford, Robbins-Monro scale per chain, mid-warmup reset, per-chain
    L = scale * chol(pooled covariance), optional early freeze.  With
    `collect`, the batches after the mid-warmup reset also return every
    proposal with its exact log-likelihood (surrogate training pairs).
    `langevin`: MALA batches (mode 3, the factor as preconditioner), scale
    1.65 / d^(1/6) and target acceptance 0.574 (R/adaptive.R:128, 147)."""
    np_ = n_params
    nb = len(sizes)
    states = ad.WelfordState.init(n_chains, np_, sd_mat ** 2)
    scale0 = 1.65 / np_ ** (1.0 / 6.0) if langevin else 2.38 / math.sqrt(np_)
    scales = np.full(n_chains, scale0)
    cur_init = init_mat.copy()
    cur_sd = sd_mat.copy()
    acc_hist = np.full((n_chains, nb), np.nan)
    cur_L = None
    L_pool = None
    pool = ad.PooledWelford.init(np_, sd_mat[0] ** 2)
    target = 0.574 if langevin else ad.target_accept(np_)
    reset_batch = nb // 2
    prev_scales = scales.copy()
    streak = 0
    batches_run = 0
    used = 0
    pth, pll = [], []
    for b in range(1, nb + 1):
        use_L = np_ > 1 and cur_L is not None
        rec = collect and b > reset_batch
        pm = 3 if langevin else (2 if use_L else 0)
        r = call(cur_init, cur_sd, sizes[b - 1], seed + b - 1,
                 mode=pm, L=cur_L if use_L else None, trace=rec)
        if rec:
            pth.append(r.trace_prop.reshape(-1, np_))
            pll.append(r.trace_ll.reshape(-1))
        X = r.draws.astype(np.float64)  # the f32 draws, as R receives them
        states.update_batch(X)
        pool.update_batch(X)
        bi = b - reset_batch if b > reset_batch else b
        scales = ad.robbins_monro_scale(scales, r.accept_rate.astype(np.float64),
                                        bi, np_, target)
        cur_init = X[-1].copy()
        cov = states.covariance()
        diag_sd = np.sqrt(np.maximum(np.diagonal(cov, axis1=1, axis2=2), 1e-12))
        cur_sd = scales[:, None] * diag_sd
        acc_hist[:, b - 1] = r.accept_rate
        if b == reset_batch:
            diag_now = np.maximum(np.diag(pool.covariance()), 1e-12)
            pool = ad.PooledWelford.init(np_, diag_now)
            states = ad.WelfordState.init(n_chains, np_, diag_now)
            scales = np.full(n_chains, scale0)
        if np_ > 1:
            if pool.n >= 2:
                L_pool = ad.regularised_cholesky(pool.covariance())
            if L_pool is not None:
                cur_L = (scales[:, None, None] * L_pool[None]).reshape(-1)
        batches_run = b
        used += sizes[b - 1]
        if auto_stop and b >= max(10, reset_batch + 5):
            drift = float(np.max(np.abs(scales / prev_scales - 1)))
            acc = r.accept_rate.astype(np.float64)  # R compares in double
            if np.mean(np.abs(acc - target)) < 0.05 and drift < 0.03:
                streak += 1
            else:
                streak = 0
            if streak >= 2:
                break
        prev_scales = scales.copy()
    return WarmupResult(cur_init, cur_sd, cur_L if np_ > 1 else None, scales,
                        used, batches_run, sizes[:batches_run],
                        acc_hist[:, :batches_run],
                        np.concatenate(pth) if pth else None,
                        np.concatenate(pll) if pll else None)


def device_warmup(dm, n_params, n_chains, init_mat, sd_mat, warmup, seed,
                  auto_stop=False) -> WarmupResult:
    """.am_orchestrate_warmup (R/adaptive.R:101-254) for random-walk chains
    with every batch and between-batch step on the device, one HIP graph
    (gpum_hip_model_warmup); the result as adaptive_warmup returns it."""
    r = dm.warmup(init_mat, sd_mat, warmup, seed, auto_stop=auto_stop)
    nb = r["n_batches"]
    L = None if n_params == 1 or r["L"] is None else r["L"].reshape(-1)
    return WarmupResult(r["init"], r["sd"], L, r["scale"], r["warmup_used"],
                        nb, ad.warmup_batches(warmup)[:nb], r["accept_hist"],
                        kernel_ms=r["kernel_ms"])


def _rwm_adaptive(call, model, init_mat, sd_mat, n_iter, warmup, seed,
                  auto_stop, n_chains, surrogate=None, dm=None,
                  langevin=False, on_device=True) -> Fit:
    """R/adaptive.R:101-254 then R/api.R:537-571; with `surrogate` the
    sampling batch is two-stage delayed acceptance; with `langevin` every
    batch is MALA (mode 3) with the warmup's factor as preconditioner.  The
    random-walk warmup of the fused engine runs on the device
    (device_warmup); the others orchestrate batch by batch."""
    np_ = model.n_params
    if (on_device and dm is not None and isinstance(call, _Caller)
            and not surrogate and not langevin):
        w = device_warmup(dm, np_, n_chains, init_mat, sd_mat, warmup, seed,
                          auto_stop)
    else:
        w = adaptive_warmup(call, np_, n_chains, init_mat, sd_mat,
                            ad.warmup_batches(warmup), seed, auto_stop,
                            collect=bool(surrogate), langevin=langevin)
    sampling = max(1, n_iter - w.used)
    use_L = w.L is not None
    adaptation = {"final_proposal_sd": w.sd, "final_scales": w.scales,
                  "n_batches": w.batches_run, "batch_sizes": w.batch_sizes,
                  "accept_history": w.accept_history, "final_L": w.L}
    net = None
    if surrogate:
        from .surrogate import fit_surrogate, tune_scale
        opts = dict(surrogate) if isinstance(surrogate, dict) else {}
        scale = opts.pop("scale", "auto")
        sf = fit_surrogate(w.pairs_theta, w.pairs_loglik, seed=seed, **opts)
        info = sf.summary()
        info["used"] = sf.usable
        if sf.usable:
            net = sf.net
        adaptation["surrogate"] = info
    if net is not None:
        s_seed = int(seed + w.batches_run) & 0xFFFFFFFF
        L_da = w.L if use_L else None
        sd_da = None if use_L else w.sd
        if scale == "auto":
            factor, pilots = tune_scale(dm, net, w.init, sf.whiten, s_seed,
                                        proposal_l=L_da, proposal_sd=sd_da)
            adaptation["surrogate"]["pilots"] = pilots
        else:
            factor = float(scale)
        adaptation["surrogate"]["scale_factor"] = factor
        r = dm.run_da(net, w.init, np.ones(n_chains), sampling, s_seed,
                      proposal_sd=None if sd_da is None else sd_da * factor,
                      proposal_l=None if L_da is None else L_da * factor,
                      final_state=True)
        call.kernel_ms += r.kernel_ms
        adaptation["surrogate"].update(
            stage1_rate=float(np.mean(r.stage1_rate)),
            exact_evals=float(r.exact_evals))
        net.close()
    else:
        r = call(w.init, w.sd, sampling, seed + w.batches_run,
                 mode=3 if langevin else (2 if use_L else 0),
                 L=w.L if use_L else None)
    d = r.draws.astype(np.float64)
    return Fit(d, r.accept_rate.astype(np.float64),
               "mala" if langevin else "rwm", d.shape[0],
               w.used + d.shape[0], w.used, n_chains, np_, model.params, seed,
               adaptation=adaptation)


def _pt(call, model, init_mat, sd_mat, temperatures, n_iter, warmup,
        swap_every, adapt, seed, n_chains, ladder_size=None, dm=None,
        tiling=None) -> Fit:
    """R/api.R:395-440 and R/adaptive.R:272-413.

    With ladders of at most 16 chains the whole orchestration -- batches,
    per-chain adaptation during the warmup batches, the alternating swap
    sweeps -- runs inside one kernel launch.  Longer ladders fall back to one
    launch per batch with the sweep on the host; both draw the same Philox
    stream (one stream for the run, iteration offsets per batch)."""
    np_ = model.n_params
    G = n_chains if ladder_size is None else int(ladder_size)
    if G < 1 or n_chains % G != 0:
        raise ValueError("`ladder_size` must divide the number of chains.")
    if temperatures is None:
        temps = np.tile(ad.default_ladder(G), n_chains // G)
    else:
        temps = np.asarray(temperatures, dtype=np.float64).ravel()
        if temps.size == G and G != n_chains:
            temps = np.tile(temps, n_chains // G)
        if temps.size != n_chains:
            raise ValueError("`temperatures` must have one value per chain.")
        if not np.all(np.isfinite(temps)) or np.any(temps <= 0):
            raise ValueError("`temperatures` must be positive and finite.")
    if swap_every is None:
        swap_every = max(G, 10)
    swap_every = int(swap_every)
    if swap_every < 1:
        raise ValueError("`swap_every` must be a positive integer.")
    sizes, nb = ad.split_batches(n_iter, swap_every)
    wb = max(0, min(nb, warmup // swap_every))
    warmup_used = sum(sizes[:wb])
    if G <= 16 and dm is not None:
        r = dm.run(init_mat, sd_mat, temps, n_iter, int(seed) & 0xFFFFFFFF,
                   final_state=True, tiling=tiling,
                   batches={"swap_every": swap_every, "swap_group": G,
                            "adapt_batches": wb if adapt else 0})
        call.kernel_ms += r.kernel_ms
        draws = r.draws.astype(np.float64)
        acc_hist = r.acc_hist.astype(np.float64)
        swap_hist = r.swap_hist
        cur_sd, scales = r.final_sd, r.final_scale
    else:
        draws, acc_hist, swap_hist, cur_sd, scales = _pt_batched(
            call, np_, n_chains, init_mat, sd_mat, temps, sizes, wb, adapt,
            seed, G)
    kept = draws[warmup_used:]
    w = np.array(sizes[wb:], dtype=np.float64)
    if w.sum() > 0:
        rate = (acc_hist[wb:] * w[:, None]).sum(axis=0) / w.sum()
    else:
        rate = acc_hist[-1]
    pairs = [i for i in range(n_chains) if (i + 1) % G != 0]
    sh = swap_hist[:, pairs].T.astype(np.float64)
    sh[sh == 255] = np.nan
    return Fit(kept, rate, "pt", kept.shape[0], warmup_used + kept.shape[0],
               warmup_used, n_chains, np_, model.params, seed,
               temperatures=temps,
               adaptation={"final_proposal_sd": cur_sd, "final_scales": scales,
                           "n_batches": nb, "batch_sizes": sizes,
                           "accept_history": acc_hist.T,
                           "swap_history": sh})


def _pt_batched(call, np_, n_chains, init_mat, sd_mat, temps, sizes, wb,
                adapt, seed, G):
    """One launch per batch; adaptation and the swap sweep on the host."""
    nb = len(sizes)
    states = ad.WelfordState.init(n_chains, np_, sd_mat ** 2)
    scales = np.ones(n_chains)
    cur = init_mat.copy()
    cur_sd = sd_mat.copy()
    draws = []
    acc_hist = np.zeros((nb, n_chains))
    swap_hist = np.full((nb, n_chains), 255, np.uint8)
    t0 = 0
    for b in range(1, nb + 1):
        r = call(cur, cur_sd, sizes[b - 1], seed, temps=temps, iter_offset=t0)
        X = r.draws.astype(np.float64)
        draws.append(X)
        acc = r.accept_rate.astype(np.float64)
        acc_hist[b - 1] = acc
        if b <= wb and adapt:
            states.update_batch(X)
            scales = ad.robbins_monro_scale(scales, acc, b, np_)
            cov = states.covariance()
            cur_sd = scales[:, None] * np.sqrt(
                np.maximum(np.diagonal(cov, axis1=1, axis2=2), 1e-12))
        cur = r.final_state.copy()
        lp = r.last_log_post.copy()
        for g0 in range(0, n_chains, G):
            pairs = G - 1
            if pairs < 1:
                continue
            start = 0 if b % 2 == 1 else (1 if pairs >= 2 else 0)
            order = list(range(start, pairs, 2))
            u = uniforms(seed, b, len(order), group=g0 // G)
            for k, i in enumerate(order):
                lo, hi = g0 + i, g0 + i + 1
                delta = (lp[hi] - lp[lo]) * (1 / temps[lo] - 1 / temps[hi])
                ok = math.isfinite(delta) and math.log(u[k]) < delta
                if ok:
                    cur[[lo, hi]] = cur[[hi, lo]]
                    lp[[lo, hi]] = lp[[hi, lo]]
                swap_hist[b - 1, lo] = 1 if ok else 0
        t0 += sizes[b - 1]
    return np.concatenate(draws), acc_hist, swap_hist, cur_sd, scales


def _de_scales(gamma, de_noise, np_):
    """ter Braak's default scale and the checks of R/api.R:448-460."""
    if gamma is None:
        gamma = 2.38 / math.sqrt(2 * np_)
    gamma = float(gamma)
    if not math.isfinite(gamma) or gamma <= 0:
        raise ValueError("`gamma` must be positive and finite.")
    de_noise = float(de_noise)
    if not math.isfinite(de_noise) or de_noise < 0:
        raise ValueError("`de_noise` must be non-negative and finite.")
    return gamma, de_noise


def _de_sync(dm, model, init_mat, sd_mat, gamma, de_noise, n_iter, warmup,
             seed, n_chains) -> Fit:
    """R/api.R:464-498 (path B): one call runs every generation on the
    device (gpum_hip_run_de_sync semantics on the resident model); the
    warmup rows are trimmed, with no adaptation, as the reference does."""
    np_ = model.n_params
    gamma, de_noise = _de_scales(gamma, de_noise, np_)
    r = dm.run_da(None, init_mat, np.ones(n_chains), n_iter, seed,
                  proposal_sd=sd_mat, de_sync=(gamma, de_noise))
    draws = r.draws.astype(np.float64)[warmup:]
    return Fit(draws=draws, accept_rate=r.accept_rate, method="de",
               n_iter=n_iter - warmup, n_iter_total=n_iter, warmup=warmup,
               n_chains=n_chains, n_params=np_, params=list(model.params),
               seed=seed, kernel_ms=r.kernel_ms,
               adaptation={"de_sync": True, "gamma": gamma,
                           "de_noise": de_noise,
                           "compiled": bool(dm.jit)})


def _de(call, model, init_mat, sd_mat, gamma, de_noise, n_iter, warmup,
        de_every, adapt, seed, n_chains) -> Fit:
    """R/api.R:442-527 and R/adaptive.R:427-524 (path A)."""
    np_ = model.n_params
    gamma, de_noise = _de_scales(gamma, de_noise, np_)
    if de_every is None:
        de_every = max(n_chains, 10)
    sizes, nb = ad.split_batches(n_iter, int(de_every))
    wb = max(0, min(nb, warmup // int(de_every)))
    warmup_used = sum(sizes[:wb])
    states = ad.WelfordState.init(n_chains, np_, sd_mat ** 2)
    cur_init = init_mat.copy()
    cur_sd = sd_mat.copy()
    kept = []
    acc_acc = np.zeros(n_chains)
    iters_acc = 0
    acc_hist = np.full((n_chains, nb), np.nan)
    disp = np.full((np_, nb), np.nan)
    r = None
    for b in range(1, nb + 1):
        bsize = sizes[b - 1]
        r = call(cur_init, cur_sd, bsize, seed + b - 1, mode=1, g=gamma,
                 noise=de_noise)
        X = r.draws.astype(np.float64)
        acc_hist[:, b - 1] = r.accept_rate
        in_warmup = b <= wb
        if not in_warmup:
            kept.append(X)
            acc_acc += bsize * r.accept_rate
            iters_acc += bsize
        if in_warmup and adapt:
            states.update_batch(X)
            cov = states.covariance()
            cur_sd = np.sqrt(np.maximum(np.diagonal(cov, axis1=1, axis2=2), 1e-12))
        cur_init = X[-1].copy()
        disp[:, b - 1] = cur_init.std(axis=0, ddof=1)
    d = np.concatenate(kept) if kept else np.zeros((0, n_chains, np_))
    rate = acc_acc / iters_acc if iters_acc > 0 else r.accept_rate
    return Fit(d, rate, "de", d.shape[0], warmup_used + d.shape[0], warmup_used,
               n_chains, np_, model.params, seed,
               adaptation={"final_proposal_sd": cur_sd, "n_batches": nb,
                           "batch_sizes": sizes, "accept_history": acc_hist,
                           "disp_history": disp})
