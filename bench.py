"""Headline benchmark: plain HIP random-walk Metropolis on the gamma model.

Workload: y = 1 000 000 FP32 draws ~ Gamma(shape 2.5, rate 0.7); params
(log_a, log_b) with N(0, 10) priors; 4096 chains x 5000 iterations.  The
data are synthetic and seeded; there are no model weights.

One step = one batch of `--iters-per-step` Metropolis iterations of all 4096
chains over the full data set on the GPU (one launch of the fused chain
kernel).  The W warmup steps are W x iters-per-step warmup iterations, run
untimed as the reference's adaptive warmup with its own batch plan (10 to
40 batches of about 50, R/adaptive.R:112-118; Welford + Robbins-Monro +
pooled Cholesky, :101-254); the K timed steps continue the chains from
their exact FP64 states with the frozen per-chain Cholesky proposal.
Defaults: W = K = 10, 250 iterations per step: 2500 warmup + 2500 sampling
iterations = the 5000-iteration workload.

value  = observation-loglik evaluations per second over the timed region
         (chains x observations x iterations / wall time, inputs resident in
         HBM, draws copied back to the host inside the region).
Also reported: ESS/s (min over parameters of the summed per-chain Geyer ESS
of the sampling draws / timed wall) with the max split R-hat and the
sampling acceptance, the posterior moments against the exact posterior
(quadrature) and its Laplace approximation, the roofline of the chain
kernel, and the oracle timed on the host's cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

VALU_PEAK = 256 * 64 * 2.4e9  # FP32 lane-ops/s, non-packed (one slot per op)

# HBM bytes per launch of the headline kernel, from two rocprofv3 --pmc
# passes over this script's default launches (FETCH_SIZE scaled x2.0 and
# WRITE_SIZE x1.0 by a 1 GiB copy); the summary and commands are in
# profiles/compiled_chain_kernel_gamma_1e6obs_4096chains.md (compiled) and
# profiles/chain_kernel_gamma_1e6obs_4096chains.md (interpreter, --no-jit).
# Reported only for the launch shape it was measured on.
PROFILED_TRAFFIC = {
    ("compiled", 4096, 1_000_000, 250): {"fetch_bytes": 8_133_000 * 1024 * 2.0,
                                         "write_bytes": 66_160 * 1024 * 1.0},
    ("interpreter", 4096, 1_000_000, 250): {
        "fetch_bytes": 20_390_000 * 1024 * 2.0,
        "write_bytes": 819_300 * 1024 * 1.0},
}
# Instruction mix of the parallel-tempering launch (4000 iterations, 4096
# chains, 2e5 rows; SQ_INSTS_VALU and SQ_INSTS_VALU_TRANS_F32 over 64
# evaluations), from the rocprofv3 pass in
# profiles/pt_chain_kernel_mixture_4096chains.md.
PT_PROFILED_MIX = {"valu_insts_per_64_evals": 17.63,
                   "trans_insts_per_64_evals": 2.01}
HBM_PEAK = 8.0e12             # B/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--iters-per-step", type=int, default=250)
    p.add_argument("--chains", type=int, default=4096)
    p.add_argument("--n-obs", type=int, default=1_000_000)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="target duration of the CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-jit", dest="jit", action="store_false",
                   help="run every model on the interpreter instead of its "
                        "compiled kernels")
    p.add_argument("--vm-launches", type=int, default=2,
                   help="interpreter launches timed beside the compiled "
                        "headline (0: skip)")
    p.add_argument("--no-da", action="store_true",
                   help="skip the surrogate delayed-acceptance comparison")
    p.add_argument("--da-obs", type=int, default=4_000_000)
    p.add_argument("--da-chains", type=int, default=1024)
    p.add_argument("--da-warmup", type=int, default=600)
    p.add_argument("--da-iters", type=int, default=1000,
                   help="iterations per chunk of the R-hat-gated runs")
    p.add_argument("--da-max-iters", type=int, default=16000)
    p.add_argument("--da-train-steps", type=int, default=4000)
    p.add_argument("--no-pt", action="store_true",
                   help="skip the parallel-tempering workload")
    p.add_argument("--pt-iters", type=int, default=4000)
    p.add_argument("--pt-chunk", type=int, default=2000,
                   help="iterations per continuation launch of the PT run")
    p.add_argument("--pt-max-iters", type=int, default=16000)
    p.add_argument("--no-eval", action="store_true",
                   help="skip the FP64 pointwise / gradient, synchronous "
                        "DE and MALA sections")
    p.add_argument("--no-m1", action="store_true",
                   help="skip the reference's M1 benchmark cells")
    p.add_argument("--m1-reps", type=int, default=3,
                   help="replications per M1 cell (median reported)")
    p.add_argument("--no-gaussian", action="store_true",
                   help="skip the Gaussian-mean (8 chains) section")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the many-model (512 regressions) section")
    p.add_argument("--batch-models", type=int, default=512)
    p.add_argument("--batch-iters", type=int, default=10000,
                   help="DE-MC iterations per chain of the many-model run; "
                        "64 chains of 4000 (half warmup) leave ~55 ESS per "
                        "chain, which holds split R-hat near 1.02")
    p.add_argument("--batch-warmup", type=int, default=2000)
    p.add_argument("--only", choices=["pt", "batch", "da", "eval",
                                      "gaussian", "m1"],
                   help="run only this section and print its JSON (for "
                        "profiling one kernel family); no headline line")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only run a plain device-to-device copy of MIB MiB "
                        "(calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}), flush=True)


def _host_threads():
    """Threads for the CPU baseline: the cores this process may run on,
    capped by OMP_NUM_THREADS where the host sets it (a GPU host that
    allots each GPU's jobs a share of its cores says so there)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, target_s, start, sd, L, seed):
    """The oracle (C restatement of the reference's f64 CPU sampler with the
    per-model compiled log-likelihood mirroring its JIT) on the host cores,
    on bounded samples of the same workload (one chain per thread, the full
    data set):
      evals/s  all threads' chains for a calibrated number of iterations;
      ESS/s    the same number of chains started from the GPU run's final
               (stationary) states with its adapted proposal, run for a
               calibrated length; per-chain Geyer ESS summed, as the GPU's.
    """
    from oracle import codegen, oracle
    from gpumetropolis_amd.diagnostics import summarize

    m = w.model
    P = m.n_params
    cores = _host_threads()
    _, _, addr = codegen.compile_loglik(m.loglik.code, m.loglik.consts)
    C = cores
    init = np.tile([math.log(2.5), math.log(0.7)], (C, 1))
    sd0 = np.full((C, 2), 0.001)

    def timed(iters, init_m, sd_m, mode=0, Lm=None, s=1):
        t = time.perf_counter()
        r = oracle.run(m.loglik.code, m.loglik.consts, m.prior.code,
                       m.prior.consts, P, w.data, 1, w.n_obs, init_m, sd_m,
                       np.ones(C), iters, s, proposal_mode=mode,
                       proposal_l=Lm, n_threads=cores, loglik_fn=addr)
        return time.perf_counter() - t, r

    # One iteration is two data passes (the initial state and the proposal).
    per_pass = timed(1, init, sd0)[0] / 2.0
    iters = max(1, int(target_s / max(per_pass, 1e-3)) - 1)
    dt, _ = timed(iters, init, sd0)
    per_pass = dt / (iters + 1)
    evals = C * w.n_obs * (iters + 1)  # + the initial evaluation
    out = {"value": evals / dt, "unit": "obs-loglik evals/s", "cores": cores,
           "kind": "port",
           "host_cores": os.cpu_count(), "cpu_model": _cpu_model(),
           "threads_note": "one thread per core this process may use "
                           "(affinity, capped by OMP_NUM_THREADS where the "
                           "host sets it for one GPU's share)",
           "sample": f"{C} chains x {iters} iterations x {w.n_obs} obs "
                     f"(compiled f64 loglik, {cores} threads, {dt:.1f} s)"}
    # ESS/s leg: stationary starts, the GPU run's adapted proposal.
    mode = 2 if L is not None else 0
    Lm = None if L is None else np.asarray(L).reshape(-1, P * P)[:C].ravel()
    ess_iters = max(200, int(target_s / max(per_pass, 1e-6)))
    dt2, r = timed(ess_iters, np.asarray(start)[:C], np.asarray(sd)[:C], mode,
                   Lm, seed)
    summ = summarize(r.draws.astype(np.float64))
    out.update({
        "ess_per_s": summ["min_ess"] / dt2,
        "ess_min": summ["min_ess"], "ess_rhat_max": summ["max_rhat"],
        "ess_accept": float(np.mean(r.accept_rate)),
        "ess_sample": f"{C} chains x {ess_iters} iterations from the GPU "
                      f"chains' final states with their adapted proposal "
                      f"(mode {mode}), {dt2:.1f} s; ESS/s is chain-count "
                      "invariant on a saturated CPU"})
    return out


def gaussian_mean_section(a):
    """Plain RWM on the Gaussian mean with known sd 2 and a flat prior:
    20,000 draws of N(3.4, 2), 8 chains x 3000 iterations with the
    reference's adaptive warmup (half), overdispersed starts
    mean(y) + 2 * seq(-2, 2) and proposal sd 2.4 * 2 / sqrt(n), as the
    reference's M1 benchmark (benchmark/adapters/cpu_adapters.R:25-61).
    ESS/s is the post-warmup ESS over the whole call's wall time, as there;
    the closed-form posterior is N(mean(y), 4 / n)."""
    from gpumetropolis_amd import workloads
    from gpumetropolis_amd.api import gpu_metropolis
    from gpumetropolis_amd.diagnostics import summarize

    w = workloads.gaussian_mean()
    C, T = 8, 3000
    init = w.data.mean() + 2.0 * np.linspace(-2, 2, C).reshape(C, 1)
    psd = 2.4 * 2.0 / math.sqrt(w.n_obs)
    kw = dict(init=init, proposal_sd=psd, n_iter=T, seed=a.seed, jit=a.jit)
    gpu_metropolis(w.model, w.data, **kw)        # first call: module loads
    t0 = time.perf_counter()
    fit = gpu_metropolis(w.model, w.data, **kw)
    wall = time.perf_counter() - t0
    s = summarize(fit.draws)
    p = s["params"][0]
    gate = s["max_rhat"] < 1.01
    return {"workload": f"Gaussian mean, {w.n_obs} obs, {C} chains x {T} "
                        f"iterations ({T - fit.draws.shape[0]} adaptive warmup)",
            "engine": fit.adaptation.get("engine"),
            "wall_s": wall, "ms_per_iteration": wall / T * 1e3,
            "obs_evals_per_s": C * w.n_obs * T / wall,
            "ess": s["min_ess"], "rhat_max": s["max_rhat"],
            "ess_per_s": s["min_ess"] / wall if gate else None,
            "closed_form_mean_err_mcse": abs(p["mean"] - w.meta["post_mean"]) / p["mcse"],
            "sd_ratio": p["sd"] / w.meta["post_sd"],
            "sampling_accept": float(np.mean(fit.accept_rate))}


# The reference's published M1 cells: gpumetropolis-cuda on an RTX 4090,
# median ESS/s over replications (benchmark/full_m1_cell_summary.csv:35-43).
M1_PUBLISHED = {(10000, 1): 3040.54, (10000, 8): 24143.41,
                (10000, 64): 101775.20, (10000, 512): 207630.28,
                (10000, 4096): 259145.06, (100000, 1): 53.18,
                (100000, 8): 398.84, (100000, 64): 2873.47,
                (100000, 512): 5785.03}


def m1_cells_section(a):
    """The reference's M1 benchmark protocol on every published cell
    (benchmark/run_cell.R, benchmark/adapters/cpu_adapters.R:25-61,
    benchmark/models/m1_gaussian_mean.R): y ~ N(0.37, 1.5^2) of N rows,
    log-likelihood -((y - mu)^2) / 4.5, C chains started at
    mean(y) + 1.5 * seq(-2, 2) (mean(y) for one chain), proposal sd
    2.4 * 1.5 / sqrt(N), 4000 iterations with the API's default warmup (the
    first half), one discarded call first.  ESS is the sum over chains of
    each chain's AR spectral ESS of the kept half (coda::effectiveSize, as
    benchmark/harness_util.R:9-12), over the wall time of the whole API
    call; the median over replications is compared with the published
    median of the same cell.  The data are drawn with numpy, not R."""
    from gpumetropolis_amd import dsl
    from gpumetropolis_amd.api import gpu_metropolis
    from gpumetropolis_amd.diagnostics import ess_ar, rhat

    sigma, mu0, T = 1.5, 0.37, 4000
    model = dsl.gpum_model(f"~ -((y - mu)^2) / {2 * sigma ** 2:.10g}",
                           ["mu"], ["y"])
    cells = []
    for k, ((N, C), pub) in enumerate(M1_PUBLISHED.items()):
        rates, rhats, walls, errs = [], [], [], []
        for rep in range(1, a.m1_reps + 1):
            y = np.random.default_rng(20260000 + 1000 * k + rep).normal(
                mu0, sigma, N).reshape(N, 1)
            init = (np.full((1, 1), y.mean()) if C == 1 else
                    (y.mean() + sigma * np.linspace(-2, 2, C)).reshape(C, 1))
            kw = dict(init=init, proposal_sd=2.4 * sigma / math.sqrt(N),
                      seed=10000 * (k + 1) + rep, jit=a.jit)
            if rep == 1:
                gpu_metropolis(model, y, n_iter=64, **kw)  # discarded
            t0 = time.perf_counter()
            fit = gpu_metropolis(model, y, n_iter=T, **kw)
            wall = time.perf_counter() - t0
            d = fit.draws[:, :, 0]
            total = sum(ess_ar(d[:, c]) for c in range(C))
            rates.append(total / wall)
            walls.append(wall)
            if C > 1:
                rhats.append(rhat(d))
            # closed form N(mean(y), sigma^2 / N), in Monte Carlo errors
            mcse = sigma / math.sqrt(N) / math.sqrt(max(total, 1.0))
            errs.append(abs(float(d.mean()) - float(y.mean())) / mcse)
        med = float(np.median(rates))
        cells.append({"N": N, "C": C, "ess_per_s_median": med,
                      "wall_s_median": float(np.median(walls)),
                      "rhat_max": max(rhats) if rhats else None,
                      "closed_form_max_err_mcse": max(errs),
                      "published_rtx4090": pub, "ratio": med / pub})
        print(f"[bench-m1] N={N} C={C}: {med:.0f} ESS/s "
              f"({med / pub:.2f}x the published cell)", file=sys.stderr,
              flush=True)
    return {"protocol": "reference M1 harness: 4000 iterations, half "
                        "warmup, sum of per-chain AR spectral ESS over the "
                        f"API call's wall time, median of {a.m1_reps} "
                        "replications",
            "cells": cells,
            "min_ratio": min(c["ratio"] for c in cells)}


MLP_FLOP_PER_ROW = 2 * (16 * 256 + 256 * 256)  # matrix-core flops, forward
MFMA_BF16_PEAK = 2.5e15


def da_section(a):
    """Plain random-walk Metropolis vs surrogate delayed acceptance on the
    16-dim logistic regression (4e6 FP32 rows, 1024 chains).  Both sample
    the same number of iterations from the same warmed-up states with the
    same proposal; plain MH runs on the faster of the two plain engines
    (the stepwise sampler splits the observations over every CU; the fused
    chain kernel's 1024 chains fill a quarter of them)."""
    from gpumetropolis_amd import _native, adaptive as ad, workloads
    from gpumetropolis_amd.api import _StepCaller, adaptive_warmup
    from gpumetropolis_amd.diagnostics import rhat, summarize
    from gpumetropolis_amd.surrogate import fit_surrogate, tune_scale

    t0 = time.perf_counter()
    w = workloads.logistic(n=a.da_obs, n_chains=a.da_chains)
    m = w.model
    d, C, N = m.n_params, a.da_chains, w.n_obs
    dm = _native.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                             m.prior.consts, d, w.data, w.n_cols, N, jit=a.jit)
    info = dm.info()
    print(f"[bench-da] data + model {time.perf_counter() - t0:.1f} s: {info}",
          file=sys.stderr, flush=True)
    rng = np.random.default_rng(a.seed)
    # chains start within a few posterior sds (~1.3e-3) of the truth
    init = w.meta["true"] + rng.normal(0.0, 0.003, (C, d))
    call = _StepCaller(dm, C)
    t0 = time.perf_counter()
    wu = adaptive_warmup(call, d, C, init, np.full((C, d), 5e-4),
                         ad.warmup_batches(a.da_warmup), a.seed, collect=True)
    warm_s = time.perf_counter() - t0
    print(f"[bench-da] warmup {a.da_warmup} it {warm_s:.1f} s, accept "
          f"{wu.accept_history[:, -1].mean():.3f}, pairs {wu.pairs_loglik.size}",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    fit = fit_surrogate(wu.pairs_theta, wu.pairs_loglik, seed=a.seed,
                        steps=a.da_train_steps)
    fit_s = time.perf_counter() - t0
    print(f"[bench-da] surrogate {fit_s:.1f} s: {fit.summary()}",
          file=sys.stderr, flush=True)
    seed2 = (a.seed + wu.batches_run) & 0xFFFFFFFF
    res = {}
    # The fused chain kernel on the same states, briefly, for the record.
    r = dm.run(wu.init, wu.sd, np.ones(C), 20, seed2, proposal_mode=2,
               proposal_l=wu.L)
    fused_ms = r.kernel_ms / 20
    factor, pilots = 1.0, []
    if fit.usable:
        t0 = time.perf_counter()
        factor, pilots = tune_scale(dm, fit.net, wu.init, fit.whiten, seed2,
                                    proposal_l=wu.L)
        pilot_s = time.perf_counter() - t0
        print(f"[bench-da] scale factor {factor} ({pilot_s:.1f} s): {pilots}",
              file=sys.stderr, flush=True)
    for name, net, f in (("plain", None, 1.0),
                         ("da", fit.net if fit.usable else None, factor)):
        if name == "da" and net is None:
            break
        # Chunks of da_iters iterations, continued from the exact states,
        # until every parameter's split R-hat is below 1.01 (or the cap).
        draws, state, wall, kms, exact, t = [], wu.init, 0.0, 0.0, 0.0, 0
        s1, acc = [], []
        while True:
            # run_da returns after its stream has drained
            t0 = time.perf_counter()
            r = dm.run_da(net, state, np.ones(C), a.da_iters, seed2,
                          proposal_l=wu.L * f, final_state=True, iter_offset=t,
                          raw_draws=True)
            wall += time.perf_counter() - t0
            kms += r.kernel_ms
            exact += r.exact_evals
            s1.append(float(np.mean(r.stage1_rate)))
            acc.append(float(np.mean(r.accept_rate)))
            # library layout in the timed call, rearranged outside it (as in
            # the headline's timed loop)
            draws.append(r.draws_pct.transpose(2, 1, 0))
            state = r.final_state
            t += a.da_iters
            x = np.concatenate(draws)
            rh = max(rhat(x[:, :, j]) for j in range(d))
            print(f"[bench-da] {name}: {t} it, {wall:.1f} s, R-hat {rh:.4f}",
                  file=sys.stderr, flush=True)
            if rh < 1.01 or t >= a.da_max_iters:
                break
        summ = summarize(x.astype(np.float64))
        moved = float(np.mean(np.any(np.diff(x, axis=0) != 0, axis=(0, 2))))
        res[name] = {
            "iters": t, "scale_factor": f,
            "wall_s": wall, "ms_per_iter": wall / t * 1e3,
            "kernel_ms_per_iter": kms / t,
            "ess_min": summ["min_ess"], "ess_per_s": summ["min_ess"] / wall,
            "rhat_max": summ["max_rhat"],
            "accept": float(np.mean(acc)),
            "stage1_rate": float(np.mean(s1)),
            "exact_evals": exact,
            "obs_evals_per_s": exact * N / wall,
            "chains_that_moved": moved,
            "params": summ["params"]}
        print(f"[bench-da] {name}: {res[name]['ms_per_iter']:.2f} ms/it, ESS "
              f"{summ['min_ess']:.0f}, R-hat {summ['max_rhat']:.4f}",
              file=sys.stderr, flush=True)
        del draws, x
    out = {"workload": f"logistic regression, {d} params, {N} FP32 rows x "
                       f"{w.n_cols} cols, {C} chains, {wu.used} warmup "
                       f"iterations, then chunks of {a.da_iters} until split "
                       f"R-hat < 1.01",
           "plain_engine": "stepwise",
           "plain_fused_ms_per_iter": fused_ms,
           "da_scale_pilots": pilots,
           "warmup_s": warm_s,
           "surrogate": dict(fit.summary(),
                             train_steps=a.da_train_steps,
                             mfma_tflops=(a.da_train_steps * 1024 * 3 *
                                          MLP_FLOP_PER_ROW /
                                          (fit.train_ms / 1e3) / 1e12),
                             mfma_frac=(a.da_train_steps * 1024 * 3 *
                                        MLP_FLOP_PER_ROW /
                                        (fit.train_ms / 1e3) / MFMA_BF16_PEAK)),
           "fit_s": fit_s}
    for name in res:
        params = res[name].pop("params")
        res[name]["posterior_mean"] = [p["mean"] for p in params]
        res[name]["posterior_mcse"] = [p["mcse"] for p in params]
    out.update(res)
    if "da" in res:
        p, q = res["plain"], res["da"]
        out["ess_per_s_ratio"] = q["ess_per_s"] / p["ess_per_s"]
        out["ess_per_s_ratio_incl_training"] = (
            q["ess_min"] / (q["wall_s"] + fit_s)) / p["ess_per_s"]
        z = [abs(x - y) / math.hypot(u, v) for x, y, u, v in zip(
            p["posterior_mean"], q["posterior_mean"], p["posterior_mcse"],
            q["posterior_mcse"])]
        out["max_mean_diff_in_mcse"] = max(z)
        out["rhat_ok"] = bool(p["rhat_max"] < 1.01 and q["rhat_max"] < 1.01)
    if fit.net is not None:
        fit.net.close()
    dm.close()
    return out


def pt_section(a):
    """Parallel tempering on the bimodal mixture: 200,000 FP32 rows, 16
    geometric temperatures (beta 1 .. 0.01) x 256 replicas, swaps every 10
    iterations; batches, per-chain adaptation and swap sweeps run inside one
    launch.  The configuration's 4000 iterations (half warmup, adapting) are
    one launch; the exact states then continue, with the adapted proposals
    and the same Philox stream and sweep parity, in chunks until the cold
    chains' split R-hat is below 1.01 (as gpu_metropolis's R/adaptive.R
    orchestration would by calling again).  ESS/s is over all the launches'
    kernel time."""
    from gpumetropolis_amd import _native, adaptive as ad, workloads
    from gpumetropolis_amd.diagnostics import summarize

    wl = workloads.mixture(n_iter=a.pt_iters)
    m = wl.model
    G, R = wl.meta["n_temps"], wl.meta["n_replicas"]
    C = G * R
    se = wl.meta["swap_every"]
    temps = np.tile(wl.meta["temperatures"], R)
    rng = np.random.default_rng(a.seed)
    init = rng.normal(0.0, 4.0, (C, 2))
    dm = _native.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                             m.prior.consts, 2, wl.data, 1, wl.n_obs,
                             jit=a.jit)
    warm = a.pt_iters // 2
    sizes, nb = ad.split_batches(a.pt_iters, se)
    wb = warm // se
    seed = a.seed & 0xFFFFFFFF
    r = dm.run(init, np.full((C, 2), 0.3), temps, a.pt_iters, seed,
               final_state=True,
               batches={"swap_every": se, "swap_group": G, "adapt_batches": wb})
    first_ms = r.kernel_ms
    kms = r.kernel_ms
    sh = [r.swap_hist]
    cold = [r.draws[warm:, ::G, :]]
    state, sd, t, b = r.final_state, r.final_sd, a.pt_iters, nb
    upper0 = float(np.mean(cold[0][:, :, 0] > cold[0][:, :, 1]))

    def label_free(x):
        # The likelihood is symmetric in the labels; the posterior's two
        # modes are (mu1, mu2) = (-3, 3) and (3, -3).  On the raw labels
        # R-hat measures how often replicas cross between the mirror
        # images, not convergence; (min, max) is the invariant summary.
        return np.sort(np.concatenate(x).astype(np.float64), axis=2)

    s = summarize(label_free(cold))
    while s["max_rhat"] >= 1.01 and t < a.pt_max_iters:
        n = a.pt_chunk
        r = dm.run(state, sd, temps, n, seed, final_state=True,
                   batches={"swap_every": se, "swap_group": G,
                            "adapt_batches": 0, "iter_offset": t,
                            "batch0": b})
        kms += r.kernel_ms
        sh.append(r.swap_hist)
        cold.append(r.draws[:, ::G, :])
        state, t, b = r.final_state, t + n, b + n // se
        s = summarize(label_free(cold))
        print(f"[bench-pt] {t} it, cold R-hat {s['max_rhat']:.4f}",
              file=sys.stderr, flush=True)
    w_exec = dm.info()["w_exec"]
    dm.close()
    raw = summarize(np.concatenate(cold).astype(np.float64))
    gate = s["max_rhat"] < 1.01
    swaps = np.concatenate(sh)
    evals = C * wl.n_obs * a.pt_iters
    rate = evals / (first_ms / 1e3)
    issue = PT_PROFILED_MIX["valu_insts_per_64_evals"]
    roof = {"bound": "valu", "kernel": "gpum_jit_kernel (fused PT, compiled "
                                       "mixture program)",
            "w_exec": w_exec, "achieved": w_exec * rate / 1e12,
            "peak": VALU_PEAK / 1e12, "unit": "T VALU-slot/s",
            "frac": w_exec * rate / VALU_PEAK,
            "avg_launch_ms": first_ms,
            # what the SIMDs issued: every VALU instruction (transcendentals
            # included) as one slot per lane
            "issued_slots_per_eval": issue,
            "issue_frac": issue * rate / VALU_PEAK,
            "trans_insts_per_64_evals":
                PT_PROFILED_MIX["trans_insts_per_64_evals"]}
    return {"workload": f"bimodal mixture, {wl.n_obs} FP32 rows, {G} temperatures "
                        f"x {R} replicas, {a.pt_iters} iterations ({warm} "
                        f"adaptive warmup), swaps every {se}; continued in "
                        f"chunks of {a.pt_chunk} until the cold R-hat < 1.01",
            "obs_evals_per_s": rate,
            "roofline": roof,
            "kernel_ms_config_launch": first_ms,
            "iterations_total": t, "kernel_ms_total": kms,
            "swap_accept": float(np.mean(swaps[swaps != 255])),
            "cold_draws_in_upper_mode": upper0,
            "cold_summary": "label-invariant (min(mu1, mu2), max(mu1, mu2))",
            "cold_ess_min": s["min_ess"],
            "cold_ess_per_s": s["min_ess"] / (kms / 1e3) if gate else None,
            "cold_rhat_max": s["max_rhat"],
            "cold_rhat_max_raw_labels": raw["max_rhat"],
            "cold_mean": [p["mean"] for p in s["params"]],
            "cold_sd": [p["sd"] for p in s["params"]],
            "engine": "compiled per model" if a.jit else "interpreter"}


def batch_section(a):
    """512 independent linear regressions of log-uniform size in [1e3, 1e6]
    rows (x1..x8, FP32), DE-MC with 64 chains each, every batch of every
    model in one launch and all batches one HIP graph; tiles autotuned."""
    from gpumetropolis_amd import workloads
    from gpumetropolis_amd.batch import gpu_metropolis_batch
    from gpumetropolis_amd.diagnostics import summarize

    t0 = time.perf_counter()
    wls = workloads.regressions(n_models=a.batch_models)
    gen_s = time.perf_counter() - t0
    n_tot = sum(w.n_obs for w in wls)
    rng = np.random.default_rng(a.seed)
    # chains start around the truth (posterior sds are 1e-3 .. 3e-2)
    inits = [w.meta["true"] + 0.01 * rng.standard_normal((64, 10)) for w in wls]
    t0 = time.perf_counter()
    bf = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls],
                              n_chains=64, n_iter=a.batch_iters,
                              warmup=a.batch_warmup,
                              init=inits, proposal_sd=0.003, seed=a.seed,
                              tiling="auto", jit=a.jit)
    wall = time.perf_counter() - t0
    evals = n_tot * 64 * a.batch_iters
    # Every model against its closed-form posterior, with its R-hat and ESS.
    t0 = time.perf_counter()
    worst, rh, es, sdr, zs = 0.0, [], [], [], []
    for fit, w in zip(bf.fits, wls):
        p = workloads.regression_posterior(w.data)
        s = summarize(fit.draws)
        fit.draws = None                 # free the host copy as we go
        rh.append(s["max_rhat"])
        es.append(s["min_ess"])
        for j, row in enumerate(s["params"]):
            zs.append((row["mean"] - p["mean"][j]) / row["mcse"])
            sdr.append(row["sd"] / p["sd"][j])
    worst = float(np.max(np.abs(zs)))
    check_s = time.perf_counter() - t0
    info = dict(bf.fits[0].adaptation)
    rh, es = np.array(rh), np.array(es)
    z = np.array(zs)
    return {"workload": f"{len(wls)} linear regressions, {n_tot} FP32 rows x "
                        f"9 cols, DE-MC 64 chains each, {a.batch_iters} "
                        f"iterations ({a.batch_warmup} warmup, "
                        f"{bf.batch_sizes[0]} per batch)",
            "obs_evals_per_s": evals / (bf.kernel_ms / 1e3),
            "graph_ms": bf.kernel_ms, "wall_s_incl_setup": wall,
            "data_gen_s": gen_s,
            "tiling": list(bf.tiling), "tile_key": bf.tile_key,
            "tile_timings_ms": {"x".join(map(str, k)): v
                                for k, v in bf.tile_timings.items()},
            "workgroups_per_launch": bf.n_workgroups,
            "launches": 2 * info["n_batches"],
            "closed_form_max_mean_err_mcse": worst,
            "closed_form_frac_within_3_mcse": float(np.mean(np.abs(z) <= 3)),
            "closed_form_err_mcse_rms": float(np.sqrt(np.mean(z * z))),
            "closed_form_sd_ratio_range": [float(min(sdr)), float(max(sdr))],
            "rhat_max": float(rh.max()),
            "models_rhat_below_1_01": int((rh < 1.01).sum()),
            "ess_min": float(es.min()),
            "ess_median": float(np.median(es)),
            "ess_per_s_min_model": float(es.min() / (bf.kernel_ms / 1e3)),
            "ess_per_s_all_models": float(es.sum() / (bf.kernel_ms / 1e3)),
            "check_s": check_s,
            "mean_accept": float(np.mean([f.accept_rate.mean()
                                          for f in bf.fits]))}


def eval_section(a):
    """FP64 batch evaluation at many parameter points: the per-observation
    log-likelihood matrix (64 posterior draws x the 1e6 gamma rows, 512 MB
    of output in one device chunk) and the gradient of the summed data term
    at a Cramer-Rao stencil (33 points = centre +- h e_k, logistic-16, 1e6
    rows).  Kernel time from HIP events; the oracle times a sample of the
    same calls on one host thread."""
    from gpumetropolis_amd import _native, workloads
    from oracle import oracle

    _native.set_eval_jit(a.jit)
    out = {}
    wg = workloads.gamma(n=a.n_obs)
    mg = wg.model
    rng = np.random.default_rng(a.seed)
    pts = np.array([np.log(2.5), np.log(0.7)]) + 0.002 * rng.standard_normal((64, 2))
    args = (mg.loglik.code, mg.loglik.consts, 2, wg.data, 1, wg.n_obs)
    _native.loglik_pointwise(*args, pts[:2])  # warm the kernel
    t0 = time.perf_counter()
    _native.loglik_pointwise(*args, pts)
    wall = time.perf_counter() - t0
    ms = _native.last_eval_ms()
    evals = 64 * wg.n_obs
    t0 = time.perf_counter()
    oracle.loglik_pointwise(*args, pts[:2])
    cpu_rate = 2 * wg.n_obs / (time.perf_counter() - t0)
    # Algorithmic bytes per evaluation: the FP64 output written, and the
    # FP64 data (one column) read once per row for the 64 points.
    bpe = 8 + 8 * wg.n_cols / 64
    gbps = bpe * evals / (ms / 1e3) / 1e9
    out["pointwise"] = {
        "workload": f"gamma, {wg.n_obs} rows x 64 points, FP64",
        "compiled": _native.last_eval_compiled(),
        "kernel_ms": ms, "evals_per_s": evals / (ms / 1e3),
        "wall_s_incl_pcie": wall,
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": gbps * 1e9 / HBM_PEAK,
                     "bytes_per_eval": bpe},
        "cpu_oracle_evals_per_s": cpu_rate, "cpu_cores": 1}
    wl = workloads.logistic(n=1_000_000)
    ml = wl.model
    h = 1e-3
    centre = np.zeros(16)
    stencil = np.vstack([centre] + [centre + s * h * np.eye(16)[k]
                                    for k in range(16) for s in (1, -1)])
    args = (ml.loglik.code, ml.loglik.consts, 16, wl.data, wl.n_cols, wl.n_obs)
    _native.grad_batch(*args, stencil[:1])
    t0 = time.perf_counter()
    _native.grad_batch(*args, stencil)
    wall = time.perf_counter() - t0
    ms = _native.last_eval_ms()
    evals = len(stencil) * wl.n_obs
    t0 = time.perf_counter()
    oracle.grad_batch(*args, stencil[:1])
    cpu_rate = wl.n_obs / (time.perf_counter() - t0)
    out["gradient"] = {
        "workload": f"logistic-16, {wl.n_obs} rows x {len(stencil)} points "
                    "(Cramer-Rao stencil), FP64 reverse mode",
        "compiled": _native.last_eval_compiled(),
        "kernel_ms": ms, "obs_grads_per_s": evals / (ms / 1e3),
        "wall_s_incl_upload": wall,
        "cpu_oracle_obs_grads_per_s": cpu_rate, "cpu_cores": 1}
    return out


def de_sync_section(a):
    """Synchronous Differential Evolution (path B) on the gamma model's
    1e6 FP32 rows: 1024 chains, 200 generations, one launch sequence per
    generation (proposal, exact stage over every CU, accept)."""
    from gpumetropolis_amd import _native, workloads

    w = workloads.gamma(n=a.n_obs)
    m = w.model
    C, T = 1024, 200
    rng = np.random.default_rng(a.seed)
    init = np.array([math.log(2.5), math.log(0.7)]) + 0.002 * rng.standard_normal((C, 2))
    dm = _native.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                             m.prior.consts, 2, w.data, 1, w.n_obs, jit=a.jit)
    g = 2.38 / math.sqrt(4.0)
    sd = np.full((C, 2), 1e-4)
    dm.run_da(None, init, np.ones(C), 5, a.seed, proposal_sd=sd,
              de_sync=(g, 1e-3))
    r = dm.run_da(None, init, np.ones(C), T, a.seed, proposal_sd=sd,
                  de_sync=(g, 1e-3))
    dm.close()
    evals = C * w.n_obs * T
    return {"workload": f"DE path B, gamma, {w.n_obs} FP32 rows, {C} chains "
                        f"x {T} generations",
            "obs_evals_per_s": evals / (r.kernel_ms / 1e3),
            "ms_per_generation": r.kernel_ms / T,
            "mean_accept": float(np.mean(r.accept_rate))}


def mala_section(a):
    """MALA (Langevin proposals with the FP64 value-and-gradient kernel) on
    the 16-dim logistic regression: 1e5 FP32 rows, 256 chains, 50
    iterations, Cholesky preconditioner; the oracle times one chain's
    iterations on one host thread."""
    from gpumetropolis_amd import _native, workloads
    from oracle import oracle

    w = workloads.logistic(n=100_000, n_chains=256)
    m = w.model
    d, C, T = m.n_params, 256, 50
    rng = np.random.default_rng(a.seed)
    init = w.meta["true"] + 0.003 * rng.standard_normal((C, d))
    sd = np.full((C, d), 0.004)
    L = np.tile(0.004 * np.eye(d), (C, 1, 1)).reshape(-1)
    args = (m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts, d,
            w.data, w.n_cols, w.n_obs)
    _native.run_chains(*args, init, sd, np.ones(C), 2, a.seed, proposal_mode=3,
                       proposal_l=L)
    _native.set_eval_jit(a.jit)
    dm = _native.DeviceModel(*args)
    r = dm.run_da(None, init, np.ones(C), T, a.seed, proposal_sd=sd,
                  proposal_l=L, langevin=True)
    dm.close()
    grads = C * w.n_obs * T
    t0 = time.perf_counter()
    oracle.run(*args, init[:1], sd[:1], np.ones(1), 2, a.seed, proposal_mode=3,
               proposal_l=L[:d * d], n_threads=1)
    cpu = w.n_obs * 3 / (time.perf_counter() - t0)  # 2 iterations + the start
    return {"workload": f"MALA, logistic-16, {w.n_obs} FP32 rows, {C} chains "
                        f"x {T} iterations, Cholesky preconditioner, FP64 "
                        "value and gradient",
            "obs_grads_per_s": grads / (r.kernel_ms / 1e3),
            "ms_per_iteration": r.kernel_ms / T,
            "mean_accept": float(np.mean(r.accept_rate)),
            "cpu_oracle_obs_grads_per_s": cpu, "cpu_cores": 1}


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must "
                          "be 1 (the multi-GPU split is designed, not "
                          "launched)"}))
        sys.exit(2)
    # torch is not imported: it bundles its own HIP runtime, and the
    # library's calls already return after their stream has drained.
    from gpumetropolis_amd import _native, adaptive as ad, workloads
    from gpumetropolis_amd.api import device_warmup
    from gpumetropolis_amd.diagnostics import summarize

    if not _native.available():
        raise RuntimeError("bench needs a gfx950 GPU: " + _native.last_error())
    if a.only:
        fn = {"pt": pt_section, "batch": batch_section, "da": da_section,
              "eval": eval_section, "gaussian": gaussian_mean_section,
              "m1": m1_cells_section}[a.only]
        print(json.dumps({a.only: fn(a)}), flush=True)
        return

    w = workloads.gamma(n=a.n_obs, n_chains=a.chains)
    m = w.model
    C, P, N = a.chains, m.n_params, w.n_obs
    dm = _native.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                             m.prior.consts, P, w.data, 1, N, jit=a.jit)
    info = dm.info()
    print(f"[bench] model lowered: {info}", file=sys.stderr, flush=True)
    state = {"kernel_ms": []}

    def call(init_m, sd_m, iters, s, temps=None, mode=0, g=0.0, noise=0.0,
             L=None, trace=False, raw_draws=False):
        r = dm.run(init_m, sd_m, np.ones(C) if temps is None else temps,
                   iters, int(s) & 0xFFFFFFFF, proposal_mode=mode, gamma=g,
                   de_noise=noise, proposal_l=L, final_state=True, trace=trace,
                   raw_draws=raw_draws)
        state["kernel_ms"].append(r.kernel_ms)
        return r

    # Initial values drawn around the prior scale, as gpu_metropolis does.
    from gpumetropolis_amd.philox_host import normals
    init = normals(a.seed, C * P).reshape(C, P)
    sd0 = np.full((C, P), 0.1)
    t0 = time.perf_counter()
    plan = ad.warmup_batches(a.warmup * a.iters_per_step)
    # The reference's adaptive warmup (R/adaptive.R:101-254), every batch
    # and between-batch step on the device in one graph.
    wu = device_warmup(dm, P, C, init, sd0, a.warmup * a.iters_per_step,
                       a.seed)
    warm_s = time.perf_counter() - t0
    warm_acc = float(wu.accept_history[:, -1].mean())
    print(f"[bench] warmup {sum(plan)} it in {len(plan)} batches done in "
          f"{warm_s:.1f} s, last batch accept {warm_acc:.3f} (target "
          f"{ad.target_accept(P)})", file=sys.stderr, flush=True)

    state["kernel_ms"].clear()
    draws = []
    cur = wu.init
    # final_state of the last warmup batch is the exact FP64 state; the
    # warmup hands back the f32 draws, so continue from them as R does.
    mode = 2 if wu.L is not None else 0
    accs = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        # The draws arrive on the host in the library's (param, chain,
        # iteration) layout; they are rearranged for the diagnostics after
        # the timed region.
        r = call(cur, wu.sd, a.iters_per_step, a.seed + wu.batches_run + k,
                 mode=mode, L=wu.L, raw_draws=True)
        cur = r.final_state
        draws.append(r.draws_pct)
        accs.append(r.accept_rate)
        print(f"[bench] step {k + 1}/{a.steps} kernel {r.kernel_ms:.1f} ms",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0

    iters = a.steps * a.iters_per_step
    evals = C * N * iters
    value = evals / wall
    d = np.concatenate([x.transpose(2, 1, 0) for x in draws]).astype(np.float64)
    summ = summarize(d)
    ess_min = summ["min_ess"]
    # The exact posterior of (log_a, log_b) from the data's sufficient
    # statistics, and its Laplace approximation N(mode, I^-1).
    gp = workloads.gamma_posterior(w.data)
    mcse = np.array([p["mcse"] for p in summ["params"]])
    pmean = np.array([p["mean"] for p in summ["params"]])
    psd = np.array([p["sd"] for p in summ["params"]])
    posterior_check = {
        "reference": "exact posterior by quadrature on the sufficient "
                     "statistics; Laplace N(mode, I^-1) beside it",
        "exact_mean": gp["mean"].tolist(), "exact_sd": gp["sd"].tolist(),
        "max_err_mcse": float(np.max(np.abs(pmean - gp["mean"]) / mcse)),
        "laplace_max_err_mcse": float(np.max(np.abs(pmean - gp["mode"]) / mcse)),
        "sd_ratio": (psd / gp["sd"]).tolist()}
    kms = np.array(state["kernel_ms"])
    avg_launch_s = kms.mean() / 1e3
    evals_per_launch = C * N * a.iters_per_step
    w_exec = info["w_exec"]
    achieved_slots = w_exec * evals_per_launch / avg_launch_s
    # Data bytes one launch streams through the chips' caches into LDS:
    # every workgroup reads every column once per iteration.
    cpb = 16
    n_wg = math.ceil(C / cpb)
    bytes_per_launch = (a.iters_per_step * n_wg * N * info["n_cols_total"] * 4)
    rec = PROFILED_TRAFFIC.get(("compiled" if a.jit else "interpreter", C, N,
                                a.iters_per_step))
    traffic = rec["fetch_bytes"] + rec["write_bytes"] if rec else None
    out = {
        "metric": "ESS/s at R-hat<1.01 + obs-loglik evals/s; plain vs "
                  "surrogate-DA, 1 MI355X",
        "value": value,
        "unit": "obs-loglik evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 terms, fp64 sums",
        "data": "synthetic (seeded Gamma(2.5, 0.7) draws, FP32)",
        "config": {"workload": "plain RWM, gamma likelihood, 1e6 FP32 obs, "
                               f"{C} chains x {iters} sampling + "
                               f"{a.warmup * a.iters_per_step} warmup iterations",
                   "n_obs": N, "chains": C,
                   "iters_per_step": a.iters_per_step,
                   "parallelism": "single GPU"},
        "ess_per_s": ess_min / wall if summ["max_rhat"] < 1.01 else None,
        "rhat_max": summ["max_rhat"],
        "ess_min": ess_min,
        "sampling_accept": float(np.mean(accs)),
        "warmup_adaptation": {"iterations": sum(plan), "batches": len(plan),
                   "engine": "device: one HIP graph of every batch and "
                             "between-batch step",
                   "last_batch_accept": warm_acc,
                   "target_accept": ad.target_accept(P),
                   "kernel_ms": wu.kernel_ms, "wall_s": warm_s},
        "posterior_check": posterior_check,
        "posterior": [{"mean": p["mean"], "sd": p["sd"], "mcse": p["mcse"]}
                      for p in summ["params"]],
        "roofline": {
            "bound": "valu",
            "achieved": achieved_slots / 1e12,
            "peak": VALU_PEAK / 1e12,
            "unit": "T VALU-slot/s",
            "frac": achieved_slots / VALU_PEAK,
            "traffic": traffic,
            "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE of this "
                               "launch shape, scaled by a copy probe"
                               if traffic else None),
            "w_exec": w_exec,
            "w_bytecode": info["w_bytecode"],
            "kernel": ("gpum_jit_kernel (chain_body, compiled gamma program)"
                       if a.jit else "gpum::chain_kernel_affine<16,4,true> "
                       "(interpreter of affine programs)"),
            "avg_launch_ms": kms.mean(),
            "stream_bytes_per_launch": bytes_per_launch,
            "stream_GBps": bytes_per_launch / avg_launch_s / 1e9,
        },
    }
    if a.jit and a.vm_launches > 0:
        # The same launches on the interpreter (identical draws), for the
        # record of what compiling the model buys.
        dm.set_jit(False)
        vm_ms = []
        for k in range(a.vm_launches):
            r = dm.run(wu.init, wu.sd, np.ones(C), a.iters_per_step,
                       a.seed + wu.batches_run + k, proposal_mode=mode,
                       proposal_l=wu.L, final_state=True)
            vm_ms.append(r.kernel_ms)
        vm_avg = float(np.mean(vm_ms))
        out["interpreter"] = {
            "avg_launch_ms": vm_avg,
            "obs_evals_per_s": evals_per_launch / (vm_avg / 1e3),
            "valu_frac": w_exec * evals_per_launch / (vm_avg / 1e3) / VALU_PEAK}
    out["engine"] = "compiled per model" if a.jit else "interpreter"
    dm.close()
    if not a.no_gaussian:
        out["gaussian_mean"] = gaussian_mean_section(a)
    if not a.no_m1:
        out["m1_cells"] = m1_cells_section(a)
    if not a.no_da:
        out["surrogate_da"] = da_section(a)
    if not a.no_pt:
        out["parallel_tempering"] = pt_section(a)
    if not a.no_batch:
        out["batched_models"] = batch_section(a)
    if not a.no_eval:
        out["batch_evaluation"] = eval_section(a)
        out["de_sync"] = de_sync_section(a)
        out["mala"] = mala_section(a)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds, cur, wu.sd, wu.L,
                                           a.seed + wu.batches_run + a.steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
