"""Benchmark: min-ESS/sec over reactions on one MI355X.

Default workload ("genome"): the 5,800 x 10,600 synthetic genome-scale model
(3.2 nonzeros per column, 40% irreversible, 2% fixed, bounds +-(100..1000)
around an interior null-space point), 21,200 warmup points stored with the
6 decimals of the reference's "%f" warmup CSV, 20,000 chains x 5,000 accepted
steps. The default sampler is the learned one: a plain-ACHR training run
(1,000 steps, every chain traced), the direction network trained on those
200,000 samples, then the timed run along the learned direction bank; the
plain FP64 sampler runs the same file for comparison (speedup, moment
agreement, split-R-hat). "ecoli" is the 72 x 95 fixture model, 500 chains x
1,000 steps; "batch" is 64 models of 100..13,000 reactions in one job
(aggregate chain-steps/s); "precision" the projection rounds in FP64, FP32
and split-BF16x3 on 65,536 chains; "gemm" the residual round's GEMM alone at
that shape (kernel TFLOP/s against the MFMA peak); "convergence" short runs against a long
plain run; "warmup" the 21,200 flux-variability LPs of the genome model's
warmup set (LPs converged per second); "putida" the reference's own P. putida
model through both command lines, end to end, as its paper times it.

Besides the plain FP64 run from the warmup as read (the reference's
behaviour: a "%f" warmup is off S x = 0 by its rounding, so every check
reprojects), the default line carries a second plain FP64 run from the same
warmup moved onto {S x = 0, fixed reactions at their bound}, the speedup
against both, the rate with the training time charged to the one file, KS
statistics against each, and a `gate` object that states the headline
target's conditions (moments within 1% of range, drift <= 1e-9, mixed
samplers, 10x) as booleans.

One bench step = one file of the reference CLI (all chains, all steps), with
inputs already resident in HBM. The timed region is bracketed by device
synchronisation. Traces of `--trace-chains` chains (every 50 steps) give the
bulk ESS per reaction; min-ESS/sec scales the traced chains' ESS to all P
chains (chains are independent and identically distributed).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "achr.cu_amd"))

PROJ = {"fp64": 0, "res_bf16x3": 1, "res_f32": 2, "null_f32": 3, "null_bf16x3": 4}
# HBM bytes of one full launch of the learned step kernel (10 check periods: with
# FP64 rounds an unflagged chain runs through 10 checks per launch), measured with
# rocprofv3 FETCH_SIZE / WRITE_SIZE and calibrated on a 1 GiB device copy
# (profiles/step_kernel_bank_genome_pmc_traffic.md), with the algorithmic bytes of
# that launch, keyed by (reactions, chains, check period); other shapes report null.
PMC_TRAFFIC = {(10600, 20000, 100): (1761.1e9, 1699.4e9, "profiles/step_kernel_bank_genome_pmc_traffic.md")}
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def genome_model(seed: int = 2):
    import achr
    from achr import synth

    def nullspace_fn(rows, cols, vals, m, n):
        t = time.time()
        mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
        d = mdl.compute_nullspace()
        log(f"null space on GPU (rocSOLVER syevd of S^T S, n={n}): d={d} in {time.time() - t:.1f}s")
        return mdl.nullspace()

    t = time.time()
    m = synth.config2(nullspace_fn, seed=seed)
    log(f"model {m.m} x {m.n}, nnz {len(m.vals)} built in {time.time() - t:.1f}s")
    t = time.time()
    W = synth.round_like_csv(synth.chord_warmup(m.N, m.x0, m.lb, m.ub, 2 * m.n, seed=seed + 100))
    log(f"warmup {W.shape} in {time.time() - t:.1f}s")
    return m, W


def ecoli_model():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import load_c1
    from achr import synth

    S, lb, ub, Wf, _ = load_c1()
    N = synth.nullspace_dense(S)
    r, c = np.nonzero(S)
    m = synth.SynthModel(S, r.astype(np.int32), c.astype(np.int32), S[r, c], S.shape[0], S.shape[1], lb, ub,
                         None, N)
    return m, Wf


def cpu_baseline(m, W, N, P_cpu, S_cpu, threads, gpu_ess_per_step):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from achr import synth

    S_dense = None
    rp = np.zeros(m.m + 1, dtype=np.int32)
    order = np.lexsort((m.cols, m.rows))
    rows, cols, vals = m.rows[order], m.cols[order], m.vals[order]
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    del S_dense
    t = time.time()
    O.sample(warm=W, lb=m.lb, ub=m.ub, rowptr=rp, colidx=cols.astype(np.int32), vals=vals, N=N, m=m.m, P=P_cpu,
             S=S_cpu, seed=1, lanes=1024 if m.n > 4096 else 32, threads=threads)
    dt = time.time() - t
    steps_per_s = P_cpu * S_cpu / dt
    # one core: an eighth of the chains (at least 2), same steps
    P1 = max(2, P_cpu // 8)
    t = time.time()
    O.sample(warm=W, lb=m.lb, ub=m.ub, rowptr=rp, colidx=cols.astype(np.int32), vals=vals, N=N, m=m.m, P=P1,
             S=S_cpu, seed=1, lanes=1024 if m.n > 4096 else 32, threads=1)
    dt1 = time.time() - t
    return {"value": gpu_ess_per_step * steps_per_s, "unit": "min-ESS/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C, OpenMP) {P_cpu} chains x {S_cpu} steps of the same model and warmup in {dt:.1f}s "
                      f"= {steps_per_s:.4g} chain-steps/s; min-ESS/s = that rate x the GPU run's min-ESS per "
                      f"chain-step (both run the identical, bit-equal algorithm)",
            "chain_steps_per_s": steps_per_s,
            "single_core": {"chain_steps_per_s": P1 * S_cpu / dt1, "value": gpu_ess_per_step * P1 * S_cpu / dt1,
                            "sample": f"{P1} chains x {S_cpu} steps in {dt1:.1f}s on 1 thread"},
            "host": host_cpu()}


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "usable_cores": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}


def measure_copy_gbs(mib=1024, reps=5):
    """Device-to-device copy rate (read + write bytes / time, best of `reps`):
    the HBM rate a plain streaming kernel reaches on this GPU."""
    import ctypes as C
    import achr
    achr.lib()
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
    best = float("inf")
    try:
        if hip.hipMemset(a, 1, nbytes) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        for _ in range(reps):
            t = time.perf_counter()
            if hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():
                raise RuntimeError("hipMemcpy failed")
            best = min(best, time.perf_counter() - t)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * (mib << 20) / best / 1e9


def measure_read_gbs(gib=4, reps=5):
    """Read-only streaming rate (one pass over `gib` GiB of device memory,
    best of `reps`): the yardstick for kernels that mostly read."""
    import ctypes as C
    import achr
    v = C.c_double()
    achr._lib.check(achr.lib().achr_read_bandwidth(gib << 30, reps, C.byref(v)), "achr_read_bandwidth")
    return v.value


def measure(smp, P, S, steps, warmup, seed0=1):
    """Warmup files (20 steps each) then `steps` timed files; returns (wall, timing) of the slowest."""
    for i in range(warmup):
        smp.run(P, min(S, 20), seed=1000 + i)
        smp.timing()
    times, timings = [], []
    for k in range(steps):
        t0 = time.perf_counter()
        smp.run(P, S, seed=seed0 + k, file=k)
        tm = smp.timing()  # waits for the end event on the sampler's stream
        times.append(time.perf_counter() - t0)
        timings.append(tm)
        log(f"file {k}: {times[-1]:.3f}s wall, kernels {tm}")
    j = int(np.argmax(times))
    return times[j], timings[j]


def moving(lb, ub, *point_sets):
    """Reactions with a range that the chains actually move: ub > lb and a
    sample spread above 1e-9 of the range in at least one point set.
    Reactions pinned by S once the fixed reactions are fixed only carry
    rounding noise, which would dominate standard-error and R-hat maxima."""
    rng = ub - lb
    spread = np.max([p.std(1) for p in point_sets], axis=0)
    return (rng > 0) & (spread > 1e-9 * rng)


def sample_quality(m, smp, S, trace_stride_chains, P):
    from achr import stats as st
    stats_ = smp.stats()
    pts = smp.points()
    trace = smp.trace()
    live = moving(m.lb, m.ub, pts)
    ess = st.bulk_ess(trace[:, :, live])
    min_ess = float(np.nanmin(ess))
    if m.S is not None:
        drift = float(np.abs(m.S @ pts).max())
    else:
        import scipy.sparse as sp
        drift = float(np.abs(sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, m.n)) @ pts).max())
    rhat = st.rhat(trace[:, :, live])
    return {"stats": stats_, "pts": pts, "min_ess": min_ess, "ess_per_step": min_ess / (trace.shape[0] * S),
            "traced": int(trace.shape[0]), "drift": drift, "max_rhat": float(np.nanmax(rhat))}


def roofline_for(args_proj, m, n, d, P, S, q, tm, bank=False):
    """Roofline of the dominant kernel. Step kernel: 8n B per attempt (the
    gathered warmup point or bank direction) plus, per chain and resumed
    segment, the state read and written back: x and the centre (32n B), or
    x alone in bank mode (16n B)."""
    reproj = int(q["stats"][:, 1].sum())
    attempts = P * S + int(q["stats"][:, 0].sum())
    segs = tm["step_launches"]
    state = 16.0 if bank else 32.0
    if bank and args_proj == "fp64" and reproj == 0:
        # FP64 rounds: a chain no check flags stops only after every 10th check of a launch
        # (the library's check budget), so it resumed checks // 10 times; the launches
        # enqueued after the chains finished are empty
        resumed = int((q["stats"][:, 2] // 10).sum())
    else:
        resumed = P * max(segs - 1, 0)  # every chain stops at every check
    step_bytes = 8.0 * n * attempts + state * n * resumed
    step_ms, proj_ms = tm["ms_step"], tm["ms_proj"]
    if args_proj == "fp64":
        proj_flops = 4.0 * n * d * reproj
        kname, peak, per = "dgemm_mfma_kernel (batched N (N^T X), FP64 MFMA)", PEAK_FP64_TFLOPS, "4*n*d"
    else:
        proj_flops = reproj * (0.9 * 2.0 * n * m.m + 0.1 * 4.0 * n * d)
        kname = f"gemm_kernel ({args_proj}: X -= S+ (S X)) + FP64 rounds"
        peak = PEAK_BF16_TFLOPS / 3 if "bf16" in args_proj else PEAK_FP32_TFLOPS
        per = "2*n*m (residual) / 4*n*d (FP64 round)"
    if proj_ms > step_ms:
        ach = proj_flops / (proj_ms * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": kname, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                "frac": ach / peak, "traffic": None, "launches": tm["proj_launches"],
                "per_launch": f"{per} flops per flagged chain; {reproj} chain reprojections over "
                              f"{tm['proj_launches']} rounds"}
    ach = step_bytes / (step_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "step_kernel", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": ach / PEAK_HBM_GBS, "traffic": None, "launches": segs,
            "per_launch": f"8n B per attempt (direction gather) + {int(state)}n B per chain per resumed segment; "
                          f"{attempts} attempts, {resumed} chain resumptions, {segs} launches"}


def check_kernel_rate(n, q, tm):
    """The residual-check kernel of the batched rounds (plain chains of 512
    or more lanes): it reads each checked chain's x once (8n B) and sums the
    rows of S from registers or L2."""
    if not tm.get("check_launches"):
        return None
    checks = int(q["stats"][:, 2].sum())
    bytes_ = 8.0 * n * checks
    return {"ms": tm["ms_check"], "launches": tm["check_launches"], "checks": checks,
            "GBs_algorithmic": bytes_ / (tm["ms_check"] * 1e-3) / 1e9, "per_check": "8n B (the chain's x)"}


def gpu_nullspace(rows, cols, vals, m, n):
    import achr
    mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
    mdl.compute_nullspace()
    return mdl.nullspace()


def batch_model(n, seed):
    """One model of the batch workload with the basis of its free null space
    (GPU null spaces)."""
    from achr import synth
    return synth.batch_model(n, seed, gpu_nullspace)


def batch_warmup(sm, Nf, seed):
    """2n warmup points on free null-space chords (exact FP64; each chord
    direction mixes 8 free basis vectors)."""
    from achr import synth
    return synth.chord_warmup(None, sm.x0, sm.lb, sm.ub, 2 * sm.n, seed=seed, free=Nf, mix=8)


def csr_of_model(sm):
    from achr import synth
    return synth.csr(sm)


def run_batch(args):
    """64 models, n log-uniform in [100, 13000], 2,000 chains x 2,000 steps
    each, in one batch: autotuned lanes per size class, one persistent launch
    per schedule bucket and segment, segments replayed from HIP graphs."""
    import achr
    import scipy.sparse as sp
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    count, P, S, seed = args.models, args.chains or 2000, args.chain_steps or 2000, 3
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(100), np.log(13000), count)).round().astype(int)
    t0 = time.time()
    models, smps, info, cpu_sample = [], [], [], []
    cpu_t, cpu_steps = 0.0, 0
    threads = min(16, os.cpu_count() or 1)
    built = []
    for k, n in enumerate(ns):
        built.append(batch_model(int(n), seed * 1000 + k))
    log(f"{count} models and free null spaces in {time.time() - t0:.1f}s")
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=8) as pool:
        futs = [pool.submit(batch_warmup, sm, Nf, seed * 1000 + k + 1) for k, (sm, Nf) in enumerate(built)]
        warms = [f.result() for f in futs]
    log(f"warmups in {time.time() - t0:.1f}s")
    for k, (sm, _) in enumerate(built):
        W = warms[k]
        warms[k] = None
        mdl = achr.Model.from_coo(sm.n, sm.m, sm.rows, sm.cols, sm.vals, sm.lb, sm.ub)
        mdl.set_nullspace(sm.N)
        smp = achr.Sampler(mdl, W, achr.Options(model_id=k))
        if not args.no_cpu:
            rp, ci, va = csr_of_model(sm)
            t = time.time()
            O.sample(warm=W, lb=sm.lb, ub=sm.ub, rowptr=rp, colidx=ci, vals=va, N=sm.N, m=sm.m, P=args.cpu_chains,
                     S=args.cpu_steps, seed=1, model=k, lanes=smp.schedule()[0], threads=threads)
            cpu_t += time.time() - t
            cpu_steps += args.cpu_chains * args.cpu_steps
        models.append((sm, mdl))
        smps.append(smp)
        info.append({"n": int(sm.n), "m": int(sm.m), "d": int(sm.N.shape[1])})
        del W
        if k % 8 == 7:
            log(f"{k + 1}/{count} samplers in {time.time() - t0:.1f}s")
    del built
    t_build = time.time() - t0
    batch = achr.Batch(smps)
    t = time.time()
    probe = batch.autotune([P] * count, probe_steps=args.probe_steps, seed=1)
    t_tune = time.time() - t
    lanes = batch.lanes()
    classes = {}
    for k in range(count):
        cl = int(np.floor(np.log2(info[k]["n"])))
        classes.setdefault(cl, {"n_range": [info[k]["n"], info[k]["n"]], "models": 0, "lanes": int(lanes[k]),
                                "probe_ms": {str(achr.Batch.TPC[j]): round(float(probe[k, j]), 3)
                                             for j in range(7) if probe[k, j] > 0}})
        c = classes[cl]
        c["models"] += 1
        c["n_range"] = [min(c["n_range"][0], info[k]["n"]), max(c["n_range"][1], info[k]["n"])]
    log(f"autotune {t_tune:.1f}s: {classes}")
    graph = not args.no_graph
    for i in range(args.warmup):
        batch.run([P] * count, min(S, 20), seed=1000 + i, graph=graph)
    walls, tms = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        batch.run([P] * count, S, seed=seed + i, file=i, graph=graph)
        walls.append(time.perf_counter() - t)
        tms.append(batch.timing())
        log(f"batch file {i}: {walls[-1]:.3f}s wall, {tms[-1]}")
    j = int(np.argmax(walls))
    wall, tm = walls[j], tms[j]
    total_steps = count * P * S
    pts_b = [s.points() for s in smps]
    st_b = [s.stats() for s in smps]
    # the same models one at a time (the sampler's own launches), another seed
    t = time.perf_counter()
    for s in smps:
        s.run(P, S, seed=seed + 100)
    for s in smps:
        s.timing()
    wall_solo = time.perf_counter() - t
    log(f"solo runs: {wall_solo:.3f}s")
    dmean, dstd, zmean, drift = [], [], [], []
    for k, s in enumerate(smps):
        sm = models[k][0]
        pb, ps = pts_b[k], s.points()
        rng_ = sm.ub - sm.lb
        live = moving(sm.lb, sm.ub, pb, s.points())
        dm = np.abs(pb.mean(1) - ps.mean(1))[live]
        se = np.sqrt((pb.var(1) + ps.var(1))[live] / P)
        dmean.append(float((dm / rng_[live]).max()))
        dstd.append(float((np.abs(pb.std(1) - ps.std(1))[live] / rng_[live]).max()))
        zmean.append(float((dm / np.maximum(se, 1e-300)).max()))
        Ssp = sp.csr_matrix((sm.vals, (sm.rows, sm.cols)), shape=(sm.m, sm.n))
        drift.append(float(np.abs(Ssp @ pb).max()))
    attempts = sum(int(st[:, 0].sum()) + P * S for st in st_b)
    reproj = sum(int(st[:, 1].sum()) for st in st_b)
    sum_n = sum(i["n"] for i in info)
    step_bytes = sum(8.0 * i["n"] * (P * S + int(st_b[k][:, 0].sum())) + 32.0 * i["n"] * P * (tm["segments"] - 1)
                     for k, i in enumerate(info))
    ach = step_bytes / (tm["ms_total"] * 1e-3) / 1e9
    out = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "metric": f"aggregate chain-steps/s ({count}-model batch, plain FP64 ACHR)",
           "value": total_steps / wall, "unit": "chain-steps/s", "ms_per_step": wall * 1e3,
           "config": {"workload": f"batch: {count} synthetic models, n log-uniform in [100, 13000] "
                                  f"(sum {sum_n}), m = 0.55 n, {P} chains x {S} steps each, 2n exact warmup points",
                      "models": count, "chains_per_model": P, "chain_steps": S, "reactions_sum": sum_n,
                      "reactions_min": int(min(ns)), "reactions_max": int(max(ns))},
           "batch": {**tm, "graph": graph, "autotune_s": t_tune, "build_s": t_build, "size_classes": classes},
           "solo": {"s_total": wall_solo, "chain_steps_per_s": total_steps / wall_solo,
                    "batch_speedup": wall_solo / wall},
           "moment_agreement_vs_solo_fp64": {
               "mean_frac_max": max(dmean), "mean_frac_median_model": float(np.median(dmean)),
               "std_frac_max": max(dstd), "std_frac_median_model": float(np.median(dstd)),
               "mean_z_max": max(zmean), "note": "per model: max over reactions of |batch - solo| / (ub - lb), "
                                                  "independent seeds; z = |d mean| / its standard error"},
           "drift_inf_max": max(drift), "reprojections": reproj, "attempts": attempts,
           "roofline": {"bound": "hbm", "kernel": "step_batch_kernel (whole graph-replayed run)", "achieved": ach,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None,
                        "per_launch": "8n B per attempt + 32n B per chain per resumed segment, summed over models; "
                                      "time = the whole run including projection rounds"}}
    if not args.no_cpu:
        out["cpu_baseline"] = {"value": cpu_steps / cpu_t, "unit": "chain-steps/s", "cores": threads, "kind": "port",
                               "sample": f"oracle (C, OpenMP) {args.cpu_chains} chains x {args.cpu_steps} steps of "
                                         f"each of the {count} models in {cpu_t:.1f}s"}
    print(json.dumps(out), flush=True)


def warmup_roofline(n, m, B, rep):
    """HBM roofline of the batched PDHG iteration: per iteration the primal
    step reads and writes x and the running sum and writes x_bar (40 n B), the
    dual step reads and writes y and its sum (32 m B), and each step reads the
    other side's iterate once (8 m B + 8 n B); the KKT checks every 64
    iterations are counted in the time, not in the bytes."""
    per_iter = (48.0 * n + 40.0 * m) * B
    ach = per_iter * rep["iterations"] / (rep["ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "primal_kernel + dual_kernel (batched PDHG)", "achieved": ach,
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None,
            "per_launch": f"(48 n + 40 m) B bytes per iteration, n={n}, m={m}, B={B} LPs"}


def run_warmup(args):
    """Warmup generation for the genome model: the 2n flux-variability LPs of
    the reference's createWarmupPts (min and max of every reaction subject to
    S x = 0 and the bounds) solved at once by the batched PDHG solver, with an
    iteration cap that keeps the run near a minute; value = LPs converged to
    the solver tolerance per second. CPU baseline: scipy HiGHS (dual simplex,
    one thread) on a sample of the same LPs."""
    import achr
    import scipy.sparse as sp
    from scipy.optimize import linprog
    m, _ = genome_model()
    n = m.n
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    iters = args.warmup_iters
    log(f"warmup generation: {2 * n} LPs, at most {iters} iterations")
    achr.warmup_points(model, 2 * n, max_iters=64)  # untimed: device and code warm
    t = time.perf_counter()
    W, fva, conv, rep = achr.warmup_points(model, 2 * n, max_iters=iters)
    wall = time.perf_counter() - t
    log(f"warmup: {rep}, wall {wall:.1f}s")
    S = sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, n))
    # HiGHS on a sample of the same LPs (about a minute each on one core);
    # its optima also check the converged GPU columns
    rng = np.random.default_rng(0)
    sample = rng.choice(2 * n, size=args.cpu_lps, replace=False)
    cpu, solved, checked, worst = 0.0, 0, 0, 0.0
    for col in sample:
        i, sgn = col // 2, (1.0 if col % 2 == 0 else -1.0)
        c = np.zeros(n)
        c[i] = sgn
        log(f"HiGHS: LP {col} ({'min' if sgn > 0 else 'max'} of reaction {i})")
        t = time.perf_counter()
        r = linprog(c, A_eq=S, b_eq=np.zeros(m.m), bounds=list(zip(m.lb, m.ub)), method="highs-ds",
                    options={"time_limit": 170.0})
        cpu += time.perf_counter() - t
        log(f"HiGHS: status {r.status} in {time.perf_counter() - t:.1f}s")
        if r.status != 0:
            continue
        solved += 1
        if conv[col]:
            checked += 1
            worst = max(worst, abs(sgn * r.fun - fva[i, col % 2]) / (1 + abs(r.fun)))
    # an LP that hits the time limit bounds the rate from above
    cpu_rate = max(solved, 1) / cpu if cpu > 0 else None
    out = {"n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "metric": "flux-variability LPs converged per second (genome model warmup generation)",
           "value": rep["converged"] / wall, "unit": "LPs/s", "ms_per_step": wall * 1e3,
           "config": {"workload": f"warmup: {m.m}x{n} model, {2 * n} FVA LPs, batched PDHG, {iters} iterations at most",
                      "lps": 2 * n, "max_iters": iters},
           "converged": rep["converged"], "iterations": rep["iterations"], "solve_ms": rep["ms"],
           "checked_vs_highs": {"lps": checked, "max_rel_err": worst},
           "roofline": warmup_roofline(n, m.m, 2 * n, rep)}
    if args.cpu_lps > 0:
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "LPs/s", "cores": 1, "kind": "port",
                               "sample": f"scipy HiGHS dual simplex, {solved} of {args.cpu_lps} sampled LPs solved in "
                                         f"{cpu:.1f}s" + ("" if solved else " (time limit: the value is an upper bound)")}
    print(json.dumps(out), flush=True)


def run_precision(args):
    """Reduced-precision projection on the genome model with 65,536
    concurrent chains: the projection rounds of every 10th step in FP64 (the
    oracle's arithmetic, bit for bit), or on the matrix cores in FP32 or
    split-BF16x3 (residual form x - S+ (S x)) with an FP64 round every 100
    steps. Same seed for all three, so the reduced-precision runs are
    compared with the FP64 run chain by chain."""
    import achr
    import scipy.sparse as sp
    P, S = args.chains or 65536, args.chain_steps or 1000
    if args.load_inputs:
        m, W = load_inputs(args.load_inputs)
    else:
        m, W = genome_model()
    n, d = m.n, m.N.shape[1]
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(m.N)
    Ssp = sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, m.n))
    rng = m.ub - m.lb
    live = rng > 0
    out = {"n_gpus": 1, "steps": 1, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 bounds / f32, bf16x3 projection", "data": "synthetic"}
    res, ref = {}, None
    for name in ("fp64", "res_f32", "res_bf16x3"):
        t = time.time()
        smp = achr.Sampler(model, W, achr.Options(proj_prec=PROJ[name], fp64_every=args.fp64_every,
                                                  reproject_mode=2))
        log(f"{name}: sampler ready in {time.time() - t:.1f}s")
        wall, tm = measure(smp, P, S, 1, args.warmup, seed0=7)
        pts, st = smp.points(), smp.stats()
        del smp
        r = {"s_per_file": wall, "chain_steps_per_s": P * S / wall, "kernel_ms": tm,
             "drift_inf": float(np.abs(Ssp @ pts).max()), "reprojections": int(st[:, 1].sum()),
             "checks": int(st[:, 2].sum())}
        if ref is None:
            ref = pts
        else:
            A, B = pts[live], ref[live]
            dm = np.abs(A.mean(1) - B.mean(1))
            dsd = np.abs(A.std(1) - B.std(1))
            r["vs_fp64"] = {"mean_frac_max": float((dm / rng[live]).max()),
                            "std_frac_max": float((dsd / rng[live]).max()),
                            "point_abs_diff_max_frac": float((np.abs(A - B).max(1) / rng[live]).max()),
                            "note": "same seed: chain-by-chain against the FP64 run, fractions of ub - lb"}
            del A, B
        # projection arithmetic: FP64 rounds 4 n d flop per flagged chain;
        # residual rounds 2 n m (the S+ contraction; S x comes from the check)
        fp64_rounds = tm["proj_launches"] if name == "fp64" else tm["proj_launches"] // args.fp64_every
        low_rounds = tm["proj_launches"] - fp64_rounds
        per_round_chains = r["reprojections"] / max(tm["proj_launches"], 1)
        r["proj_tflops"] = (per_round_chains * (fp64_rounds * 4.0 * n * d + low_rounds * 2.0 * n * m.m)
                            / (tm["ms_proj"] * 1e-3) / 1e12)
        res[name] = r
        log(f"{name}: {r}")
    base = res["fp64"]["s_per_file"]
    for name in ("res_f32", "res_bf16x3"):
        res[name]["speedup_vs_fp64"] = base / res[name]["s_per_file"]
    best = res["res_bf16x3"]
    out.update({"metric": "chain-steps/s (genome model, 65,536 chains, split-BF16x3 projection rounds)",
                "value": best["chain_steps_per_s"], "unit": "chain-steps/s", "ms_per_step": best["s_per_file"] * 1e3,
                "config": {"workload": f"precision: {m.m}x{n} model (d={d}), {P} chains x {S} steps, "
                                       f"FP64 round every {args.fp64_every * 10} steps", "chains": P, "chain_steps": S},
                "variants": res})
    print(json.dumps(out), flush=True)


def run_gemm(args):
    """The residual projection round's GEMM alone, at the precision
    workload's shape: 65,536 residual rows over K = 5,800 metabolites times
    the packed S+ (N = 10,600 reactions), subtracted from FP64 chains, on
    synthetic operands, in FP32 MFMA and split-BF16x3. Kernel time from HIP
    events on the kernel's stream (packing and host copies excluded). No
    solver runs in the process, so counter runs can profile it."""
    import achr
    P = args.chains or 65536
    K, N = 5800, 10600
    rng = np.random.default_rng(3)
    A = rng.standard_normal((P, K), dtype=np.float32)
    B = rng.standard_normal((K, N), dtype=np.float32) * np.float32(1e-2)
    C = np.zeros((P, N))
    flop = 2.0 * P * K * N
    res = {}
    for prec, name, passes, peak in ((0, "f32", 1, 157.3), (1, "bf16x3", 3, 2500.0)):
        times = []
        for _ in range(args.warmup + args.steps):
            _, ms = achr.gemm_sub_f64(A, B, C, prec=prec, timing=True)
            times.append(ms)
        ms = float(np.median(times[args.warmup:]))
        tf = flop / (ms * 1e-3) / 1e12
        res[name] = {"kernel_ms": ms, "launch_ms": times, "useful_tflops": tf, "mfma_tflops": tf * passes,
                     "mfma_peak_tflops": peak, "frac": tf * passes / peak}
        log(f"gemm_sub {name}: {res[name]}")
    b = res["bf16x3"]
    out = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 (split-bf16x3 MFMA)", "data": "synthetic",
           "metric": "TFLOP/s of the residual projection round's GEMM (useful work, split-BF16x3)",
           "value": b["useful_tflops"], "unit": "TFLOP/s", "ms_per_step": b["kernel_ms"],
           "config": {"workload": f"gemm: {P} residual rows x {K} metabolites x {N} reactions, FP64 update",
                      "rows": P, "K": K, "N": N},
           "roofline": {"bound": "mfma", "kernel": "gemm_sub_kernel<split-BF16x3>", "achieved": b["mfma_tflops"],
                        "peak": 2500.0, "unit": "TFLOP/s", "frac": b["frac"], "traffic": None,
                        "per_launch": "3 BF16 MFMA passes x 2 P K N flop"},
           "variants": res}
    print(json.dumps(out), flush=True)


def run_convergence(args):
    """Moment error of plain and learned ACHR against a long plain run on the
    genome model: the reference is plain ACHR with --ref-steps steps (residual
    split-BF16x3 projection rounds with FP64 every 100 steps, within 3e-5 of
    range of the FP64 rounds); each short run is scored per reaction, as a
    fraction of ub - lb, against it. All runs start at the warmup centre."""
    import achr
    P = args.chains or 20000
    S_short, S_long = args.chain_steps or 5000, args.ref_steps
    if args.load_inputs:
        m, W = load_inputs(args.load_inputs)
    else:
        m, W = genome_model()
    n = m.n
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(m.N)
    rng = m.ub - m.lb
    live = rng > 0
    tc = min(args.trace_chains, P)

    def summary(pts, ref):
        A, B = pts[live], ref[live]
        dm = np.abs(A.mean(1) - B.mean(1)) / rng[live]
        ds = np.abs(A.std(1) - B.std(1)) / rng[live]
        return {"mean_frac_max": float(dm.max()), "std_frac_max": float(ds.max()),
                "mean_frac_q99": float(np.quantile(dm, 0.99)), "std_frac_q99": float(np.quantile(ds, 0.99)),
                "mean_frac_median": float(np.median(dm)), "std_frac_median": float(np.median(ds)),
                "reactions_over_1pct": int(((dm > 0.01) | (ds > 0.01)).sum())}

    def run(smp, S, seed, thin):
        from achr import stats as st
        smp.set_run_options(tc, thin, smp.opts.check_every)
        t = time.perf_counter()
        smp.run(P, S, seed=seed)
        tm = smp.timing()
        wall = time.perf_counter() - t
        tr = smp.trace()
        rh = st.rhat(tr[:, :, live])
        return smp.points(), {"s": wall, "kernel_ms": tm["ms_total"], "max_rhat": float(np.nanmax(rh))}

    res = {}
    plain = achr.Sampler(model, W, achr.Options(proj_prec=PROJ["res_bf16x3"], reproject_mode=2))
    ref, info = run(plain, S_long, 11, max(S_long // 100, 1))
    res[f"plain_{S_long}_reference"] = info
    log(f"reference: {info}")
    pts, info = run(plain, S_short, 12, max(S_short // 100, 1))
    res[f"plain_{S_short}"] = {**info, **summary(pts, ref)}
    log(f"plain short: {res[f'plain_{S_short}']}")
    del plain, pts
    learned = achr.Sampler(model, W, achr.Options(lanes=0, trace_chains=P, trace_thin=args.train_thin,
                                                  proj_prec=PROJ["res_bf16x3"], reproject_mode=2))
    learned.run(P, args.train_steps, seed=99)
    learned.timing()
    rep = learned.learn_directions(rank=args.rank, width=512, iters=args.iters, batch=2048, lr=1e-3, seed=7)
    learned.opts.check_every = args.check_every
    for S in (S_short, S_long):
        pts, info = run(learned, S, 13, max(S // 100, 1))
        res[f"learned_{S}"] = {**info, **summary(pts, ref)}
        log(f"learned {S}: {res[f'learned_{S}']}")
    out = {"n_gpus": 1, "steps": 1, "warmup": 0, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "metric": f"max per-reaction |mean| / |std| error vs a {S_long}-step plain run (fraction of ub - lb)",
           "value": res[f"learned_{S_short}"]["std_frac_max"], "unit": "fraction of range",
           "config": {"workload": f"convergence: genome {m.m}x{n}, {P} chains, short runs {S_short} steps, "
                                  f"reference {S_long} steps of plain ACHR", "chains": P},
           "learn": rep, "runs": res}
    print(json.dumps(out), flush=True)


_MODEL_FIELDS = ("rows", "cols", "vals", "lb", "ub", "x0", "N")


def save_inputs(path, m, W):
    os.makedirs(path, exist_ok=True)
    for f in _MODEL_FIELDS:
        np.save(os.path.join(path, f + ".npy"), getattr(m, f))
    np.save(os.path.join(path, "dims.npy"), np.array([m.m, m.n]))
    np.save(os.path.join(path, "W.npy"), W)


def load_inputs(path):
    from achr import synth
    v = {f: np.load(os.path.join(path, f + ".npy")) for f in _MODEL_FIELDS}
    mm, n = (int(x) for x in np.load(os.path.join(path, "dims.npy")))
    m = synth.SynthModel(None, v["rows"], v["cols"], v["vals"], mm, n, v["lb"], v["ub"], v["x0"], v["N"])
    log(f"inputs loaded from {path}")
    return m, np.load(os.path.join(path, "W.npy"))


def calibrate_copy(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime (a
    blit kernel reading and writing exactly that many bytes)."""
    import ctypes as C
    import achr
    achr.lib()  # loads the HIP runtime the library links
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
    rc = hip.hipMemset(a, 1, nbytes) or hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError(f"calibration copy failed ({rc})")
    log(f"calibration copy: {mib} MiB device to device")


def compare_endpoints(m, A_pts, B_pts, P, ks_samples=4096):
    """Per-reaction agreement of two sets of P independent chain endpoints:
    |mean| and |std| differences as fractions of ub - lb and in standard
    errors, and the two-sample KS statistic on the first ks_samples
    endpoints of each set (with its critical values at 1%, for one reaction
    and Bonferroni over all of them)."""
    from achr import stats as st
    rng = m.ub - m.lb
    live = moving(m.lb, m.ub, A_pts, B_pts)
    A, B = A_pts[live], B_pts[live]
    dm_abs = np.abs(A.mean(1) - B.mean(1))
    dm = dm_abs / rng[live]
    ds = np.abs(A.std(1) - B.std(1)) / rng[live]
    se_m = np.sqrt((A.var(1) + B.var(1)) / P)
    se_s = np.sqrt((A.var(1) + B.var(1)) / (2 * P))
    dsd_abs = np.abs(A.std(1) - B.std(1))
    k = min(ks_samples, A.shape[1], B.shape[1])
    D = st.ks_2samp(A[:, :k].T, B[:, :k].T)
    nlive = int(live.sum())
    c1 = np.sqrt(-0.5 * np.log(0.01 / 2)) * np.sqrt(2.0 / k)
    cb = np.sqrt(-0.5 * np.log(0.01 / (2 * nlive))) * np.sqrt(2.0 / k)
    return {"mean_frac_max": float(dm.max()), "std_frac_max": float(ds.max()),
            "mean_frac_median": float(np.median(dm)), "std_frac_median": float(np.median(ds)),
            "mean_z_max": float((dm_abs / np.maximum(se_m, 1e-300)).max()),
            "std_z_max": float((dsd_abs / np.maximum(se_s, 1e-300)).max()),
            "mean_z_q50_q99": [float(np.quantile(dm_abs / np.maximum(se_m, 1e-300), x)) for x in (0.5, 0.99)],
            "reactions_mean_frac_over_1pct": int((dm > 0.01).sum()),
            "reactions_std_frac_over_1pct": int((ds > 0.01).sum()),
            "ks": {"D_max": float(D.max()), "D_median": float(np.median(D)), "samples_per_side": k,
                   "critical_1pct_one_reaction": float(c1), "critical_1pct_bonferroni": float(cb),
                   "reactions_over_bonferroni": int((D > cb).sum())},
            "z_max_of_pure_noise_approx": float(np.sqrt(2.0 * np.log(nlive))),
            "moving_reactions": nlive,
            "note": "|learned - plain| over the live reactions, as a fraction of ub - lb and in standard errors "
                    "of the difference (P independent chain endpoints each); KS on the first endpoints of each"}


def gate(out, q, plain_qs):
    """The headline target's conditions, stated as booleans: per-reaction
    mean and std within 1% of range of plain ACHR (the reference sampler), drift <= 1e-9,
    and every sampler mixed (max split-R-hat < 1.01 and < 1.1)."""
    me = out.get("moment_err_vs_plain", {})
    rhats = [q["max_rhat"]] + [p["max_rhat"] for p in plain_qs]
    g = {"mean_le_1pct_of_range": bool(me.get("mean_frac_max", np.inf) <= 0.01),
         "std_le_1pct_of_range": bool(me.get("std_frac_max", np.inf) <= 0.01),
         "drift_le_1e-9": bool(q["drift"] <= 1e-9),
         "max_rhat_learned": q["max_rhat"], "max_rhat_plain": [p["max_rhat"] for p in plain_qs],
         "rhat_lt_1.1": bool(max(rhats) < 1.1), "rhat_lt_1.01": bool(max(rhats) < 1.01),
         "speedup_ge_10x": bool(out.get("speedup_vs_plain", 0) >= 10.0),
         "speedup_ge_10x_vs_projected_warmup": bool(out.get("speedup_vs_plain_projected_warmup", 0) >= 10.0),
         "speedup_ge_10x_incl_training": bool(out.get("speedup_vs_plain_incl_training", 0) >= 10.0)}
    g["all"] = bool(g["mean_le_1pct_of_range"] and g["std_le_1pct_of_range"] and g["drift_le_1e-9"]
                    and g["rhat_lt_1.1"] and g["speedup_ge_10x"])
    g["reference_mixed"] = bool(max(rhats) < 1.1)
    g["note"] = ("The moment comparison is against plain ACHR, which is itself unmixed at this length when "
                 "reference_mixed is false: min-ESS then measures the chains' transient motion from their start, and "
                 "moment differences compare two transients, not an error against the uniform law. No sampler "
                 "run here mixes on the 10,600-reaction model within a feasible run.")
    return g


def run_putida(args):
    """The reference's own model end to end through the two command lines, as
    its paper times it (paper/paper.md:71-73: wall time with the null space,
    1,000 steps per point): createWarmupPts P_Putida.mps 2120 (the 2n flux-
    variability LPs on the GPU), then ACHRCuda P_Putida.mps <warmup> 1 1000
    1000 and 1 10000 1000 (--lossless for bit-exact output). Reports each
    run's wall time and chain-steps/s, the drift |S x| of the samples, and a
    16-chain file checked bit for bit against the oracle (with the null
    space the library computes, handed to both)."""
    import subprocess
    import tempfile
    import achr
    import scipy.sparse as sp
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    mps = os.path.join(ROOT, "tests", "golden", "P_Putida.mps")
    work = tempfile.mkdtemp(prefix="putida_")
    model = achr.Model.read_mps(mps)
    n, m_, nnz, _ = model.dims
    rp, ci, va = model.csr()
    lb, ub = model.bounds()
    Ssp = sp.csr_matrix((va, ci, rp), shape=(m_, n))
    src = os.path.join(work, "P_Putida.mps")
    with open(mps, "rb") as f, open(src, "wb") as g:
        g.write(f.read())
    t = time.perf_counter()
    r = subprocess.run([achr.WARMUP_CLI_PATH, src, str(2 * n)], cwd=work, capture_output=True, text=True)
    t_warm = time.perf_counter() - t
    if r.returncode:
        raise RuntimeError(r.stderr)
    conv_line = [l for l in r.stdout.splitlines() if "converged" in l]
    log(f"createWarmupPts: {t_warm:.1f}s; {conv_line}")
    wcsv = os.path.join(work, f"P_Putida{2 * n}warmup.csv")
    runs = {}
    for P in (1000, 10000):
        t = time.perf_counter()
        r = subprocess.run([achr.CLI_PATH, src, wcsv, "1", str(P), "1000", "--lossless"], cwd=work,
                           capture_output=True, text=True)
        wall = time.perf_counter() - t
        if r.returncode:
            raise RuntimeError(r.stderr)
        pts = achr.read_points_bin(os.path.join(work, "File0.bin"))
        lanes_line = [l for l in r.stdout.splitlines() if "threads per chain" in l]

        def cli_seconds(tag):
            # the command line's own clock: "<tag> done in X seconds." from its start
            hit = [l for l in r.stdout.splitlines() if l.startswith(tag + " done in")]
            return float(hit[0].split()[-2]) if hit else None
        runs[P] = {"points": P, "steps": 1000, "wall_s": wall, "chain_steps_per_s": P * 1000 / wall,
                   "cli_setup_s": cli_seconds("Null space"), "cli_total_s": cli_seconds("Sampling"),
                   "drift_inf": float(np.abs(Ssp @ pts).max()),
                   "in_bounds": bool((pts >= lb[:, None] - 1e-12).all() and (pts <= ub[:, None] + 1e-12).all()),
                   "cli": lanes_line[0] if lanes_line else ""}
        log(f"ACHRCuda 1 x {P} x 1000: {runs[P]}")
    # 16 chains against the oracle, the null space handed to both
    model.compute_nullspace()
    N = model.nullspace()
    npath = os.path.join(work, "N.bin")
    np.asfortranarray(N).ravel(order="F").tofile(npath)
    r = subprocess.run([achr.CLI_PATH, src, wcsv, "1", "16", "1000", "--lossless"], cwd=work,
                       env=dict(os.environ, ACHR_NULLSPACE=npath), capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr)
    g = achr.read_points_bin(os.path.join(work, "File0.bin"))
    lanes = int([l for l in r.stdout.splitlines() if "threads per chain" in l][0].split(",")[1].split()[0])
    Wc = np.loadtxt(wcsv, delimiter=",")
    threads = min(16, os.cpu_count() or 1)
    o, ost, _ = O.sample(warm=Wc, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=m_, P=16, S=1000, seed=0,
                         lanes=lanes, threads=threads)
    # the CPU baseline: the oracle over a bounded sample of the same file's chains
    P_cpu = args.cpu_chains if args.cpu_chains != 32 else 2048
    t = time.perf_counter()
    O.sample(warm=Wc, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=m_, P=P_cpu, S=1000, seed=0,
             lanes=lanes, threads=threads)
    t_cpu = time.perf_counter() - t
    parity = {"chains": 16, "steps": 1000, "lanes": lanes, "bit_equal": bool(np.array_equal(g, o)),
              "max_abs_diff": float(np.abs(g - o).max()), "reprojections_per_chain": float(ost[:, 1].mean())}
    log(f"oracle parity: {parity}")
    big = runs[10000]
    out = {"n_gpus": 1, "steps": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "the reference's P. putida model (VFWarmup/lib/P_Putida.mps)",
           "metric": "chain-steps/s, ACHRCuda end to end (null space + sampling + File0), P. putida",
           "value": big["chain_steps_per_s"], "unit": "chain-steps/s", "ms_per_step": big["wall_s"] * 1e3,
           "config": {"workload": f"putida: {m_} x {n} model, createWarmupPts {2 * n} then ACHRCuda 1 x P x 1000",
                      "reactions": n, "metabolites": m_},
           "warmup_generation": {"lps": 2 * n, "wall_s": t_warm, "report": conv_line},
           "runs": {str(k): v for k, v in runs.items()}, "oracle_parity": parity,
           "paper_context": "Tesla K40, 30,000 warmup points: 17 s (1,000 points) and 138 s (10,000 points), "
                            "paper/paper.md:71-73; other hardware, context only",
           "cpu_baseline": {"value": P_cpu * 1000 / t_cpu, "unit": "chain-steps/s", "cores": threads, "kind": "port",
                            "sample": f"oracle (C, OpenMP) {P_cpu} chains x 1,000 steps of the same model and "
                                      f"warmup in {t_cpu:.1f}s (sampling only: no null space, no file I/O)"}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed files (bench steps)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed files (short: 20 chain steps)")
    ap.add_argument("--workload", choices=["genome", "ecoli", "batch", "precision", "gemm", "convergence", "warmup",
                                               "putida"],
                    default="genome")
    ap.add_argument("--warmup-iters", type=int, default=8000, help="warmup workload: PDHG iteration cap")
    ap.add_argument("--cpu-lps", type=int, default=1, help="warmup workload: LPs HiGHS times (about 1 min each)")
    ap.add_argument("--ref-steps", type=int, default=20000, help="convergence workload: steps of the reference run")
    ap.add_argument("--models", type=int, default=64, help="batch workload: number of models")
    ap.add_argument("--probe-steps", type=int, default=20, help="batch workload: autotuner probe length")
    ap.add_argument("--no-graph", action="store_true", help="batch workload: launch the segments eagerly")
    ap.add_argument("--sampler", choices=["learned", "plain"], default="learned")
    ap.add_argument("--chains", type=int, default=None)
    ap.add_argument("--chain-steps", type=int, default=None)
    ap.add_argument("--trace-chains", type=int, default=64)
    ap.add_argument("--trace-thin", type=int, default=50)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--proj", choices=list(PROJ), default="fp64",
                    help="arithmetic of the batched projection rounds of plain ACHR")
    ap.add_argument("--fp64-every", type=int, default=10)
    ap.add_argument("--train-steps", type=int, default=1000, help="plain-ACHR steps that produce the training set")
    ap.add_argument("--train-thin", type=int, default=100)
    ap.add_argument("--train-fp64-every", type=int, default=0,
                    help="every n-th projection round of the training run in FP64 (0: none; the residual-form "
                         "rounds recompute S x in FP64 each time, so they do not let the residual grow)")
    ap.add_argument("--rank", type=int, default=512)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--check-every", type=int, default=100, help="residual check period of the learned sampler")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain FP64 ACHR comparison run")
    ap.add_argument("--no-projected-plain", action="store_true",
                    help="skip the plain FP64 run from the projected warmup")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--save-inputs", default=None, help="genome workload: save model, warmup and bank (.npy) here")
    ap.add_argument("--load-inputs", default=None,
                    help="genome workload: load them from here instead (no solver or training in the process)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (a known-traffic kernel for counters)")
    ap.add_argument("--cpu-chains", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (chains shard independently across "
              f"GPUs, see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
    if args.workload == "precision":
        run_precision(args)
        return
    if args.workload == "gemm":
        run_gemm(args)
        return
    if args.workload == "convergence":
        run_convergence(args)
        return
    if args.workload == "warmup":
        run_warmup(args)
        return
    if args.workload == "putida":
        run_putida(args)
        return
    if args.workload == "batch":
        if args.cpu_chains == 32 and args.cpu_steps == 300:
            args.cpu_chains, args.cpu_steps = 16, 40
        run_batch(args)
        return

    import achr
    from achr import stats as st

    if args.workload == "genome":
        if args.load_inputs:
            m, W = load_inputs(args.load_inputs)
        else:
            m, W = genome_model()
            if args.save_inputs:
                save_inputs(args.save_inputs, m, W)
        P, S = 20000, 5000
    else:
        m, W = ecoli_model()
        P, S = 500, 1000
    P = args.chains or P
    S = args.chain_steps or S
    N = m.N
    n, d = m.n, N.shape[1]
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    tc = min(args.trace_chains, P)
    out = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    cfg = {"workload": f"{args.workload}: {m.m}x{m.n} model, {W.shape[1]} warmup points (%f), {P} chains x {S} "
                       f"steps", "chains": P, "chain_steps": S, "reactions": n, "metabolites": m.m, "null_dim": d}
    plain_q = None
    if args.sampler == "plain" or not args.no_compare:
        t = time.time()
        opts = achr.Options(lanes=args.lanes, trace_chains=tc, trace_thin=args.trace_thin,
                            proj_prec=PROJ[args.proj], fp64_every=args.fp64_every,
                            reproject_mode=2 if args.proj != "fp64" else 0)
        smp = achr.Sampler(model, W, opts)
        lanes, mode = smp.schedule()
        log(f"plain sampler ready in {time.time() - t:.1f}s: {lanes} threads/chain, reprojection mode {mode}")
        wall_p, tm_p = measure(smp, P, S, args.steps if args.sampler == "plain" else 1, args.warmup)
        plain_q = sample_quality(m, smp, S, tc, P)
        plain_value = plain_q["min_ess"] * (P / plain_q["traced"]) / wall_p
        plain = {"min_ess_per_s": plain_value, "s_per_file": wall_p, "chain_steps_per_s": P * S / wall_p,
                 "min_ess_traced": plain_q["min_ess"], "drift_inf": plain_q["drift"],
                 "max_rhat": plain_q["max_rhat"], "kernel_ms": tm_p, "projection": args.proj,
                 "reprojections": int(plain_q["stats"][:, 1].sum()), "checks": int(plain_q["stats"][:, 2].sum()),
                 "warmup_set": "as read (the reference's behaviour: \"%f\" points, every check reprojects)",
                 "check_kernel": check_kernel_rate(n, plain_q, tm_p)}
        log(f"plain: {plain}")
        del smp
    proj_q = None
    if args.sampler == "learned" and not args.no_compare and not args.no_projected_plain:
        # the same plain FP64 sampler from the warmup moved onto {S x = 0,
        # fixed reactions at their bound}: the reprojection the "%f" file forces
        # is no longer part of the comparison
        t = time.time()
        smp = achr.Sampler(model, W, achr.Options(lanes=args.lanes, trace_chains=tc, trace_thin=args.trace_thin,
                                                  reproject_mode=2))
        smp.project_warmup()
        t_proj = time.time() - t
        wall_q, tm_q = measure(smp, P, S, 1, 0)
        proj_q = sample_quality(m, smp, S, tc, P)
        plain_proj = {"min_ess_per_s": proj_q["min_ess"] * (P / proj_q["traced"]) / wall_q, "s_per_file": wall_q,
                      "chain_steps_per_s": P * S / wall_q, "min_ess_traced": proj_q["min_ess"],
                      "drift_inf": proj_q["drift"], "max_rhat": proj_q["max_rhat"], "kernel_ms": tm_q,
                      "reprojections": int(proj_q["stats"][:, 1].sum()), "checks": int(proj_q["stats"][:, 2].sum()),
                      "warmup_set": f"projected onto S x = 0 with the fixed reactions at their bound ({t_proj:.1f}s, "
                                f"untimed), FP64 projection rounds for the checks that still flag",
                      "check_kernel": check_kernel_rate(n, proj_q, tm_q)}
        log(f"plain, projected warmup: {plain_proj}")
        del smp
    if args.sampler == "plain":
        out.update({"metric": f"min-ESS/sec over reactions (plain ACHR, projection {args.proj})",
                    "value": plain_value, "unit": "min-ESS/s", "ms_per_step": wall_p * 1e3})
        cfg.update({"sampler": "plain ACHR", "lanes": lanes, "reproject_mode": mode})
        out["config"] = cfg
        out["roofline"] = roofline_for(args.proj, m, n, d, P, S, plain_q, tm_p)
        out.update({k: v for k, v in plain.items() if k != "kernel_ms"})
        ess_per_step_ref = plain_q["ess_per_step"]
    else:
        bank_file = os.path.join(args.load_inputs, "bank.npy") if args.load_inputs else None
        if bank_file and os.path.exists(bank_file):
            # a bank learned and saved by an earlier run (--save-inputs)
            # (FP64 projection rounds: no pseudo-inverse to build, so no rocSOLVER in the process; bank chains
            # start on S x = 0 and never reproject, so the rounds' precision does not matter)
            smp = achr.Sampler(model, W, achr.Options(lanes=args.lanes, reproject_mode=2))
            smp.set_directions(np.load(bank_file))
            smp.use_directions(True)
            t_train_run = t_learn = 0.0
            rep = {"bank": "loaded"}
        else:
            # 1. training set: plain ACHR (residual BF16x3 projection) with every chain traced
            t = time.time()
            topts = achr.Options(lanes=args.lanes, trace_chains=P, trace_thin=args.train_thin,
                                 proj_prec=PROJ["res_bf16x3"], reproject_mode=2,
                                 fp64_every=args.train_fp64_every if args.train_fp64_every > 0 else 1 << 30)
            smp = achr.Sampler(model, W, topts)
            smp.run(P, args.train_steps, seed=99)
            smp.timing()
            t_train_run = time.time() - t
            import scipy.sparse as sp
            train_drift = float(np.abs(sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, m.n))
                                       @ smp.points()[:, :1024]).max())
            # 2. rounding map + network + direction bank
            t = time.time()
            rep = smp.learn_directions(rank=args.rank, width=512, iters=args.iters, batch=2048, lr=1e-3, seed=7)
            t_learn = time.time() - t
            rep["train_drift_inf"] = train_drift
            bank = smp.directions()
            if args.save_inputs:
                np.save(os.path.join(args.save_inputs, "bank.npy"), bank)
            # The learned file runs on FP64 projection rounds, as a loaded bank does: bank chains start
            # on S x = 0 and never reproject, so the rounds' precision changes no result, and with FP64
            # rounds a chain runs through up to 10 unflagged checks per launch instead of stopping at each
            # (reduced-precision rounds keep every chain's checks aligned with the rounds).
            del smp
            smp = achr.Sampler(model, W, achr.Options(lanes=args.lanes, reproject_mode=2))
            smp.set_directions(bank)
            smp.use_directions(True)
            del bank
        log(f"training run {t_train_run:.1f}s, learn {t_learn:.1f}s: {rep}")
        # 3. timed learned run
        smp.set_run_options(tc, args.trace_thin, args.check_every)
        lanes, mode = smp.schedule()
        wall, tm = measure(smp, P, S, args.steps, args.warmup)
        q = sample_quality(m, smp, S, tc, P)
        value = q["min_ess"] * (P / q["traced"]) / wall
        out.update({"metric": "min-ESS/sec over reactions (learned+batched vs plain ACHR)", "value": value,
                    "unit": "min-ESS/s", "ms_per_step": wall * 1e3})
        cfg.update({"sampler": "learned direction bank (3-layer MLP rounding map) + batched chains",
                    "lanes": lanes, "reproject_mode": mode, "check_every": args.check_every,
                    "bank_size": 2 * n, "rank": args.rank})
        out["config"] = cfg
        out.update({"chain_steps_per_s": P * S / wall, "min_ess_traced": q["min_ess"], "traced_chains": q["traced"],
                    "drift_inf": q["drift"], "max_rhat": q["max_rhat"],
                    "learn": {"train_run_s": t_train_run, "learn_s": t_learn,
                              "train_fp64_every": args.train_fp64_every, **rep},
                    "kernel_ms": tm})
        out["roofline"] = roofline_for("fp64", m, n, d, P, S, q, tm, bank=True)
        out["roofline"]["measured_copy_GBs"] = measure_copy_gbs()
        out["roofline"]["measured_read_GBs"] = measure_read_gbs()
        meas = PMC_TRAFFIC.get((n, P, args.check_every))
        if meas and out["roofline"]["kernel"] == "step_kernel":
            out["roofline"]["traffic"] = meas[0]
            out["roofline"]["traffic_note"] = (f"HBM bytes of one full launch, {10 * args.check_every} steps (10 "
                                               f"checks), of the learned step kernel (FETCH_SIZE x 2.000 + "
                                               f"WRITE_SIZE x 1.000, calibrated on a 1 GiB copy); algorithmic "
                                               f"{meas[1]:.4g} B for that launch; {meas[2]}")
        if plain_q is not None:
            out["plain_fp64"] = plain
            out["speedup_vs_plain"] = value / plain["min_ess_per_s"]
            out["moment_err_vs_plain"] = compare_endpoints(m, q["pts"], plain_q["pts"], P)
        if proj_q is not None:
            out["plain_fp64_projected_warmup"] = plain_proj
            out["speedup_vs_plain_projected_warmup"] = value / plain_proj["min_ess_per_s"]
            out["moment_err_vs_plain_projected_warmup"] = compare_endpoints(m, q["pts"], proj_q["pts"], P)
        # the charge of the training (plain run + network + bank) for the
        # config's one 20,000 x 5,000 file
        t_train = t_train_run + t_learn
        out["value_incl_training"] = q["min_ess"] * (P / q["traced"]) / (wall + t_train)
        if plain_q is not None:
            out["speedup_vs_plain_incl_training"] = out["value_incl_training"] / plain["min_ess_per_s"]
        if proj_q is not None:
            out["speedup_vs_plain_projected_warmup_incl_training"] = (out["value_incl_training"]
                                                                      / plain_proj["min_ess_per_s"])
        out["gate"] = gate(out, q, [x for x in (plain_q, proj_q) if x is not None])
        ess_per_step_ref = plain_q["ess_per_step"] if plain_q is not None else None
    if not args.no_cpu and ess_per_step_ref is not None:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(m, W, N, args.cpu_chains, args.cpu_steps, threads, ess_per_step_ref)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
