"""MGMC sampling benchmark on one MI355X.

Workload: the 128^3 3D shifted-Laplace posterior (kappa^2 = 10, Dirichlet
boundaries) with 32 top-hat ball measurements (radius 0.05, centres
U[0.1, 0.9]^3, variance 1e-4, y ~ N(0, 1)), MGMC V-cycle with SSOR 1+1 Gibbs
smoothing (omega = 1), 5 levels (coarsest 7^3 = 343 unknowns, Cholesky).
A step is one MGMC cycle over the whole batch of chains.

value = IACT-adjusted independent samples per second of the batched FP32 form
(64 chains): K * steps / (t_wall * tau_int). t_wall covers exactly `steps`
replays of the captured cycle (observables recorded into device memory; the
buffer and the graph exist before the clock starts). tau_int is the largest
Sokal estimate over the probe vertices, taken from a fixed run of --tau-steps
cycles per chain after the timed one, so that it does not depend on --steps.
The plain single-chain FP64 form runs --fp64-steps cycles (timed, and its
IACT sample) in the same job and is reported beside it, with the ratio, and
both forms' probe covariances are compared with the exact posterior.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "IACT-adjusted indep. samples/s at 128^3 posterior: batched FP32 vs plain FP64"
HBM_PEAK_GBS = 8000.0


def problem(n, m, seed):
    import multigridmc_amd as mg
    lat = mg.Lattice(n, n, n)
    centres, y, var = mg.seeded_ball_measurements(lat, m, 0.05, 1e-4, seed)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.05, var)
    # f = Q mean_exact = B Sigma^{-1} y for a zero prior mean (src/measurements/sampling_time.cc:19-28)
    f = op.get_B() @ (y / op.get_Sigma())
    return op, f, centres, y, var


def tau_max(obs):
    """Largest and median Sokal tau_int over the probes of obs[sample, chain, probe]."""
    import multigridmc_amd as mg
    taus = [mg.tau_int_chains(obs[:, :, p])[0] for p in range(obs.shape[2])]
    return float(max(taus)), float(np.median(taus))


def tau_max_window(obs, window=10):
    """Largest fixed-window estimate over the probes (the reference's
    Statistics::tau_int, src/auxilliary/statistics.cc:65-79), each probe's
    value the mean over its chains."""
    import multigridmc_amd as mg
    return float(max(np.mean([mg.tau_int_window(obs[:, c, p], window) for c in range(obs.shape[1])])
                     for p in range(obs.shape[2])))


def covariance_error(obs, mean_exact, cov_exact, tau):
    """Error of the probe statistics of obs[sample, chain, probe] against the
    exact posterior: relative Frobenius error of the probe covariance, its
    Monte Carlo noise floor (for Gaussian samples Var C_ij = (C_ii C_jj +
    C_ij^2) / n_eff, n_eff = samples / tau), the RMS and largest relative
    variance errors beside their noise, and the largest mean error in
    posterior standard deviations."""
    z = obs.reshape(-1, obs.shape[2])
    n = z.shape[0]
    mu = z.mean(axis=0)
    d = z - mu
    cov = d.T @ d / n
    norm = np.linalg.norm(cov_exact)
    err = float(np.linalg.norm(cov - cov_exact) / norm)
    dg = np.diag(cov_exact)
    floor = float(np.sqrt(np.sum(np.outer(dg, dg) + cov_exact ** 2) * tau / n) / norm)
    rel_var = (np.diag(cov) - dg) / dg
    return {"rel_frobenius": err, "mc_noise_floor": floor, "samples_per_probe": n,
            # relative variance errors: Monte Carlo noise sqrt(2 tau / n) each
            "rms_rel_variance_error": float(np.sqrt(np.mean(rel_var ** 2))),
            "rel_variance_mc_noise": float(np.sqrt(2.0 * tau / n)),
            "max_rel_variance_error": float(np.max(np.abs(rel_var))),
            "max_mean_error_in_sd": float(np.max(np.abs(mu - mean_exact) / np.sqrt(dg)))}


def new_sampler(op, f, params, nchains, precision, probes, seed, tuning=None):
    import multigridmc_amd as mg
    s = mg.MultigridMCSampler(op, params, nchains=nchains, precision=precision, seed=seed)
    for k, v in (tuning or {}).items():
        s.set_tuning(k, v)
    s.fix_rhs(f)
    s.set_state(np.zeros((nchains, op.lattice.Nvertex)))
    s.set_probes(probes)
    return s


def timed_run(s, warmup, steps, rows, use_graph):
    """Warm-up and `steps` timed cycles, both recording the probe observables
    into one device buffer of `rows` rows, so that the buffer exists and the
    cycle graph is captured before the clock starts; the timed region holds
    graph replays only, bracketed by device synchronisations."""
    import torch
    obs = torch.empty((rows, s.nchains, s.nprobe), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s.run_dev(max(warmup, 1), obs.data_ptr(), use_graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run_dev(steps, obs.data_ptr(), use_graph)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return t, obs[:steps].cpu().numpy()


def repeat_runs(s, steps, n, use_graph):
    """`n` more timed runs of `steps` cycles each (one device buffer; one
    untimed cycle first captures the graph for it): ms per cycle of each."""
    import torch
    obs = torch.empty((max(steps, 1), s.nchains, s.nprobe), dtype=torch.float64, device="cuda")
    s.run_dev(1, obs.data_ptr(), use_graph)
    out = []
    for _ in range(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run_dev(steps, obs.data_ptr(), use_graph)
        torch.cuda.synchronize()
        out.append(1e3 * (time.perf_counter() - t0) / steps)
    return out


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n, m, seed, budget_s, nthreads):
    """The oracle (FP64 restatement of the reference: lexicographic SSOR Gibbs
    sweeps with dense SMW matrices, CLCG noise) on the host cores."""
    from oracle import pyoracle as po
    import multigridmc_amd as mg
    lat = mg.Lattice(n, n, n)
    centres, y, var = mg.seeded_ball_measurements(lat, m, 0.05, 1e-4, seed)
    ora = po.Operator.fd(3, (n, n, n), 10.0)
    ora.add_measurements(centres, 0.05, var)
    omg = po.Multigrid(ora, 5, ordering=po.LEX)
    f = ora.csr(1) @ (y / var)
    N = ora.N
    # single thread: ms per sample (src/measurements/sampling_time.cc:36-44)
    x, _, t1 = omg.chains(f, np.zeros((1, N)), 1, nthreads=1)
    ns1 = max(1, min(20, int(0.4 * budget_s / max(t1, 1e-3))))
    x, _, t1 = omg.chains(f, x, ns1, nthreads=1)
    ms_single = 1e3 * t1 / ns1
    # all cores: one chain per thread (src/measurements/convergence.cc:29-63)
    ns = max(1, min(10, int(0.5 * budget_s / max(ms_single * 1e-3 * 1.5, 1e-3))))
    _, _, tm = omg.chains(f, np.zeros((nthreads, N)), ns, nthreads=nthreads)
    rate = nthreads * ns / tm
    return {
        "value": rate,
        "unit": "chain-samples/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"oracle MGMC (lexicographic SSOR 1+1, dense SMW, FP64) at {n}^3 with {m} measurements: "
                  f"{nthreads} chains x {ns} cycles on {nthreads} threads; single thread {ns1} cycles",
        "ms_per_sample_1thread": ms_single,
    }


def mixed_batch(args):
    print(json.dumps(mixed_batch_run(args.steps, args.warmup, not args.no_graph, args.persistent)), flush=True)


def mixed_batch_run(steps, warmup, use_graph, persistent, forms=("merged", "merged_persistent", "streams")):
    """The uneven batch of the batched-chains configuration: 16 posterior
    instances on lattices drawn from {32^3, 48^3, 64^3, 96^3, 128^3} (seeded),
    8-32 ball measurements each, 4 FP32 chains per instance, V(1,1) SSOR down
    to at most 8 cells per direction. Each form runs --warmup cycles, then
    --steps timed cycles (graph replays, device synchronisation on both
    sides); the line reports chain-samples per second of the whole batch."""
    import torch
    import multigridmc_amd as mg
    from multigridmc_amd import measurements as ms
    from multigridmc_amd.batch import SamplerBatch
    rng = np.random.default_rng(2024)
    inst = []
    for k in range(16):
        n = int(rng.choice([32, 48, 64, 96, 128]))
        lat = mg.Lattice(n, n, n)
        m = int(rng.integers(8, 33))
        centres, y, var = mg.seeded_ball_measurements(lat, m, 0.05, 1e-4, 2024 + k)
        op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.05, var)
        inst.append(dict(op=op, params=mg.MultigridParameters(nlevel=ms._nlevel(op)), nchains=4,
                         f=op.get_B() @ (y / var), observables=[(np.array([op.lattice.Nvertex // 2]), np.array([1.0]))],
                         seed=77 + k))
    chains = sum(i["nchains"] for i in inst)
    out = {"workload": "mixed batch of 16 instances, lattices " +
                       ",".join(str(i["op"].lattice.n[0]) for i in inst) + ", 4 FP32 chains each",
           "chains": chains, "steps": steps, "warmup": warmup, "graph": use_graph}
    for name, mode, pers in (("merged", "merged", 0), ("merged_persistent", "merged", persistent),
                             ("streams", "streams", 0)):
        if name not in forms:
            continue
        b = SamplerBatch(inst, mode=mode, persistent=pers)
        bufs = [torch.empty((max(steps, warmup), s.nchains, 1), dtype=torch.float64, device="cuda")
                for s in b.samplers]
        torch.cuda.synchronize()
        b.run_dev(warmup, bufs, use_graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run_dev(steps, bufs, use_graph)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        st = b.stats()
        out[name] = {"ms_per_cycle": 1e3 * t / steps, "chain_samples_per_s": chains * steps / t}
        if st:
            out[name]["launches_per_cycle"] = st[0]
            out[name]["separate_launches_per_cycle"] = st[1]
        b.close()
        del bufs
    if "merged" in out and "streams" in out:
        out["speedup_merged_vs_streams"] = out["streams"]["ms_per_cycle"] / out["merged"]["ms_per_cycle"]
    return out


def config3(args):
    """The mixed-precision configuration at scale: 256^3, -div(a grad) +
    kappa^2 with a piecewise-constant a in [1, 10] on a seeded 8^3 block grid,
    64 balls of radius 0.03 (variance 1e-4), 6 levels. (1) The loop solver
    with FP32 cycles and FP64 residuals to 1e-10 relative, against the same
    solve with FP64 cycles; (2) 32 FP32 chains x --config3-samples cycles at
    16 seeded probe vertices, against the exact posterior moments from FP64
    multigrid solves (tolerance 1e-11). Prints one JSON line."""
    import multigridmc_amd as mg
    from multigridmc_amd import measurements as ms
    from multigridmc_amd.operators import block_diffusion
    lat = mg.Lattice(256, 256, 256)
    a = block_diffusion(lat, 8, 1.0, 10.0, 33)
    centres, y, var = mg.seeded_ball_measurements(lat, 64, 0.03, 1e-4, 31)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0, diffusion=a), centres, 0.03, var)
    N = lat.Nvertex
    p = mg.MultigridParameters(nlevel=6)
    out = {"workload": "256^3, block diffusion a in [1,10] on 8^3 blocks, 64 balls r=0.03 var=1e-4, 6 levels"}
    b = np.random.default_rng(3).standard_normal(N)
    sols = {}
    for precision in ("fp32", "fp64"):
        solver = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p, precision=precision), rtol=1e-10, atol=1e300,
                               maxiter=100)
        t0 = time.perf_counter()
        sols[precision] = solver.apply(b)
        t = time.perf_counter() - t0
        h = solver.history
        out[f"solve_{precision}_cycles"] = {"iterations": len(h) - 1, "final_rel_residual": float(h[-1] / h[0]),
                                            "seconds": t}
        print(f"config3: solve with {precision} cycles done", file=sys.stderr, flush=True)
    out["solve_fp32_cycles_vs_fp64_rel_error"] = float(np.linalg.norm(sols["fp32"] - sols["fp64"]) /
                                                       np.linalg.norm(sols["fp64"]))
    probes = np.sort(np.random.default_rng(6).choice(N, size=16, replace=False))
    obsv = [(np.array([q]), np.array([1.0])) for q in probes]
    mean_exact, cov_exact = ms.observed_moments(op, np.zeros(N), y, obsv, params=p, rtol=1e-11)
    print("config3: exact probe moments done", file=sys.stderr, flush=True)
    K, ns = 32, args.config3_samples
    s = new_sampler(op, op.get_B() @ (y / op.get_Sigma()), p, K, "fp32", probes, 19)
    s.run(50, record=False)
    chunks, done, t = [], 0, 0.0
    while done < ns:  # in parts, with a progress line after each
        m = min(500, ns - done)
        t0 = time.perf_counter()
        chunks.append(s.run(m))
        t += time.perf_counter() - t0
        done += m
        print(f"config3: {done} of {ns} cycles sampled", file=sys.stderr, flush=True)
    obs = np.concatenate(chunks)
    tau = max(mg.tau_int_chains(obs[:, :, q], mean=mean_exact[q])[0] for q in range(len(probes)))
    out["sampling"] = {"chains": K, "samples_per_chain": ns, "probes": len(probes), "ms_per_cycle": 1e3 * t / ns,
                       "tau_int_max": tau,
                       "covariance_error_vs_fp64_exact": covariance_error(obs, mean_exact, cov_exact, tau)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--m", type=int, default=32)
    ap.add_argument("--fp64-steps", type=int, default=20000,
                    help="cycles of the single-chain FP64 run (timed; also its IACT sample)")
    ap.add_argument("--tau-steps", type=int, default=400, help="cycles per chain of the FP32 IACT estimate")
    ap.add_argument("--stat-steps", type=int, default=2000,
                    help="cycles per chain of the FP32 statistics run (covariance error; >= --tau-steps)")
    ap.add_argument("--no-exact", action="store_true", help="skip the exact probe moments (FP64 MG solves)")
    ap.add_argument("--probes", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--repeats", type=int, default=5,
                    help="further timed runs of each form after the measurement (median reported)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-batch timing in the metric line")
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of hipGraph replay")
    ap.add_argument("--colour-phases", action="store_true",
                    help="sweep level 0 with one launch per colour instead of the single-pass kernel")
    ap.add_argument("--mixed-batch", action="store_true",
                    help="only time the mixed batch of 16 instances (32^3..128^3, 8-32 balls, 4 FP32 chains each): "
                         "merged launches (plain and persistent) against one stream per instance")
    ap.add_argument("--config3", action="store_true",
                    help="only run the 256^3 block-diffusion configuration once: the mixed-precision solve and "
                         "32 FP32 chains x --config3-samples with probe errors against the FP64 exact posterior")
    ap.add_argument("--config3-samples", type=int, default=5000)
    ap.add_argument("--persistent", type=int, default=1,
                    help="workgroups per CU of the persistent merged launches (--mixed-batch)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        c = torch.empty_like(a)

        def rate(op, nbytes):
            for _ in range(3):
                op()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                op()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 20
            return ms, nbytes / (ms * 1e-3) / 1e9

        # the streaming rates this device reaches outside the sampler, timed
        # with HIP events: read + write (copy), two reads + one write (the
        # level-0 sweep's mix: x and f read, x written), read only
        ms, gbps = rate(lambda: b.copy_(a), 8 * n)
        ms2, gbps2 = rate(lambda: torch.add(a, b, alpha=2.0, out=c), 12 * n)
        ms3, gbps3 = rate(lambda: torch.sum(a), 4 * n)
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copy_ms": ms,
                          "copy_GBps_read_plus_write": gbps,
                          "two_reads_one_write_ms": ms2, "two_reads_one_write_GBps": gbps2,
                          "read_only_sum_ms": ms3, "read_only_GBps": gbps3}))
        return

    if args.mixed_batch:
        mixed_batch(args)
        return
    if args.config3:
        config3(args)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the chains would shard "
              f"across GPUs, one process per GPU, with no per-step communication; no multi-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)

    import torch  # before the library: one HIP runtime in the process
    import multigridmc_amd as mg
    from multigridmc_amd import measurements

    n, K = args.n, args.chains
    op, f, centres, y, var = problem(n, args.m, args.seed)
    N = op.lattice.Nvertex
    rng = np.random.default_rng(args.seed + 1)
    probes = np.sort(rng.choice(N, size=args.probes, replace=False))
    params = mg.MultigridParameters(nlevel=5, smoother="SSOR", npresmooth=1, npostsmooth=1, omega=1.0, cycle=1,
                                    coarse_solver="Cholesky")
    g = not args.no_graph
    tuning = {"fused_sweep": 0} if args.colour_phases else {}

    # batched FP32 (the metric's form): exactly --steps timed cycles
    s32 = new_sampler(op, f, params, K, "fp32", probes, 1, tuning)
    t32, _ = timed_run(s32, args.warmup, args.steps, max(args.warmup, args.steps, 1), g)

    # IACT and covariance on a fixed run that does not depend on --steps: the
    # chains continue from the timed run (in equilibrium after the warm-up)
    obs32 = s32.run(args.stat_steps, use_graph=g)
    tau32, tau32_med = tau_max(obs32[:args.tau_steps])
    tau32_long, _ = tau_max(obs32)
    ess32 = K * args.steps / (t32 * tau32)

    # dominant kernel: the level-0 sweep, timed with HIP events on the handle's
    # stream around each of its launches over 10 eager cycles, enqueued behind
    # 20 graph-replayed cycles so that the host has queued every timed launch
    # before the device reaches it (a host stall would otherwise fall inside
    # an event pair)
    pbuf = torch.empty((20, K, s32.nprobe), dtype=torch.float64, device="cuda")
    s32.run_dev(1, pbuf.data_ptr(), g)
    torch.cuda.synchronize()
    s32.run_dev(20, pbuf.data_ptr(), g)
    s32.profile(1)
    s32.run_dev(10, pbuf.data_ptr(), False)
    ms_sweep, nlaunch = s32.profile_result()
    s32.profile(0)
    w = 4
    Nc = (n // 2 - 1) ** 3
    if args.colour_phases:
        kernel = "k_sweep<float,3,star,philox> (level-0 red-black colour phase)"
        # one colour: read x (all, for the neighbours), f and write x of half the vertices
        bytes_per_launch = w * N * K + w * N * K // 2 + w * N * K // 2
        bytes_shared_f = w * N * K + w * N * K // 2 + w * N
    else:
        kernel = "k_sweep_rb3<float,philox> (level-0 red-black sweep, one pass)"
        # per directional sweep and chain 3 w N_0 (read x, read f, write x); the
        # first post-smoothing launch of the cycle (one of 2 (npre + npost)) also
        # reads the level-1 correction it adds while loading
        nlaunch_cycle = 2 * (params.npresmooth + params.npostsmooth)
        bytes_per_launch = 3 * w * N * K + w * Nc * K / nlaunch_cycle
        bytes_shared_f = 2 * w * N * K + w * N + w * Nc * K / nlaunch_cycle
    achieved = bytes_per_launch / (ms_sweep * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", ("sweep_l0_fp32_128cube_k64_traffic.json" if args.colour_phases
                                         else "sweep_rb3_l0_fp32_128cube_k64_traffic.json"))
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    # plain single-chain FP64: its timed run is the fixed-length IACT run
    s64 = new_sampler(op, f, params, 1, "fp64", probes, 2, tuning)
    t64, obs64 = timed_run(s64, args.warmup, args.fp64_steps, max(args.warmup, args.fp64_steps), g)
    tau64, tau64_med = tau_max(obs64)
    ess64 = args.fp64_steps / (t64 * tau64)

    # repeats: --repeats more timed runs of each form (FP64: at most 2,000
    # cycles each), reported as their median beside the single timed run
    repeats = None
    if args.repeats > 0:
        r32 = repeat_runs(s32, args.steps, args.repeats, g)
        n64 = min(args.fp64_steps, 2000)
        r64 = repeat_runs(s64, n64, args.repeats, g)
        m32, m64 = float(np.median(r32)), float(np.median(r64))
        repeats = {"n": args.repeats, "fp32_steps_each": args.steps, "fp32_ms_per_step": r32,
                   "fp32_ms_per_step_median": m32, "fp32_value_at_median": 1e3 * K / (m32 * tau32),
                   "fp64_steps_each": n64, "fp64_ms_per_step": r64, "fp64_ms_per_step_median": m64,
                   "fp64_value_at_median": 1e3 / (m64 * tau64)}

    # exact probe moments from FP64 multigrid solves (the posterior mean and
    # b_p^T Q^{-1} b_q), against which both forms' samples are checked
    cov_exact = None
    if not args.no_exact:
        obsv = [(np.array([p]), np.array([1.0])) for p in probes]
        mean_exact, cov_exact = measurements.observed_moments(op, np.zeros(N), y, obsv)

    out = {
        "metric": METRIC,
        "value": ess32,
        "unit": "indep. samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 shifted-Laplace posterior, kappa^2=10, {args.m} ball measurements r=0.05 var=1e-4, "
                        f"MGMC V(1,1) SSOR omega=1, 5 levels, Cholesky 343 coarse, {K} chains FP32",
            "chains": K,
            "lattice": [n, n, n],
            "measurements": args.m,
            "nlevel": 5,
            "probes": args.probes,
        },
        "tau_int": {"fp32_batched_max": tau32, "fp32_batched_median": tau32_med,
                    "fp32_samples_per_probe": K * args.tau_steps,
                    "fp32_batched_max_long_run": tau32_long, "fp32_long_run_samples_per_probe": K * args.stat_steps,
                    "fp64_single_max": tau64, "fp64_single_median": tau64_med,
                    "fp64_samples_per_probe": args.fp64_steps,
                    "estimator": "Sokal automatic window c=5, chain-averaged autocovariance about the pooled mean",
                    "reference_window10_fp32_max": tau_max_window(obs32[:args.tau_steps]),
                    "reference_window10_fp64_max": tau_max_window(obs64)},
        "fp64_single_chain": {"value": ess64, "unit": "indep. samples/s", "steps": args.fp64_steps,
                              "ms_per_step": 1e3 * t64 / args.fp64_steps},
        "speedup_vs_fp64_single": ess32 / ess64,
        "repeats": repeats,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel,
                     "avg_launch_ms": ms_sweep, "launches_timed": nlaunch,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "achieved_f_read_once": bytes_shared_f / (ms_sweep * 1e-3) / 1e9},
        "cpu_baseline": None,
    }
    if cov_exact is not None:
        out["covariance_error_vs_fp64_exact"] = {
            "fp32_batched": covariance_error(obs32, mean_exact, cov_exact, tau32_long),
            "fp64_single": covariance_error(obs64, mean_exact, cov_exact, tau64)}
    if not args.no_mixed:
        # the uneven part of the batched configuration beside the headline:
        # 16 instances of different sizes as merged launches against one
        # stream per instance (bench.py --mixed-batch for all forms)
        try:
            out["mixed_batch"] = mixed_batch_run(100, 10, g, 0, forms=("merged", "streams"))
        except Exception as e:  # the metric line must still print
            out["mixed_batch"] = {"error": str(e)}
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        try:
            out["cpu_baseline"] = cpu_baseline(n, args.m, args.seed, args.cpu_budget, nthreads)
            out["cpu_baseline"]["cpu_model"] = host_cpu_model()
        except Exception as e:  # the GPU line must still print
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
