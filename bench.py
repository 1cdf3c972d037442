"""Benchmark: MG-PCG time to rel-res 1e-8 and K.u GDOF/s on one MI355X.

Default workload: the ~125 M-DOF irregular domain (union of 24 seeded
ellipsoids in a 512^3 voxel box, fill 0.30; nodes of voxels below z = 8
fixed; 2000 unit loads on surface nodes above z = 400; E = 1; 7 multigrid
levels, span 2; weighted Jacobi omega 0.6; tolerance 1e-8), the largest
single-GPU configuration.  One step = one full solve from U = 0 to FP64
true relative residual 1e-8 with F and U resident in HBM; the line's value
is the mixed-precision solve (FP32 inner PCG, V-cycle and K.u; FP64 reliable
updates, scalars and coarsest solve), with the plain FP64 solve of the same
problem beside it.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2", choices=["cfg0", "cfg1", "cfg2", "cfg3", "cfg4", "part_r256",
                                                           "part_r256_default"],
                    help="cfg2: ~125 M-DOF ellipsoid union in 512^3 (default); cfg0: 64x32x32 solid cuboid; "
                         "cfg1: 256x128x128 GRF cuboid; "
                         "cfg3: 16 load cases on a 256^3 GRF cuboid, batched mixed vs 16 sequential FP64 "
                         "solves; cfg4: 32 uneven cuboids solved together (one launch per kernel over all instances, "
                         "replayed as a graph) plus sensitivities; "
                         "part_r256: the 256x160x112 voxel model data/part_R256.TopVoxel (E=1, 6 levels); "
                         "part_r256_default: the same on the reference's default hierarchy (4 levels, "
                         "nonDyadic: spans 4, 2; 20x32x14-element coarsest level, 16.5 k free DOFs)")
    ap.add_argument("--n3", type=int, default=256, help="cube edge of the batched workload")
    ap.add_argument("--hw-queues", type=int, default=None,
                    help="hardware queues for the multi-instance workload (sets GPU_MAX_HW_QUEUES, <= 32; "
                         "matters for the per-instance graphs of SGLD_MULTI_GRAPHS=1, where 4 measured fastest: "
                         "0.126 s vs 0.187 s with 8, 0.175 s with 16)")
    ap.add_argument("--size", default="256x128x128")
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--cpu-its", type=int, default=3, help="PCG iterations in the CPU sample (1 above 50 M DOF)")
    ap.add_argument("--cpu-full", type=int, default=0, help="also time this many whole oracle solves (median)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ref-case", action="store_true",
                    help="also time the 64x32x32 reference case (GPU and oracle whole solves) beside the workload")
    ap.add_argument("--probe-its", type=int, default=0,
                    help="only a fixed-length solve of this many iterations per precision plus a 256 MiB "
                         "torch copy (for rocprofv3 --pmc runs)")
    ap.add_argument("--kbyu-reps", type=int, default=100, help="timed back-to-back K.u launches")
    ap.add_argument("--kbyu-warmup", type=int, default=300,
                    help="untimed K.u launches first: after idle time the FP64 kernel runs a ~30 ms "
                         "clock transient (10-launch windows 80 -> 92 -> 72 us)")
    ap.add_argument("--kbyu-only", action="store_true",
                    help="only K.u launches plus a 256 MiB torch copy (for rocprofv3 --pmc runs)")
    return ap.parse_args()


def measured_copy_GBs(reps=10):
    """Achievable HBM bandwidth on this GPU: a 1 GiB device-to-device copy
    (read + write), HIP-event timed; the spec peak stays the roofline's peak."""
    import torch
    src = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * 128 * 1024 * 1024 * 8 / ms / 1e6


def measured_bf16_gemm_TFLOPs(n=8192, reps=10):
    """Achievable dense BF16 matrix-core rate: an n^3 torch matmul (rocBLAS /
    hipBLASLt), HIP-event timed."""
    import torch
    a = torch.randn(n, n, dtype=torch.bfloat16, device="cuda:0")
    b = torch.randn(n, n, dtype=torch.bfloat16, device="cuda:0")
    for _ in range(3):
        c = a @ b
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        c = a @ b
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b, c
    torch.cuda.empty_cache()
    return 2.0 * n ** 3 / ms / 1e9


def time_kbyu(a, s, prec):
    """ms per launch of `a.kbyu_reps` back-to-back K.u launches (HIP events on
    the solver stream) after `a.kbyu_warmup` untimed ones."""
    if a.kbyu_warmup > 0:
        s.bench_kbyu(prec, a.kbyu_warmup)
    return s.bench_kbyu(prec, a.kbyu_reps)


def kbyu_only(a, s, n, ne):
    import torch
    from sgldbench_amd import solver
    ms64 = time_kbyu(a, s, solver.PREC_FP64)
    ms32 = time_kbyu(a, s, solver.PREC_MIXED)
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kbyu_only": True, "fp64_ms": ms64, "fp32_ms": ms32, "reps": a.kbyu_reps,
                      "copy_bytes": src.numel() * 8, "dofs": n, "elements": ne}), flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads the oracle runs with (OMP_NUM_THREADS, else the CPUs this
    process may run on) and that affinity count."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return int(os.environ.get("OMP_NUM_THREADS", aff)), aff


def cpu_baseline(p, its_fp64, sample_its=3, one_thread=True, lean=False, full_solves=0):
    """Oracle (C/OpenMP restatement of the reference algorithm) on the host:
    `sample_its` FP64 PCG iterations of the same workload with OMP_NUM_THREADS
    threads bound close (after its set-up and the initial z = M r), scaled to
    the GPU FP64 iteration count; optionally one iteration on one thread and
    `full_solves` complete solves (median)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")  # read when the OpenMP runtime starts
    from oracle import oracle as orc
    threads, aff = cpu_threads()
    orc.set_threads(threads)
    t0 = time.time()
    m = orc.Mesh(p["voxels"], p["num_levels"])
    m.set_fixed(p["fixed"])
    m.set_modulus(p["E"])
    m.assemble(lean=lean)
    t_setup = time.time() - t0
    progress("oracle set-up %.1f s (%d threads); timing %d iteration(s)" % (t_setup, threads, sample_its))
    t_it, t_k, t_v = m.time_pcg_iterations(p["F"], p["omega"], sample_its)
    progress("oracle: %.2f s per iteration" % (t_it / sample_its))
    per_it = t_it / sample_its
    n = p["F"].size
    t_solve = per_it * its_fp64
    out = {"value": n / t_solve / 1e9, "unit": "GDOF/s", "cores": threads, "kind": "port",
           "sample": "%d FP64 PCG iteration(s) of the same workload on the oracle, %d threads "
                     "(%.2f s/iteration: K.u %.2f s, V-cycle %.2f s; set-up %.1f s not counted), "
                     "scaled to the GPU's %d FP64 iterations" % (sample_its, threads, per_it, t_k / sample_its,
                                                                 t_v / sample_its, t_setup, its_fp64),
           "solve_s_est": t_solve, "s_per_iteration": per_it, "kbyu_s_per_iteration": t_k / sample_its,
           "vcycle_s_per_iteration": t_v / sample_its, "setup_s": t_setup,
           "cpu": cpu_model(), "affinity_cpus": aff, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "lean_level2_storage": bool(lean)}
    if one_thread:
        orc.set_threads(1)
        t1, t1k, t1v = m.time_pcg_iterations(p["F"], p["omega"], 1)
        orc.set_threads(threads)
        out["one_thread"] = {"s_per_iteration": t1, "kbyu_s": t1k, "vcycle_s": t1v, "solve_s_est": t1 * its_fp64}
    if full_solves:
        ts, its = [], None
        for _ in range(full_solves + 1):  # the first is a warm-up
            t = time.time()
            _, its, _, rc = m.pcg(p["F"], p["omega"], p["tol"], p.get("max_it", 500))
            ts.append(time.time() - t)
            progress("oracle full solve %d/%d: %.1f s, %d its" % (len(ts), full_solves + 1, ts[-1], its))
        med = float(np.median(ts[1:]))
        out["full_solves"] = {"count": full_solves, "median_s": med, "all_s": ts[1:], "its": its,
                              "converged": rc == 0, "gdofs": n / med / 1e9}
    return out


def provenance(p):
    """SHA-256 (first 16 hex digits) of the generated workload arrays."""
    import hashlib
    h = {}
    for k in ("voxels", "E", "F", "fixed"):
        a = np.ascontiguousarray(np.asarray(p[k]))
        h[k] = hashlib.sha256(a.tobytes()).hexdigest()[:16]
    return h


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


BF16_PEAK_TFLOPS = 2500.0  # dense


def run_multi(a):
    """32 uneven cuboid instances (sizes from {32,48,64,96,128} per axis), each
    with its own SIMP density: one step = all instances solved to FP64 true
    rel-res 1e-8 (mixed precision) together - every kernel of the PCG
    iteration launched once over all instances, reliable updates on the
    device, batches of iterations replayed as one graph - then the fused
    compliance sensitivities -p x^(p-1) u_e^T Ke u_e of every instance.
    Compared with the same work done instance by instance."""
    # instances replay on their own streams, which the runtime maps onto its
    # hardware queues (GPU_MAX_HW_QUEUES, read when it initialises)
    if a.hw_queues is not None:
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(1, a.hw_queues)))
    import torch
    from sgldbench_amd import problems, solver
    ins = problems.cfg4()
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    S, dF, dU, dx, ddc, tune = [], [], [], [], [], []
    t0 = time.time()
    for d in ins:
        s = solver.FEASolver(d["voxels"], d["num_levels"])
        s.set_fixed(d["fixed"])
        s.set_modulus(d["E"])
        s.assemble(d["omega"])
        t = s.autotune(reps=3)
        # The autotuner times single launches.  In the merged multi-instance
        # launches throughput counts, and each level's restriction is one
        # launch only if every instance runs the same kernel there:
        # * the dense node-centric K.u wins single launches on small meshes by
        #   latency, the mode-space K.u costs less per element in merged ones
        #   (32 uneven cuboids, mixed: 85.7 -> 77.0 ms of graphs);
        # * the restriction is pinned to LDS-tiled into level 1 and
        #   node-centric below for every instance.
        if t["kbyu_variant"] != "tiled":
            s.set_schedule("tiled", 12, "tiled-first")
            t = dict(t, kbyu_variant="tiled (pinned for the merged launches)", z_chunk=12)
        else:
            s.set_schedule("tiled", 0, "tiled-first")  # keeps the autotuned z-chunk of each precision
        t = dict(t, restriction="tiled into level 1, node-centric below (pinned for the merged launches)")
        tune.append(t)
        S.append(s)
        dF.append(torch.from_numpy(d["F"]).to(dev))
        dU.append(torch.zeros_like(dF[-1]))
        dx.append(torch.from_numpy(d["x"]).to(dev))
        ddc.append(torch.zeros_like(dx[-1]))
    setup_s = time.time() - t0
    n_tot = sum(d["F"].size for d in ins)
    progress("uneven batch: %d instances, %d DOF in total, set-up + autotune %.1f s" % (len(ins), n_tot, setup_s))

    def sens_all():
        return [s.sensitivity_device(u.data_ptr(), x.data_ptr(), c.data_ptr(), penal=d["penal"])
                for s, u, x, c, d in zip(S, dU, dx, ddc, ins)]

    penal = ins[0]["penal"]
    assert all(d["penal"] == penal for d in ins)

    def sens_merged():  # one launch over all instances' elements
        return solver.multi_sensitivity_device(S, [u.data_ptr() for u in dU], [x.data_ptr() for x in dx],
                                               [c.data_ptr() for c in ddc], penal=penal).tolist()

    def step_multi(batch, prec=solver.PREC_MIXED):
        for u in dU:
            u.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its, rr, lat, ok = solver.multi_pcg_device(S, [f.data_ptr() for f in dF], [u.data_ptr() for u in dU],
                                                   tol=1e-8, maxIT=2000, precision=prec, batch=batch)
        comp = sens_merged()
        torch.cuda.synchronize()
        return time.perf_counter() - t, its, rr, lat, ok, comp

    def step_sequential(prec):
        for u in dU:
            u.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its, rr = [], []
        for s, f, u in zip(S, dF, dU):
            i, r = s.pcg_device(f.data_ptr(), u.data_ptr(), 1e-8, 2000, prec)
            its.append(i)
            rr.append(r)
        comp = sens_all()
        torch.cuda.synchronize()
        return time.perf_counter() - t, np.array(its), np.array(rr), comp

    # schedule: iteration batch length of the replayed graph, picked by measurement
    trial = {}
    for b in (4, 8, 16):
        trial[b] = min(step_multi(b)[0] for _ in range(2))
        progress("graph batch %d: %.3f s" % (b, trial[b]))
    batch = min(trial, key=trial.get)
    for _ in range(a.warmup):
        step_multi(batch)
    runs = []
    for k in range(a.steps):
        runs.append(step_multi(batch))
        progress("multi step %d: %.3f s, max its %d, max rel-res %.2e" % (k, runs[-1][0], runs[-1][1].max(),
                                                                        runs[-1][2].max()))
    multi64 = [step_multi(batch, solver.PREC_FP64) for _ in range(max(1, a.steps // 2))]
    t_multi64 = float(np.mean([r[0] for r in multi64]))
    progress("multi FP64: %.3f s" % t_multi64)
    seq = {}
    for prec, name in ((solver.PREC_MIXED, "sequential_mixed"), (solver.PREC_FP64, "sequential_fp64")):
        ts = []
        for _ in range(max(1, a.steps // 2)):
            ts.append(step_sequential(prec))
        seq[name] = dict(s_per_step=float(np.mean([r[0] for r in ts])), its=ts[-1][1].tolist(),
                         max_relres=float(max(r[2].max() for r in ts)))
        progress("%s: %.3f s" % (name, seq[name]["s_per_step"]))
    t_multi = float(np.mean([r[0] for r in runs]))
    lat = np.array([r[3] for r in runs])  # steps x instances
    per_inst = [dict(name=d["name"], dofs=int(d["F"].size), levels=d["num_levels"], its=int(runs[-1][1][i]),
                     p50_ms=1e3 * float(np.percentile(lat[:, i], 50)), p99_ms=1e3 * float(np.percentile(lat[:, i], 99)),
                     kbyu=tune[i]["kbyu_variant"], z_chunk=tune[i]["z_chunk"], restriction=tune[i]["restriction"])
                for i, d in enumerate(ins)]
    out = {
        "metric": "MG-PCG time to rel-res 1e-8 & K.u GDOF/s, FP64 vs mixed/batched, 1 MI355X",
        "value": n_tot / t_multi / 1e9,
        "unit": "GDOF/s (DOF solved to rel-res 1e-8 per second, incl. sensitivities)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * t_multi,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded sizes and Gaussian random density fields)",
        "config": {"workload": ("%d cuboids, edges from {32,48,64,96,128} (seeded), %d DOF in total, "
                                "cantilever BCs, SIMP p=3 densities, 4-5 levels, mixed precision, tol 1e-8, "
                                "merged launches over all instances replayed as a graph, then sensitivities"
                                % (len(ins), n_tot)),
                   "path": "multi-instance merged launches, mixed", "graph_batch": batch, "instances": len(ins)},
        "multi": {"s_per_step": t_multi, "graph_batch_trials_s": trial, "fp64_s_per_step": t_multi64,
                  "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default"),
                  "speedup_vs_sequential_mixed": seq["sequential_mixed"]["s_per_step"] / t_multi,
                  "speedup_vs_sequential_fp64": seq["sequential_fp64"]["s_per_step"] / t_multi,
                  "sequential_mixed": seq["sequential_mixed"], "sequential_fp64": seq["sequential_fp64"],
                  "latency_ms_p50_over_instances": 1e3 * float(np.percentile(lat, 50)),
                  "latency_ms_p99_over_instances": 1e3 * float(np.percentile(lat, 99)),
                  "setup_and_autotune_s": setup_s, "per_instance": per_inst},
    }
    print(json.dumps(out), flush=True)


def run_batched(a):
    """16 load cases on an n^3 GRF cuboid through the batched entry point:
    sequential mixed-precision solves (its default), the interleaved 16-case
    iteration (BF16 matrix-core K.u) and 16 sequential FP64 solves of the same
    cases.  One step = all 16 cases solved to FP64 true rel-res 1e-8; the
    line's value is the faster mixed mode."""
    import torch
    from sgldbench_amd import problems, solver
    n3 = a.n3
    p = problems.cfg3(n3, 16)
    k = p["F"].shape[0]
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    s = solver.FEASolver(p["voxels"], p["num_levels"], nonDyadic=p.get("non_dyadic", 0))
    s.set_fixed(p["fixed"])
    s.set_modulus(p["E"])
    t0 = time.time()
    s.assemble(p["omega"])
    setup_s = time.time() - t0
    n, ne = s.numDOFs, s.numElements
    progress("batched workload: %d^3, %d DOF x %d cases, assembled in %.2f s" % (n3, n, k, setup_s))
    dF = torch.from_numpy(p["F"]).to(dev)
    dU = torch.zeros_like(dF)

    def solve(prec):
        dU.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its, rr = s.pcg_batched_device(k, dF.data_ptr(), dU.data_ptr(), p["tol"], 5000, prec)
        torch.cuda.synchronize()
        return time.perf_counter() - t, its, rr

    res, U = {}, {}
    for prec, name in ((solver.PREC_FP64, "fp64_sequential"), (solver.PREC_MIXED, "mixed_sequential"),
                       (solver.PREC_INTERLEAVED, "interleaved_mixed")):
        for _ in range(a.warmup):
            solve(prec)
        runs = []
        for _ in range(a.steps):
            runs.append(solve(prec))
            progress("%s: %.3f s, its %s, max rel-res %.2e" % (name, runs[-1][0], runs[-1][1].tolist(),
                                                              runs[-1][2].max()))
        ts = [r[0] for r in runs]
        res[name] = dict(s_per_step=float(np.mean(ts)), its=runs[-1][1].tolist(),
                         max_relres=float(max(r[2].max() for r in runs)),
                         converged=all(bool(np.all(r[2] <= p["tol"])) for r in runs))
        U[name] = dU.cpu().numpy()
    F = p["F"]
    c64 = np.einsum("ij,ij->i", F, U["fp64_sequential"])

    def agreement(name):
        c = np.einsum("ij,ij->i", F, U[name])
        urel = np.linalg.norm(U[name] - U["fp64_sequential"], axis=1) / np.linalg.norm(U["fp64_sequential"], axis=1)
        return {"compliance_rel_diff_max": float(np.max(np.abs(c - c64) / np.abs(c64))),
                "u_rel_diff_max": float(urel.max())}

    if a.kbyu_warmup > 0:  # each batched launch is ~6 ms: a tenth of the launches covers the transient
        s.bench_kbyu_batched(max(1, a.kbyu_warmup // 10))
    ms_b = s.bench_kbyu_batched(a.kbyu_reps)
    ms64 = time_kbyu(a, s, solver.PREC_FP64)
    ms32 = time_kbyu(a, s, solver.PREC_MIXED)
    useful = 1152.0 * k * ne                       # flops of 16 dense 24x24 element products
    executed = 6 * 2 * 32 * 32 * 16 * (ne / 2.0) * 1.0  # 6 MFMA 32x32x16 per element pair (hi/lo, padding)
    bytes_b = 2 * n * k * 4 + ne * 4
    bytes32 = 2 * n * 4 + ne * 4
    t_s = res["fp64_sequential"]["s_per_step"]
    best = min(("mixed_sequential", "interleaved_mixed"), key=lambda m: res[m]["s_per_step"]
               if res[m]["converged"] else float("inf"))
    t_b = res[best]["s_per_step"]
    copy_gbs = measured_copy_GBs()
    gemm_tf = measured_bf16_gemm_TFLOPs()
    if best == "mixed_sequential":  # dominant kernel: the single-case FP32 K.u
        roof = {"bound": "hbm", "achieved": bytes32 / ms32 / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": bytes32 / ms32 / 1e6 / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_kbyu_owner<float>", "bytes_per_launch": bytes32,
                "measured_copy_GBs": copy_gbs, "frac_of_measured": bytes32 / ms32 / 1e6 / copy_gbs}
    else:
        roof = {"bound": "mfma", "achieved": useful / ms_b / 1e9, "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": useful / ms_b / 1e9 / BF16_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_bkbyu", "flops_per_launch": useful,
                "hbm_frac": bytes_b / ms_b / 1e6 / HBM_PEAK_GBS,
                "measured_bf16_gemm_TFLOPs": gemm_tf, "frac_of_measured": useful / ms_b / 1e9 / gemm_tf}
    out = {
        "metric": "MG-PCG time to rel-res 1e-8 & K.u GDOF/s, FP64 vs mixed/batched, 1 MI355X",
        "value": k * n / t_b / 1e9,
        "unit": "GDOF/s (DOF x load cases solved to rel-res 1e-8 per second)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * t_b, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if best == "mixed_sequential" else "bf16x3+f32+f64",
        "data": "synthetic (seeded Gaussian random density field, seeded load directions)",
        "config": {"workload": ("cuboid %d^3 voxels, %d DOF, SIMP E from GRF, x=0 fixed, %d load cases "
                                "(unit force on the x=%d face, random directions), 6 levels, omega 0.6, tol 1e-8"
                                % (n3, n, k, n3)),
                   "path": "batched entry point, " + best.replace("_", " "), "dofs": n, "elements": ne,
                   "cases": k},
        "batched": {"mixed_sequential": dict(res["mixed_sequential"], **agreement("mixed_sequential")),
                    "interleaved_mixed": dict(res["interleaved_mixed"], **agreement("interleaved_mixed")),
                    "fp64_sequential": res["fp64_sequential"],
                    "per_case_speedup": t_s / t_b,
                    "per_case_speedup_interleaved": t_s / res["interleaved_mixed"]["s_per_step"],
                    "setup_s": setup_s},
        "kbyu": {"batched_ms": ms_b, "fp64_single_ms": ms64, "fp32_single_ms": ms32,
                 "batched_per_case_ms": ms_b / k, "per_case_speedup_vs_fp64": ms64 * k / ms_b,
                 "per_case_speedup_vs_fp32": ms32 * k / ms_b,
                 "batched_GBs": bytes_b / ms_b / 1e6,
                 "useful_TFLOPs": useful / ms_b / 1e9, "executed_mfma_TFLOPs": executed / ms_b / 1e9,
                 "measured_bf16_gemm_TFLOPs": gemm_tf, "measured_copy_GBs": copy_gbs},
        "roofline": roof,
    }
    if not a.no_cpu:
        p1 = dict(p)
        p1["F"] = p["F"][0]
        cb = cpu_baseline(p1, int(np.mean(res["fp64_sequential"]["its"])), sample_its=a.cpu_its)
        cb["sample"] += "; 16 cases = 16 sequential solves"
        cb["solve_s_est"] *= k
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def kernel_bytes(key, n, n2, ne, kernel=""):
    """Algorithmic bytes of one launch of a finest-level pass (each input read
    once, each output written once; node flags and chunk lists excluded):
    n DOFs of level 1, n2 of level 2, ne elements.  The residual update fused
    with the first restriction (k_update_restrict) also writes level 2's
    residual."""
    s = 8 if key.endswith("64") else 4
    dw = 8 if s == 8 else 2  # Jacobi weights: FP64, or bfloat16 in the FP32 inner solve
    if key.startswith("dir"):       # read x_2, r, dinv, p, x; write p, x
        return n * (5 * s + dw) + n2 * s
    if key.startswith("kbyu"):      # read u, E; write y
        return 2 * n * s + ne * s
    if key.startswith("update_r"):  # read r, q, dinv; write r (and r_2 when fused)
        return n * (3 * s + dw) + (n2 * s if "restrict" in kernel else 0)
    if key.startswith("restrict"):  # read r_1; write r_2
        return (n + n2) * s
    raise KeyError(key)


TRAFFIC_FILES = {"cfg1": "kbyu_traffic_cuboid256x128x128.md", "cfg2": "traffic_ellipsoids512.md"}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy
    kernel: lines "traffic_bytes_per_launch[<kernel>]: <bytes>"), or None."""
    name = TRAFFIC_FILES.get(workload)
    if not name:
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch[%s]:" % kernel):
                return float(line.split("]:")[1])
    except OSError:
        pass
    return None


def load_workload(a):
    """(problem dict, workload description, data description)."""
    from sgldbench_amd import problems
    if a.workload == "cfg2":
        p = problems.cfg2()
        ny, nx, nz = p["voxels"].shape
        desc = ("ellipsoid union in %dx%dx%d box (fill %.3f), %d DOF, E=1, z<8 fixed, 2000 unit loads, "
                "%d levels, omega 0.6, tol 1e-8" % (nx, ny, nz, float(p["voxels"].mean()), p["F"].size,
                                                    p["num_levels"]))
        return p, desc, "synthetic (seeded ellipsoid union)"
    if a.workload in ("part_r256", "part_r256_default"):
        from sgldbench_amd import topvoxel
        default = a.workload == "part_r256_default"
        p = topvoxel.setup(topvoxel.load_packed(os.path.join(ROOT, "tests", "golden", "part_R256.npz")),
                           num_levels=None if default else 6)
        p["non_dyadic"] = 1 if default else 0
        ny, nx, nz = p["voxels"].shape
        desc = ("voxel model part_R256 (256x160x112, padded %dx%dx%d), %d DOF, E=1, its 13042 fixations "
                "and 7788 loads, %d levels%s, omega 0.6, tol 1e-8"
                % (nx, ny, nz, p["F"].size, p["num_levels"],
                   " (the reference's rule; nonDyadic: 3 levels, spans 4, 2)" if default else ""))
        return p, desc, "the reference's voxel model data/part_R256.TopVoxel (packed copy)"
    if a.workload == "cfg0":  # the reference's own CPU-sized case: E = 1, 4 levels
        p = problems.cfg0()
        ny, nx, nz = p["voxels"].shape
        desc = ("solid cuboid %dx%dx%d voxels, %d DOF, E=1, cantilever, %d levels, omega 0.6, tol 1e-8"
                % (nx, ny, nz, p["F"].size, p["num_levels"]))
        return p, desc, "synthetic (solid cuboid, E = 1)"
    nx, ny, nz = (int(v) for v in a.size.split("x"))
    p = problems.cfg1(nx, ny, nz, a.levels)
    desc = ("cuboid %dx%dx%d voxels, %d DOF, SIMP E from GRF, cantilever, %d levels, omega 0.6, tol 1e-8"
            % (nx, ny, nz, p["F"].size, p["num_levels"]))
    return p, desc, "synthetic (seeded Gaussian random density field)"


def reference_case(a):
    """The reference's CPU-runnable case (64x32x32 solid cantilever, 4
    levels): median of 5 whole solves on the GPU (FP64 and mixed) and on the
    oracle's host threads."""
    import torch
    from sgldbench_amd import problems, solver
    p = problems.cfg0()
    s = solver.FEASolver(p["voxels"], p["num_levels"], nonDyadic=p.get("non_dyadic", 0))
    s.set_fixed(p["fixed"])
    s.set_modulus(p["E"])
    s.assemble(p["omega"])
    dF = torch.from_numpy(p["F"]).to("cuda:0")
    dU = torch.zeros_like(dF)
    out = {"workload": "solid cuboid 64x32x32, 212355 DOF, E=1, cantilever, 4 levels, omega 0.6, tol 1e-8"}
    for prec, name in ((solver.PREC_FP64, "gpu_fp64"), (solver.PREC_MIXED, "gpu_mixed")):
        ts = []
        for _ in range(6):
            dU.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            its, rr = s.pcg_device(dF.data_ptr(), dU.data_ptr(), p["tol"], 500, prec)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        out[name] = {"median_ms": 1e3 * float(np.median(ts[1:])), "its": its, "relres": rr}
    s.close()
    if not a.no_cpu:
        cb = cpu_baseline(p, out["gpu_fp64"]["its"], sample_its=3, one_thread=False, full_solves=5)
        out["cpu_oracle"] = {"median_s": cb["full_solves"]["median_s"], "its": cb["full_solves"]["its"],
                             "threads": cb["cores"], "s_per_iteration": cb["s_per_iteration"]}
    return out


def run_single(a):
    """One-RHS workloads (cfg0, cfg1, cfg2, part_r256): FP64 and mixed
    MG-PCG solves from U = 0 to FP64 true rel-res 1e-8 with F, U in HBM."""
    import torch
    from sgldbench_amd import solver
    t_gen = time.time()
    p, desc, data = load_workload(a)
    n = p["F"].size
    ne = int(p["E"].size)
    prov = provenance(p)
    progress("problem %s: %d DOF, %d elements (generated in %.1f s)" % (a.workload, n, ne, time.time() - t_gen))
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    s = solver.FEASolver(p["voxels"], p["num_levels"], nonDyadic=p.get("non_dyadic", 0))
    s.set_fixed(p["fixed"])
    s.set_modulus(p["E"])
    if a.kbyu_only:
        kbyu_only(a, s, n, ne)
        return
    t0 = time.time()
    s.assemble(p["omega"])
    setup_s = time.time() - t0
    n2 = int(s.levels[1]["numDOFs"])
    progress("assembled in %.2f s (level 2: %d DOF)" % (setup_s, n2))
    dF = torch.from_numpy(p["F"]).to(dev)
    dU = torch.zeros(n, dtype=torch.float64, device=dev)

    def solve(prec, max_it=5000):
        dU.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its, rr = s.pcg_device(dF.data_ptr(), dU.data_ptr(), p["tol"], max_it, prec)
        torch.cuda.synchronize()
        return time.perf_counter() - t, its, rr, s.stats()

    if a.probe_its:  # a short, fixed-length solve per precision for counter runs
        for prec in (solver.PREC_FP64, solver.PREC_MIXED):
            solve(prec, a.probe_its)
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe_its": a.probe_its, "copy_bytes": src.numel() * 8, "dofs": n}), flush=True)
        return

    # Per precision: warm-up solves, then `steps` timed solves (the line's
    # figures; no per-kernel events), then two solves with HIP events around
    # every finest-level pass (the per-kernel figures and the roofline).
    res, ktimes = {}, {}
    for prec, name in ((solver.PREC_FP64, "fp64"), (solver.PREC_MIXED, "mixed")):
        for _ in range(a.warmup):
            solve(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        runs = []
        for _ in range(a.steps):
            runs.append(solve(prec))
            progress("%s solve: %.3f s, %d its, rel-res %.2e" % (name, runs[-1][0], runs[-1][1], runs[-1][2]))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        s.kernel_timing(True)
        inst = [solve(prec)[0] for _ in range(2)]
        ktimes[name] = s.kernel_times()
        s.kernel_timing(False)
        its = runs[-1][1]
        steps_ms = sorted(1e3 * r[0] for r in runs)
        res[name] = dict(ms_per_step=1e3 * wall / a.steps, its=its, relres=max(r[2] for r in runs),
                         outer=runs[-1][3]["outer"], converged=all(r[2] <= p["tol"] for r in runs),
                         gdofs=n / (wall / a.steps) / 1e9, gdofs_x_its=n * its / (wall / a.steps) / 1e9,
                         ms_min=steps_ms[0], ms_median=steps_ms[len(steps_ms) // 2], ms_max=steps_ms[-1],
                         ms_instrumented=1e3 * float(np.mean(inst)), instrumented_solves=len(inst))
    U = dU.cpu().numpy()
    comp_mixed = float(p["F"] @ U)
    solve(solver.PREC_FP64)
    U64 = dU.cpu().numpy()
    comp_fp64 = float(p["F"] @ U64)
    ms64 = time_kbyu(a, s, solver.PREC_FP64)
    ms32 = time_kbyu(a, s, solver.PREC_MIXED)
    copy_gbs = measured_copy_GBs()
    filt = None
    if a.workload != "cfg2":  # density filter of the same mesh (rMin 2.5: 5^3 cone), kernel time
        dxf = torch.rand(ne, dtype=torch.float64, device=dev)
        dyf = torch.zeros_like(dxf)
        fms = []
        for _ in range(4):
            s.density_filter_device(dxf.data_ptr(), 2.5, dyf.data_ptr())
            fms.append(s.stats()["filter_ms"])
        filt = {"rMin": 2.5, "ms": min(fms[1:]), "GB_s_algorithmic": 2 * ne * 8 / min(fms[1:]) / 1e6}
    # the line reports the faster of the converged paths (mixed on the large
    # workloads; FP64 on the 212 k-DOF case, where reliable updates and launch
    # overhead outweigh the cheaper FP32 passes)
    best = min((m for m in ("mixed", "fp64") if res[m]["converged"]), key=lambda m: res[m]["ms_per_step"],
               default="fp64")
    t_best = res[best]["ms_per_step"] / 1e3
    # per-kernel figures of the timed solves; the roofline names the kernel
    # with the largest share of the headline path's time
    kern = {}
    for path in ("fp64", "mixed"):
        kern[path] = {}
        for key, v in ktimes[path].items():
            if not v["work_launches"]:
                continue
            avg = v["work_ms"] / v["work_launches"]
            b = kernel_bytes(key, n, n2, ne, v["kernel"])
            kern[path][key] = {"kernel": v["kernel"], "launches": v["launches"], "work_launches": v["work_launches"],
                               "avg_us": 1e3 * avg, "avg_us_all_launches": 1e3 * v["ms"] / v["launches"],
                               "bytes_per_launch": b, "GB_s": b / avg / 1e6, "frac": b / avg / 1e6 / HBM_PEAK_GBS,
                               "share_of_solve": v["work_ms"] / (res[path]["ms_instrumented"]
                                                                 * res[path]["instrumented_solves"])}
    dom = max(kern[best], key=lambda k: kern[best][k]["avg_us"] * kern[best][k]["work_launches"])
    d = kern[best][dom]
    roof = {"bound": "hbm", "achieved": d["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
            "traffic": measured_traffic(a.workload, d["kernel"]), "kernel": d["kernel"], "path": best,
            "bytes_per_launch": d["bytes_per_launch"], "avg_us": d["avg_us"], "launches": d["work_launches"],
            "measured_copy_GBs": copy_gbs, "frac_of_measured": d["GB_s"] / copy_gbs}
    bytes64, bytes32 = kernel_bytes("kbyu64", n, n2, ne), kernel_bytes("kbyu32", n, n2, ne)
    out = {
        "metric": "MG-PCG time to rel-res 1e-8 & K.u GDOF/s, FP64 vs mixed/batched, 1 MI355X",
        "value": n / t_best / 1e9,
        "unit": "GDOF/s (DOF solved to rel-res 1e-8 per second)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": res[best]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if best == "mixed" else "f64",
        "data": data,
        "config": {"workload": desc, "path": best, "dofs": n, "elements": ne, "level2_dofs": n2,
                   "sha256_16": prov},
        "solve": {"fp64": res["fp64"], "mixed": res["mixed"],
                  "speedup_mixed_vs_fp64": res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"],
                  "compliance_fp64": comp_fp64, "compliance_mixed": comp_mixed,
                  "compliance_rel_diff": abs(comp_mixed - comp_fp64) / abs(comp_fp64),
                  "u_rel_diff": float(np.linalg.norm(U - U64) / np.linalg.norm(U64)),
                  "setup_s": setup_s},
        "kernels": kern,
        "kbyu": {"fp64_ms": ms64, "fp32_ms": ms32, "fp64_gdofs": n / ms64 / 1e6,
                 "fp32_gdofs": n / ms32 / 1e6, "fp64_GBs": bytes64 / ms64 / 1e6,
                 "fp32_GBs": bytes32 / ms32 / 1e6, "note": "back-to-back launches, HIP events"},
        "roofline": roof,
    }
    if filt:
        out["density_filter"] = filt
    s.close()
    del dF, dU
    torch.cuda.empty_cache()
    if a.workload == "part_r256_default" and not a.no_cpu:
        # the oracle factors its coarsest level with an unblocked dense
        # Cholesky: hours at 16.5 k free DOFs
        out["cpu_baseline"] = {"value": None, "note": "not sampled: the oracle's coarsest factor at 16.5 k DOFs"}
    elif not a.no_cpu:
        big = n > 50_000_000
        out["cpu_baseline"] = cpu_baseline(p, res["fp64"]["its"], sample_its=1 if big else a.cpu_its,
                                           one_thread=not big, lean=big, full_solves=a.cpu_full)
    if a.workload != "cfg0" and a.ref_case:
        out["reference_case"] = reference_case(a)
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: the solver runs on one GPU (--gpus %d requested); "
              "a sharded multi-GPU solve is not implemented." % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.workload == "cfg3":
        run_batched(a)
    elif a.workload == "cfg4":
        run_multi(a)
    else:
        run_single(a)


if __name__ == "__main__":
    main()
