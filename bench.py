"""Benchmark: MG-PCG solve time to 1e-6 relative residual (mixed vs FP64) and
stiffness-apply throughput on the 256^3 narrow-band sphere shell.

A step is one complete mixed-precision solve (FP32 MG-PCG inside an FP64
defect-correction loop) of K(rho) u = f from a zero initial guess to
||P(f - K u)||_inf <= 1e-6 ||P f||_inf, inputs resident in HBM.  Prints one
JSON line on stdout; logs go to stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--levels 5]
                  [--no-cpu] [--cpu-runs 5]

Beside the timed GPU solves it reports the accuracy of both GPU solves
against the FP64 oracle's solve to 1e-10 and, as the CPU baseline, the
oracle's own solve time to the same tolerance on the host's cores (median of
--cpu-runs); --no-cpu skips the oracle and compares with the GPU's own FP64
solve instead.

With --opt-loop it instead times the SIMP/OC optimization loop on the
512x256x256 band (cells within 24 cells of the bounding-box surface):
--opt-iters iterations with the narrow band updated from iteration 5, once
with mixed-precision and once with FP64 solves, and prints seconds per
optimization step and both final compliances.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MG-PCG solve time to 1e-6 rel. residual, mixed vs FP64; MDOF/s stiffness apply"
HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE_FP32 = 28.0  # u 12 + f 12 + E 4 (algorithmic, per active node)
BYTES_PER_NODE_FP64 = 56.0
GS_SWEEP_BYTES_PER_NODE_FP32 = 40.0  # one 8-colour sweep: u read + write 24, b 12, E 4 (algorithmic)
GS_KERNEL = "k_gs0_colour4<float, 8>"
APPLY_KERNEL = "k_apply_cell<float, 2>"
# what the mixed solve stores and sums in
ACCUMULATION = ("FP32 storage and FP32 sums throughout the inner MG-PCG (the Krylov operator q = P K s cell by cell "
                "in strain form, where rigid motions cancel exactly); FP64 outer residual, PCG scalars and dots, "
                "Galerkin assembly and bottom solve")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["shell", "ellipsoid"], default="shell",
                    help="shell: 256^3 sphere shell (default); ellipsoid: 2048x1024x1024 ellipsoidal shell")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed oracle solves (median reported)")
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--max-iterations", type=int, default=500, help="PCG iterations per pass (FP64 solve: 4x)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the FP32 apply and a 512 MiB copy (5 launches each) for rocprofv3 --pmc passes")
    ap.add_argument("--opt-loop", action="store_true", help="time the OC optimization loop on the 512x256x256 band")
    ap.add_argument("--fem-solve-calls", type=int, default=0,
                    help="time N calls of the drop-in fem_solve on the shell (wire format in and out, keep_state; "
                         "call k > 0 scales the densities by 1 - 0.02 k); SPG_SETUP_TRACE=2 prints the stages")
    ap.add_argument("--opt-iters", type=int, default=20)
    ap.add_argument("--opt-precisions", default="mixed,fp64")
    ap.add_argument("--opt-band-start", type=int, default=5)
    ap.add_argument("--opt-tol", type=float, default=1e-4,
                    help="solve tolerance in the loop (the reference optimizer's cg_tolerance default)")
    ap.add_argument("--opt-warm", action="store_true",
                    help="warm-start every solve from the previous displacement (the caller's warm_starting with "
                         "forced_reuse; off by default, as in the reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the block set over several GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)
    if args.opt_loop:
        opt_loop(args)
        return
    if args.fem_solve_calls > 0:
        fem_solve_calls(args)
        return

    import torch

    from spgrid_topo_opt_amd import _lib
    from spgrid_topo_opt_amd import workloads as W

    t0 = time.perf_counter()
    if args.workload == "ellipsoid":
        prob = W.config("cfg5", progress=lambda m: log(f"[bench] {m}"))
        workload = ("ellipsoidal shell in a 2048x1024x1024 grid, semi-axes 0.5/0.25/0.25, 16 cells thick, "
                    "density U[0.05,1], clamp at the -x pole, point load at the +x pole")
    else:
        prob = W.config("cfg2", n=args.n, mg_levels=args.levels)
        workload = (f"sphere shell {args.n}^3, cells with 0.4<=|x|<=0.5, density U[0.05,1] + 2 box filters, "
                    "clamp at south pole, point load at north pole")
    log(f"[bench] problem {prob.name}: {prob.block_base.shape[0]} blocks, built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    s = _lib.Solver(prob)
    setup_s = time.perf_counter() - t0
    log(f"[bench] solver setup {setup_s:.2f}s levels={s.levels}")

    if args.traffic_probe:
        traffic_probe(s, prob)
        return

    # kernel throughput first: it does not depend on how the solve converges
    nn = s.info()["active_nodes"]
    t_ap32 = s.bench_apply(0, args.apply_reps)  # the PCG operator: FP32, cell by cell in strain form
    t_ap64 = s.bench_apply(1, args.apply_reps)
    # the same operator as node rows: FP64 row sums, difference form, direct FP32 sums
    os.environ["SPG_APPLY_CELL"] = "0"
    t_ap32r = s.bench_apply(0, args.apply_reps)
    os.environ["SPG_DIFF_SUMS"] = "1"
    t_ap32d = s.bench_apply(0, args.apply_reps)
    os.environ["SPG_DIFF_SUMS"] = "0"
    os.environ["SPG_ACC64_MODES"] = "0"
    t_ap32f = s.bench_apply(0, args.apply_reps)
    for k in ("SPG_ACC64_MODES", "SPG_DIFF_SUMS", "SPG_APPLY_CELL"):
        del os.environ[k]
    t_gs32 = s.bench_kernel(_lib.OP_GS_SWEEP, 0, 0, 0, 20)  # 8 colour launches
    t_vc32 = s.bench_kernel(_lib.OP_VCYCLE, 0, 0, 0, 5)
    t_vc64 = s.bench_kernel(_lib.OP_VCYCLE, 0, 1, 0, 5)
    log(f"[bench] V-cycle fp32 {t_vc32 * 1e3:.3f} ms, fp64 {t_vc64 * 1e3:.3f} ms")
    gbs32 = BYTES_PER_NODE_FP32 * nn / t_ap32 / 1e9
    gbs64 = BYTES_PER_NODE_FP64 * nn / t_ap64 / 1e9
    log(f"[bench] apply fp32 {t_ap32 * 1e6:.1f} us ({3 * nn / t_ap32 / 1e6:.0f} MDOF/s, {gbs32:.0f} GB/s); "
        f"fp64 {t_ap64 * 1e6:.1f} us ({3 * nn / t_ap64 / 1e6:.0f} MDOF/s, {gbs64:.0f} GB/s)")

    pr = 2 if args.workload == "ellipsoid" else 0  # per-iteration residuals on stderr for the long solve

    def mixed():
        return s.solve_mixed(tolerance=args.tol, want_output=False, print_residuals=pr,
                             max_iterations=args.max_iterations, raise_on_error=False)[1]

    def fp64():
        return s.solve(tolerance=args.tol, precision=1, max_passes=1, max_iterations=4 * args.max_iterations,
                       want_output=False, print_residuals=pr, raise_on_error=False)[1]

    for _ in range(args.warmup):
        info = mixed()
        log(f"[bench] warmup mixed: {info['solve_seconds']:.4f}s it={info['iterations']} passes={info['passes']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [mixed() for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    info = infos[-1]
    log(f"[bench] mixed solve {t_mixed * 1e3:.2f} ms/step, iterations {info['iterations']}, passes {info['passes']}, "
        f"rel residual {info['r_linf'] / info['r0_linf']:.2e}")

    fp64()  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    i64 = fp64()
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t0
    log(f"[bench] fp64 solve {t_fp64 * 1e3:.2f} ms, iterations {i64['iterations']}, "
        f"rel residual {i64['r_linf'] / i64['r0_linf']:.2e}")

    copy_gbs = copy_bandwidth()
    log(f"[bench] device copy {copy_gbs:.0f} GB/s (read + write)")

    # the CPU leg: the FP64 oracle (the reference's algorithm restated in C,
    # OpenMP) solves the same problem to 1e-10 (the accuracy reference, and
    # the warm-up), then to args.tol five times (the CPU baseline)
    cpu, ref = (None, None) if args.no_cpu else cpu_leg(prob, args.tol, args.cpu_runs)
    # accuracy of the GPU mixed and FP64 solves (to args.tol) against that
    # reference: relative L2 displacement error and relative compliance error
    acc = accuracy_report(s, args, pr, ref)
    log(f"[bench] vs {acc['reference']['source']}: mixed u err {acc['mixed']['u_rel_l2']:.2e}, "
        f"c err {acc['mixed']['c_rel']:.2e}; fp64 u err {acc['fp64']['u_rel_l2']:.2e}, c err {acc['fp64']['c_rel']:.2e}")

    tr_gs, tr_ap = measured_traffic(GS_KERNEL, nn), measured_traffic(APPLY_KERNEL, nn)
    out = {
        "metric": METRIC,
        "value": t_mixed,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-inner/fp64-outer",
        "accumulation": ACCUMULATION,
        "data": "synthetic",
        "config": {
            "workload": workload,
            "active_cells": info["active_cells"],
            "active_nodes": nn,
            "dofs": 3 * nn,
            "mg_levels": s.levels,
            "tolerance": args.tol,
        },
        "fp64_solve_s": t_fp64,
        "mixed_speedup_vs_fp64": t_fp64 / t_mixed,
        "iterations_mixed": info["iterations"],
        "passes_mixed": info["passes"],
        "iterations_fp64": i64["iterations"],
        "rel_residual_mixed": info["r_linf"] / info["r0_linf"],
        "converged_mixed": bool(info["converged"]),
        "converged_fp64": bool(i64["converged"]),
        "rel_residual_fp64": i64["r_linf"] / i64["r0_linf"],
        "setup_s": setup_s,
        "hbm_footprint_gb": info["hbm_bytes"] / 1e9,
        "apply_mdof_s_fp32": 3 * nn / t_ap32 / 1e6,
        "apply_mdof_s_fp64": 3 * nn / t_ap64 / 1e6,
        "apply_us_fp32": t_ap32 * 1e6,
        "apply_us_fp32_rows_fp64_sums": t_ap32r * 1e6,
        "apply_us_fp32_difference_form": t_ap32d * 1e6,
        "apply_us_fp32_fp32_sums": t_ap32f * 1e6,
        "apply_mdof_s_fp32_fp32_sums": 3 * nn / t_ap32f / 1e6,
        "apply_us_fp64": t_ap64 * 1e6,
        "vcycle_ms_fp32": t_vc32 * 1e3,
        "vcycle_ms_fp64": t_vc64 * 1e3,
        # the kernel with the largest share of the timed step: the level-0
        # colour Gauss-Seidel launches (8 per sweep, 2 sweeps per V-cycle)
        "roofline": {
            "bound": "hbm",
            "kernel": GS_KERNEL + " (level-0 Gauss-Seidel colour launch; 8 per sweep)",
            "achieved": GS_SWEEP_BYTES_PER_NODE_FP32 / 8 * nn / (t_gs32 / 8) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": GS_SWEEP_BYTES_PER_NODE_FP32 * nn / t_gs32 / 1e9 / HBM_PEAK_GBS,
            "traffic": tr_gs,
            # the measured bytes at the measured time, against the same peak
            "traffic_frac": None if tr_gs is None else tr_gs / (t_gs32 / 8) / 1e9 / HBM_PEAK_GBS,
            "us_per_launch": t_gs32 / 8 * 1e6,
            "algorithmic_bytes_per_launch": GS_SWEEP_BYTES_PER_NODE_FP32 / 8 * nn,
        },
        "roofline_apply": {
            "bound": "hbm",
            "kernel": APPLY_KERNEL + " (level-0 stiffness apply of the PCG operator)",
            "achieved": gbs32,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs32 / HBM_PEAK_GBS,
            "traffic": tr_ap,
            "traffic_frac": None if tr_ap is None else tr_ap / t_ap32 / 1e9 / HBM_PEAK_GBS,
            "us_per_launch": t_ap32 * 1e6,
            "algorithmic_bytes_per_launch": BYTES_PER_NODE_FP32 * nn,
        },
        "copy_bandwidth_gbs": copy_gbs,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_reference": acc,
    }
    print(json.dumps(out), flush=True)


def accuracy_report(s, args, pr, ref):
    """Relative L2 displacement error and relative compliance error of the
    GPU mixed and FP64 solves (both to args.tol) against `ref`, the FP64
    oracle's solution to 10^-10 (displacements in the output block order and
    the oracle's compliance sum_e u_e^T K0 u_e E(rho_e)).  Without the CPU
    leg the reference is the GPU's own FP64 solve to 10^-10."""
    import numpy as np
    import torch

    dc = torch.empty(s.nblocks * 64, dtype=torch.float64, device="cuda")

    def run(**kw):
        out, info = s.solve(want_output=True, raise_on_error=False, print_residuals=pr, **kw)
        c = s.sensitivity_device(dc.data_ptr())
        return out[..., :3].copy(), c, info

    if ref is None:
        u_ref, c_ref, i_ref = run(tolerance=1e-10, precision=1, max_passes=1, max_iterations=8 * args.max_iterations)
        head = {"source": "GPU FP64 solve (no CPU leg)", "tolerance": 1e-10, "iterations": i_ref["iterations"],
                "converged": bool(i_ref["converged"]), "compliance": c_ref}
    else:
        u_ref, c_ref = ref["u"], ref["compliance"]
        head = {"source": "FP64 oracle solve", "tolerance": ref["tolerance"], "iterations": ref["iterations"],
                "converged": ref["converged"], "compliance": c_ref}
    out = {"reference": head}
    for name, kw in (("mixed", dict(precision=0, max_passes=30, max_iterations=args.max_iterations)),
                     ("fp64", dict(precision=1, max_passes=1, max_iterations=4 * args.max_iterations))):
        u, c, info = run(tolerance=args.tol, **kw)
        out[name] = {"u_rel_l2": float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)),
                     "c_rel": abs(c - c_ref) / abs(c_ref), "iterations": info["iterations"],
                     "rel_residual": info["r_linf"] / info["r0_linf"]}
    return out


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Read + write GB/s of a plain device-to-device copy (torch)."""
    import torch

    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    b.synchronize()
    t = a.elapsed_time(b) / 1e3 / reps
    del x, y
    torch.cuda.empty_cache()
    return 2 * nbytes / t / 1e9


def host_cpu():
    """CPU model, logical and physical core counts of this host."""
    model, phys = "unknown", set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if cur:
                        phys.add((cur.get("physical id"), cur.get("core id")))
                        cur = {}
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                cur[k] = v
                if k == "model name":
                    model = v
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_leg(prob, tol, runs):
    """The FP64 oracle (C, OpenMP) on the same problem: one solve to 1e-10
    (warm-up, and the accuracy reference), then `runs` timed solves to `tol`
    from zero; setup (flags, hierarchy, Galerkin, bottom factorization) and
    solve are reported separately, the solve as the median."""
    import statistics

    os.environ.setdefault("OMP_PROC_BIND", "close")  # read when the OpenMP runtime starts
    from oracle import oracle as O

    host = host_cpu()  # before the OpenMP runtime binds this thread to one core
    saved = os.sched_getaffinity(0)
    cores = int(os.environ.get("OMP_NUM_THREADS", len(saved)))
    t0 = time.perf_counter()
    o = O.Oracle(prob)
    setup = time.perf_counter() - t0
    u, _, info = o.solve(tolerance=1e-10, max_iterations=2000)
    c_ref, _, _ = o.sensitivity(u)
    ref = {"u": o.pack(u)[..., :3], "compliance": c_ref, "tolerance": 1e-10, "iterations": info["iterations"],
           "converged": bool(info["converged"])}
    log(f"[bench] cpu oracle: setup {setup:.1f}s, reference solve to 1e-10 in {info['solve_seconds']:.1f}s "
        f"({info['iterations']} iterations)")
    secs, its = [], []
    for _ in range(runs):
        _, _, inf = o.solve(tolerance=tol, max_iterations=2000)
        secs.append(inf["solve_seconds"])
        its.append(inf["iterations"])
    os.sched_setaffinity(0, saved)  # the bound master thread returns to all cores
    med = statistics.median(secs)
    log(f"[bench] cpu oracle: solve to {tol:g} median {med:.2f}s of {runs} ({its[0]} iterations) on {cores} threads")
    del o
    cpu = {
        "value": med,
        "unit": "s",
        "cores": cores,
        "kind": "port",
        "sample": f"the same problem solved by the FP64 oracle to {tol:g} relative L-inf residual from zero, "
                  f"median of {runs} solves after one warm-up solve (to 1e-10), setup excluded",
        "setup_s": setup,
        "solve_s_runs": secs,
        "iterations": its[0],
        "host": host,
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
    }
    return cpu, ref


COPY_BYTES = 512 << 20
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_shell256.json")


def measured_traffic(kernel, active_nodes):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    same counter reads for a copy of known size); None when no summary for
    this workload is present."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if k is None or t.get("active_nodes") != active_nodes:
        return None
    return k["fetch_kib"] * 1024 * t["fetch_scale"] + k["write_kib"] * 1024 * t["write_scale"]


def traffic_probe(s, prob):
    """A few launches of the hot FP32 kernels (level-0 apply, one level-0
    Gauss-Seidel sweep, the level-0 and level-2 boundary phases) and three of
    a plain device copy of COPY_BYTES (read COPY_BYTES + write COPY_BYTES).
    Under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) the copy's
    per-dispatch counter gives the KiB->bytes scale applied to the others."""
    import torch

    from spgrid_topo_opt_amd import _lib

    s.bench_apply(0, 5)
    s.bench_kernel(_lib.OP_GS_SWEEP, 0, 0, 0, 1)
    s.bench_kernel(_lib.OP_GS_BOUNDARY, 0, 0, prob.boundary_iters, 1)
    if s.levels > 3:
        s.bench_kernel(_lib.OP_GS_BOUNDARY, 2, 0, prob.boundary_iters, 1)
    x = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"[bench] traffic probe done: FP32 apply x5+1, GS sweep, boundary phases, copy of {COPY_BYTES} B x3")


def fem_solve_calls(args):
    """The reference caller's path: FEMInputs serialized, `fem_solve`, FEMOutputs
    parsed, with the solver kept between calls (keep_state; the last call
    releases it).  Densities change between calls within the same active set,
    so every call after the first updates the kept hierarchy and re-solves."""
    import dataclasses

    from spgrid_topo_opt_amd import _lib
    from spgrid_topo_opt_amd import fem_interface as FI
    from spgrid_topo_opt_amd import workloads as W

    t0 = time.perf_counter()
    prob = W.config("cfg2", n=args.n, mg_levels=args.levels)
    log(f"[wire] problem {prob.name}: {prob.block_base.shape[0]} blocks, built in {time.perf_counter() - t0:.1f}s")
    calls, state = [], None
    for k in range(args.fem_solve_calls):
        last = k + 1 == args.fem_solve_calls
        p = dataclasses.replace(prob, density=prob.density * (1.0 - 0.02 * k))
        inp = FI.FEMInputs.from_problem(p, tolerance=args.tol, max_iterations=args.max_iterations,
                                        extra_ints={"keep_state": 0 if last else 1})
        if state:
            inp.set_solver_state_ptr(state)
        t0 = time.perf_counter()
        data = inp.serialize()
        t1 = time.perf_counter()
        raw = _lib.fem_solve_bytes(data)
        t2 = time.perf_counter()
        out = FI.FEMOutputs.deserialize(raw)
        t3 = time.perf_counter()
        state = out.get_solver_state_ptr()
        rec = {"serialize_s": t1 - t0, "fem_solve_s": t2 - t1, "deserialize_s": t3 - t2, "in_mb": len(data) / 1e6,
               "out_mb": len(raw) / 1e6, "success": bool(out.success), "info": out.info}
        calls.append(rec)
        log(f"[wire] call {k}: serialize {rec['serialize_s']:.3f}s fem_solve {rec['fem_solve_s']:.3f}s "
            f"deserialize {rec['deserialize_s']:.3f}s {out.info}")
        if not out.success:
            raise SystemExit(f"fem_solve failed: {out.info}")
    warm = [c["fem_solve_s"] for c in calls[1:]]
    print(json.dumps({
        "metric": "seconds per fem_solve call (wire format in and out, kept solver, new densities)",
        "value": sorted(warm)[len(warm) // 2] if warm else calls[0]["fem_solve_s"],
        "unit": "s", "n_gpus": 1, "steps": len(calls), "higher_is_better": False, "data": "synthetic",
        "config": {"workload": f"sphere shell {args.n}^3", "mg_levels": args.levels, "tolerance": args.tol,
                   "blocks": int(prob.block_base.shape[0])},
        "first_call_s": calls[0]["fem_solve_s"], "calls": calls}), flush=True)


def opt_loop(args):
    """SIMP/OC loop (volume fraction 0.1, step 0.2, filter 1.5) on the band
    workload, narrow band updated every iteration from iteration 5."""
    import torch

    from spgrid_topo_opt_amd import topo_opt as T
    from spgrid_topo_opt_amd import workloads as W

    t0 = time.perf_counter()
    prob = W.config("cfg4")
    log(f"[opt] problem {prob.name}: {prob.block_base.shape[0]} blocks, {prob.n_active_cells} cells, "
        f"built in {time.perf_counter() - t0:.1f}s")
    runs = {}
    for name in args.opt_precisions.split(","):
        prec = {"mixed": 0, "fp64": 1}[name]
        t0 = time.perf_counter()
        opt = T.TopoOpt(prob, volume_fraction=0.1, step_limit=0.2, filter_size=1.5, grid_update_start=args.opt_band_start,
                        tolerance=args.opt_tol, precision=prec, warm_starting=args.opt_warm)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        log(f"[opt] {name}: setup {setup:.2f}s")
        recs = []
        t0 = time.perf_counter()
        for it in range(args.opt_iters):
            ti = time.perf_counter()
            r = opt.iterate()
            r["step_s"] = time.perf_counter() - ti
            recs.append(r)
            log(f"[opt] {name} it {it}: c={r['compliance']:.6e} step {r['step_s']:.2f}s solve {r['solve_s']:.3f}s "
                f"pcg {r['iterations']} fem {r['fem_s']:.2f}s band {r['band_update_s']:.2f}s rebuild "
                f"{r['rebuild_s']:.2f}s filter+oc "
                f"{r['filter_oc_s']:.2f}s blocks {r['fem_blocks']} fat {r['fat_blocks']} rebuilt {r['rebuilt']}")
        total = time.perf_counter() - t0
        runs[name] = {
            "s_per_step": total / args.opt_iters,
            "setup_s": setup,
            "final_compliance": recs[-1]["compliance"],
            "solve_s_mean": sum(r["solve_s"] for r in recs) / len(recs),
            "pcg_iterations": [r["iterations"] for r in recs],
            "rebuilds": sum(bool(r["rebuilt"]) for r in recs),
            "band_update_s_mean": sum(r["band_update_s"] for r in recs) / len(recs),
            "rebuild_s_mean": sum(r["rebuild_s"] for r in recs if r["rebuilt"]) / max(1, sum(bool(r["rebuilt"]) for r in recs)),
            "fem_s_mean": sum(r["fem_s"] for r in recs) / len(recs),
            "final_fem_blocks": recs[-1]["fem_blocks"],
            "compliance_history": [r["compliance"] for r in recs],
        }
        opt.close()
        del opt
        torch.cuda.empty_cache()
    first = next(iter(runs.values()))
    out = {
        "metric": "seconds per optimization step (solve + sensitivity + filter + OC + band update)",
        "value": first["s_per_step"],
        "unit": "s/step",
        "n_gpus": 1,
        "steps": args.opt_iters,
        "higher_is_better": False,
        "data": "synthetic",
        "config": {"workload": "band 512x256x256, cells within 24 cells of the bounding-box surface, "
                               "V=0.1, step 0.2, filter 1.5, 6 MG levels",
                   "band_update_from_iteration": args.opt_band_start, "warm_start": bool(args.opt_warm),
                   "active_cells": prob.n_active_cells, "tolerance": args.opt_tol},
        "runs": runs,
    }
    if "mixed" in runs and "fp64" in runs:
        a, b = runs["mixed"]["final_compliance"], runs["fp64"]["final_compliance"]
        out["compliance_rel_diff_mixed_vs_fp64"] = abs(a - b) / abs(b)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
