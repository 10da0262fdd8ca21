"""Benchmark: MG-PCG solve time to 1e-6 relative residual (mixed vs FP64) and
stiffness-apply throughput on the 256^3 narrow-band sphere shell.

A step is one complete mixed-precision solve (FP32 MG-PCG inside an FP64
defect-correction loop) of K(rho) u = f from a zero initial guess to
||P(f - K u)||_inf <= 1e-6 ||P f||_inf, inputs resident in HBM.  Prints one
JSON line on stdout; logs go to stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--levels 5]
                  [--no-cpu] [--cpu-seconds 20]

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
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--max-iterations", type=int, default=500, help="PCG iterations per pass (FP64 solve: 4x)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the FP32 apply and a 512 MiB copy (5 launches each) for rocprofv3 --pmc passes")
    ap.add_argument("--opt-loop", action="store_true", help="time the OC optimization loop on the 512x256x256 band")
    ap.add_argument("--opt-iters", type=int, default=20)
    ap.add_argument("--opt-precisions", default="mixed,fp64")
    ap.add_argument("--opt-band-start", type=int, default=5)
    ap.add_argument("--opt-tol", type=float, default=1e-4,
                    help="solve tolerance in the loop (the reference optimizer's cg_tolerance default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is described in DESIGN.md and not implemented.", file=sys.stderr)
        sys.exit(2)
    if args.opt_loop:
        opt_loop(args)
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
    t_ap32 = s.bench_apply(0, args.apply_reps)  # the PCG operator: FP32 in/out, FP64 row sums
    t_ap64 = s.bench_apply(1, args.apply_reps)
    # the same apply in difference form with FP32 sums, then with direct FP32 sums
    os.environ["SPG_DIFF_SUMS"] = "1"
    t_ap32d = s.bench_apply(0, args.apply_reps)
    os.environ["SPG_DIFF_SUMS"] = "0"
    os.environ["SPG_ACC64_MODES"] = "0"
    t_ap32f = s.bench_apply(0, args.apply_reps)
    del os.environ["SPG_ACC64_MODES"]
    del os.environ["SPG_DIFF_SUMS"]
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

    # accuracy against a tight FP64 reference solve (10^-10) on the same
    # problem: relative L2 displacement error and relative compliance error
    acc = accuracy_report(s, args, pr)
    log(f"[bench] vs FP64 reference: mixed u err {acc['mixed']['u_rel_l2']:.2e}, c err {acc['mixed']['c_rel']:.2e}; "
        f"fp64 u err {acc['fp64']['u_rel_l2']:.2e}, c err {acc['fp64']['c_rel']:.2e}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, info["iterations"], args.cpu_seconds)

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
        "apply_us_fp32_difference_form": t_ap32d * 1e6,
        "apply_us_fp32_fp32_sums": t_ap32f * 1e6,
        "apply_mdof_s_fp32_fp32_sums": 3 * nn / t_ap32f / 1e6,
        "apply_us_fp64": t_ap64 * 1e6,
        "vcycle_ms_fp32": t_vc32 * 1e3,
        "vcycle_ms_fp64": t_vc64 * 1e3,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_apply_lds<float, double, 2> (level-0 stiffness apply of the PCG operator: FP32 in/out, "
                      "FP64 row sums)",
            "achieved": gbs32,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs32 / HBM_PEAK_GBS,
            "traffic": measured_traffic("k_apply_lds<float, double, 2>", nn),
        },
        "roofline_gs_sweep": {
            "bound": "hbm",
            "kernel": "k_gs0_colour4<float, 8> (level-0 Gauss-Seidel, 8 colour launches per sweep)",
            "achieved": GS_SWEEP_BYTES_PER_NODE_FP32 * nn / t_gs32 / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": GS_SWEEP_BYTES_PER_NODE_FP32 * nn / t_gs32 / 1e9 / HBM_PEAK_GBS,
            "us_per_sweep": t_gs32 * 1e6,
            "traffic_per_sweep": (lambda t: None if t is None else 8 * t)(
                measured_traffic("k_gs0_colour4<float, 8>", nn)),
        },
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_reference": acc,
    }
    print(json.dumps(out), flush=True)


def accuracy_report(s, args, pr):
    """Relative L2 displacement error and relative compliance error of the
    mixed and the FP64 solve (both to args.tol) against an FP64 solve to
    10^-10; the compliance is the sensitivity kernel's objective
    (sum_e u_e^T K0 u_e E(rho_e))."""
    import torch

    n = s.nblocks * 64 * 3
    dc = torch.empty(s.nblocks * 64, dtype=torch.float64, device="cuda")

    def run(**kw):
        _, info = s.solve(want_output=False, raise_on_error=False, print_residuals=pr, **kw)
        u = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        s.displacement_device(u.data_ptr())
        c = s.sensitivity_device(dc.data_ptr(), u_ptr=u.data_ptr())
        return u, c, info

    u_ref, c_ref, i_ref = run(tolerance=1e-10, precision=1, max_passes=1, max_iterations=8 * args.max_iterations)
    out = {"reference": {"tolerance": 1e-10, "iterations": i_ref["iterations"], "converged": bool(i_ref["converged"]),
                         "compliance": c_ref}}
    for name, kw in (("mixed", dict(precision=0, max_passes=30, max_iterations=args.max_iterations)),
                     ("fp64", dict(precision=1, max_passes=1, max_iterations=4 * args.max_iterations))):
        u, c, info = run(tolerance=args.tol, **kw)
        out[name] = {"u_rel_l2": float(torch.linalg.vector_norm(u - u_ref) / torch.linalg.vector_norm(u_ref)),
                     "c_rel": abs(c - c_ref) / abs(c_ref), "iterations": info["iterations"]}
    return out


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


def cpu_baseline(prob, gpu_iterations, budget_s):
    """Oracle (C, OpenMP, FP64) on the same problem: time PCG iterations for
    about `budget_s` seconds and scale to the GPU's iteration count."""
    from oracle import oracle as O

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    o = O.Oracle(prob)
    setup = time.perf_counter() - t0
    _, sec1 = o.time_pcg_iterations(1)
    k = max(1, min(50, int(budget_s / max(sec1, 1e-3))))
    it, sec = o.time_pcg_iterations(k)
    per_it = sec / max(it, 1)
    log(f"[bench] cpu oracle: setup {setup:.1f}s, {per_it:.3f}s per PCG iteration ({k} timed) on {cores} threads")
    return {
        "value": per_it * gpu_iterations,
        "unit": "s",
        "cores": cores,
        "kind": "port",
        "sample": f"{k} FP64 MG-PCG iterations of the same shell problem timed on the host; value = time per "
                  f"iteration x {gpu_iterations} iterations (the GPU mixed solve's count); oracle setup "
                  f"{setup:.1f}s not included",
        "setup_s": setup,
        "per_iteration_s": per_it,
    }


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
                        tolerance=args.opt_tol, precision=prec)
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
                f"pcg {r['iterations']} fem {r['fem_s']:.2f}s band {r['band_s']:.2f}s filter+oc "
                f"{r['filter_oc_s']:.2f}s blocks {r['fem_blocks']} fat {r['fat_blocks']} rebuilt {r['rebuilt']}")
        total = time.perf_counter() - t0
        runs[name] = {
            "s_per_step": total / args.opt_iters,
            "setup_s": setup,
            "final_compliance": recs[-1]["compliance"],
            "solve_s_mean": sum(r["solve_s"] for r in recs) / len(recs),
            "pcg_iterations": [r["iterations"] for r in recs],
            "rebuilds": sum(bool(r["rebuilt"]) for r in recs),
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
                   "band_update_from_iteration": args.opt_band_start,
                   "active_cells": prob.n_active_cells, "tolerance": args.opt_tol},
        "runs": runs,
    }
    if "mixed" in runs and "fp64" in runs:
        a, b = runs["mixed"]["final_compliance"], runs["fp64"]["final_compliance"]
        out["compliance_rel_diff_mixed_vs_fp64"] = abs(a - b) / abs(b)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
