"""Benchmark: time to reach ||f - K u|| / ||f|| <= 1e-8 with FETI-DP on one MI355X.

One step = one full numeric solve of the workload with inputs resident in HBM:
element integration + subdomain assembly + batched partial Cholesky (MFMA) +
coarse problem + RHS condensation + PCPG + recovery + FP64 global residual
check.  The symbolic setup (edge numbering, plans) is done once, untimed.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cube96j]

Prints one JSON line.  `value` is seconds per solve (lower is better) of the
fastest variant that reaches 1e-8; all variants are listed under `variants`, each
with its status (a variant that fails or falls back is reported, not dropped).
The default workload is the 96^3-hex contrast cube the metric is quoted on.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_STEPS = 2  # untimed event-profiled solves per variant (roofline and breakdown)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
# event-profile classes -> the kernel a rocprofv3 trace names for their launches
KERNEL_OF_CLASS = {"tile_column": 1, "coarse_column": 1, "tile_schur": 0, "coarse_schur": 0,
                   "tile_diag": 2, "coarse_diag": 2}
COARSE_CLASSES = ("coarse_column", "coarse_schur", "coarse_diag")
# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and WRITE_SIZE
# passes over the same workload, each scaled by what the counter reads for a 512 MiB
# device-to-device copy (FETCH x2.000, WRITE x1.000); the counters cannot be read live.
MEASURED_TRAFFIC = {("cube48j", "k_tile_update<double,1>"): 0.3279e9,
                    ("cube96j", "k_tile_update<double,1>"): 0.7324e9}
WORKLOAD_TEXT = {
    "cube48j": "48^3-hex Kuhn cube (663,552 tets, 753,552 free edges), 6x6x6 jittered "
               "subdomains 6-10 hexes wide, nu=1, kappa=1e-2, 2 Gaussian current loops, PEC",
    "cube16": "16^3-hex Kuhn cube (24,576 tets, 26,416 free edges), 2x2x2 subdomains, "
              "nu=1, kappa=1e-2, 2 Gaussian current loops, PEC",
    "cube64c": "64^3-hex Kuhn cube, 8x8x8 subdomains, log-uniform nu/kappa contrast + sphere",
    "cube32l": "32^3-hex Kuhn cube, 4x4x4 subdomains, nu log-uniform per subdomain, 1-4 loops",
    "cube96j": "96^3-hex Kuhn cube (5,308,416 tets), 12x12x12 jittered subdomains 6-10 hexes "
               "wide, log-uniform nu/kappa contrast + sphere",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cube96j")
    ap.add_argument("--variants", default="fp64,fp32,fp64+learned,fp32+learned",
                    help="comma list of fp64 | fp32, each optionally +learned (lambda0 network)")
    ap.add_argument("--cpu-sample-iters", type=int, default=0,
                    help="reference CG iterations timed for the CPU baseline (0: sized for "
                         "about 15 s of CPU work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--max-iterations", type=int, default=1000)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the (untimed) launch-variant autotuner")
    ap.add_argument("--graphs", type=int, choices=(0, 1), default=1,
                    help="PCPG iterations as captured HIP graphs (1, the solver default) or eager "
                         "launches (0); the autotuner, when on, chooses for itself")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="copy this many MiB device-to-device once before the variants, so a "
                         "counter profile sees a copy of known size")
    return ap.parse_args()


def run_variant(prob, variant, steps, warmup, max_iterations=1000, autotune=True, graphs=True,
                launch=None, keep_system=False):
    """Time `steps` factor + solve steps of one variant.  `launch` = (side_stream, graphs,
    tile_waves, coarse_group) chosen by an earlier variant's autotuner (reused instead of
    tuning again)."""
    import torch
    from emagddm_amd import FETIDPSolver, SolverStatus
    from emagddm_amd.lambda0 import LambdaMPNN, default_weights
    precision, _, extra = variant.partition("+")
    s = FETIDPSolver(precision=precision, profile=False,  # timed run: no events
                     max_iterations=max_iterations, graphs=graphs)
    s.setup_problem(prob)
    model = None
    if extra == "learned":
        path = default_weights(prob.name)
        if path is None:
            raise RuntimeError(f"no trained lambda0 weights for {prob.name}")
        model = LambdaMPNN.load(path)
        s.set_lambda0_model(model)
    elif extra:
        raise RuntimeError(f"unknown variant {variant}")
    tuning = None
    s.set_rhs_tet_J(prob.jq)          # inputs now resident in HBM
    if launch is not None:
        s.set_config_values(side_stream=launch[0], graphs=launch[1], tile_waves=launch[2])
        if launch[3] > 0:
            s.set_config_values(coarse_group=launch[3])
    elif autotune:  # untimed, like the symbolic setup: picks the fastest launch variant
        tuning = s.autotune(repetitions=1)
    for _ in range(warmup):
        s.factor()
        r = s.solve(copy_x=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for _ in range(steps):            # the solution stays in HBM
        s.factor()
        results.append(s.solve(copy_x=False))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ok = all(x.status == SolverStatus.SUCCESS and x.residual_norm <= 1e-8 for x in results)
    # the same steps with the solution copied to host memory (PCIe-inclusive, reported
    # beside the value); its last result also provides x for the comparisons
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(steps):
        s.factor()
        r = s.solve()
        ok = ok and r.status == SolverStatus.SUCCESS and r.residual_norm <= 1e-8
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    # untimed profiled passes: per-kernel-class HIP events on the solver's stream
    s.set_profile(True)
    s.reset_profile()
    for _ in range(PROFILE_STEPS):
        s.factor()
        s.solve()
    prof = s.profile()
    s.set_profile(False)
    cfg = s.get_config()
    ref_hist = s.refinement_history()
    # the free-dof system for the CPU baseline, read back before the handle's device
    # memory is released (the next variant needs it at 96^3)
    csr = s.global_system() if keep_system else None
    launch_used = (cfg.side_stream, cfg.graphs, cfg.tile_waves, cfg.coarse_group)
    s.clear()
    if model is not None:
        model.close()
    return dict(csr=csr, ref_hist=ref_hist, tuning=tuning, result=r, ok=ok,
                s_per_solve=(t1 - t0) / steps,
                s_per_solve_host_x=(t3 - t2) / steps,
                launch=launch_used,
                factor_ms=r.factor_time_ms, solve_ms=r.solve_time_ms,
                iterations=r.iterations, refinements=r.refinements,
                residual=r.residual_norm, profile=prof, steps=steps)


def roofline(v, precision, workload):
    """The dominant kernel: the tile kernel (mode 0 Schur update, 1 column step, 2/3
    diagonal factorization) with the most event time, local and coarse launches together
    as a kernel trace counts them.  The coarse problem is always FP64."""
    prof = v["profile"]
    by_mode = {}
    for cls, mode in KERNEL_OF_CLASS.items():
        fp64 = precision == "fp64" or cls in COARSE_CLASSES
        key = (mode, fp64)
        ms, n = prof[cls]
        a = by_mode.setdefault(key, [0.0, 0, 0.0, 0.0])
        a[0] += ms
        a[1] += n
        a[2] += prof["flops"].get(cls, 0.0) * PROFILE_STEPS
        a[3] += prof.get("useful_flops", {}).get(cls, 0.0) * PROFILE_STEPS
    (mode, fp64), (ms, launches, flops, useful) = max(by_mode.items(), key=lambda kv: kv[1][0])
    if launches == 0 or ms == 0:
        return None
    per_launch_flops = flops / launches
    avg_ms = ms / launches
    achieved = per_launch_flops / (avg_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS["fp64" if fp64 else "fp32"]
    kernel = f"k_tile_update<{'double' if fp64 else 'float'},{mode}>"
    traffic = MEASURED_TRAFFIC.get((workload, kernel))
    ridge = peak * 1e12 / (HBM_PEAK_GBS * 1e9)  # flop/B where the two ceilings meet
    intensity = per_launch_flops / traffic if traffic else None
    bound = "mfma" if intensity is None or intensity >= ridge else "hbm"
    out = dict(bound=bound, achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
               frac=round(achieved / peak, 4), traffic=traffic, kernel=kernel,
               flop_unit="executed 64x64 tile products (padding and full diagonal tiles "
                         "included; the TRSM as a product by the stored lower-triangular "
                         "inverse, 160 of its 256 MFMAs issued)",
               launches_per_step=launches // PROFILE_STEPS,
               flops_per_launch=per_launch_flops, avg_launch_ms=avg_ms,
               # the same launches counted in useful flops: real rows only, triangular
               # TRSM, lower half of the diagonal-tile share
               useful_flops_per_launch=useful / launches,
               useful_frac=round(useful / launches / (avg_ms * 1e-3) / 1e12 / peak, 4))
    if intensity is not None:
        out.update(flop_per_byte=round(intensity, 2), ridge_flop_per_byte=round(ridge, 2),
                   hbm_gbs=round(traffic / (avg_ms * 1e-3) / 1e9, 1))
    # the same kernel split by the matrices its launches work on: the batched subdomain
    # launches and the coarse matrix's dependent single-matrix chain
    split = {}
    for cls, m in KERNEL_OF_CLASS.items():
        if m != mode:
            continue
        ms_c, n_c = prof[cls]
        if n_c == 0 or ms_c == 0:
            continue
        fl = prof["flops"].get(cls, 0.0) * PROFILE_STEPS
        split[cls] = dict(launches_per_step=n_c // PROFILE_STEPS, avg_launch_ms=round(ms_c / n_c, 4),
                          achieved=round(fl / (ms_c * 1e-3) / 1e12, 2),
                          frac=round(fl / (ms_c * 1e-3) / 1e12 / peak, 4))
    if split:
        out["by_matrix_class"] = split
    return out


def cpu_baseline(prob, system, sample_iters):
    """Reference CGSolver (Jacobi, tol 1e-8, x0 = 0) from oracle/_ref on the same free-dof
    system (the GPU-assembled CSR, bit-identical in structure and equal to 1e-13 to the
    oracle's), timed on a bounded prefix of its iterations.  When the full-run iteration
    count was measured (tests/golden/<workload>_cg.json) the time is scaled to it; when
    that run stopped short of 1e-8 the scaled time is reported as a lower bound; with no
    record only the per-iteration cost is reported."""
    import ctypes as C
    import numpy as np
    gpath = os.path.join(ROOT, "tests", "golden", f"{prob.name}_cg.json")
    gold = json.load(open(gpath)) if os.path.exists(gpath) else None
    rp, col, val, f = system
    n = f.size
    if sample_iters <= 0:
        # about 15 s: one iteration costs ~21 ms per 753,552 rows on one core of the build
        # container and about half that on the GPU hosts
        sample_iters = int(max(20, min(1200, 30.0 / (21e-3 * n / 753552))))
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libref_solvers.so")
    if os.path.exists(lib_path):
        lib = C.CDLL(lib_path)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        x = np.zeros(n)
        it, res, sm = C.c_int(), C.c_double(), C.c_double()
        lib.ref_solve_cg(n, ip(rp), ip(col), dp(val), dp(f), C.c_double(1e-8), sample_iters,
                         dp(x), C.byref(it), C.byref(res), C.byref(sm))
        kind, ms, its, rres = "reference", sm.value, it.value, res.value
    else:
        import scipy.sparse as sp
        from oracle import solvers
        A = sp.csr_matrix((val, col, rp), shape=(n, n))
        t0 = time.perf_counter()
        _, _, its, rres = solvers.pcg_jacobi(A, f, 1e-8, sample_iters)
        kind, ms = "port", (time.perf_counter() - t0) * 1e3
    per_it = ms / max(its, 1) / 1e3
    if gold is not None and int(gold.get("status", 1)) == 0:
        full = int(gold["iterations"])
        return dict(value=round(per_it * full, 3), unit="s to 1e-8 (Jacobi-PCG, reference CGSolver)",
                    cores=1, kind=kind,
                    sample=f"{its} of {full} CG iterations on the same {n}-dof system "
                           f"({ms / 1e3:.2f} s measured, scaled by the full-run count "
                           f"measured in the build container)")
    if gold is not None:
        # the reference path did not reach 1e-8 within the recorded count: its time to
        # 1e-8 is at least that count times the measured cost of one iteration
        lim = int(gold["iterations"])
        return dict(value=round(per_it * lim, 3),
                    unit="s to 1e-8, lower bound (Jacobi-PCG, reference CGSolver)",
                    cores=1, kind=kind,
                    sample=f"{its} CG iterations on the same {n}-dof system ({ms / 1e3:.2f} s, "
                           f"{per_it * 1e3:.1f} ms each) times {lim}: the full run had not reached "
                           f"1e-8 after {lim} iterations in the build container (relative "
                           f"residual {float(gold['residual']):.2e} then)")
    note = "full-run iteration count not measured"
    return dict(value=round(per_it, 5), unit="s per Jacobi-PCG iteration (reference CGSolver)",
                cores=1, kind=kind,
                sample=f"{its} CG iterations on the same {n}-dof system ({ms / 1e3:.2f} s, "
                       f"relative residual {rres:.2e} after them); {note}")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from emagddm_amd import workloads
    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    prob = workloads.make_workload(a.workload)
    if a.calibration_copy_mib > 0:
        n = a.calibration_copy_mib << 17  # float64 elements
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    from emagddm_amd import FETIDPSolver
    variants, failed, launch, system = {}, {}, None, None
    for prec in a.variants.split(","):
        try:
            v = run_variant(prob, prec, a.steps, a.warmup, a.max_iterations, not a.no_autotune,
                            bool(a.graphs), launch,
                            keep_system=not a.no_cpu_baseline and system is None)
        except RuntimeError as e:
            print(f"[bench] {prec}: failed: {e}", file=sys.stderr, flush=True)
            failed[prec] = str(e)
            continue
        launch = launch or v["launch"]  # the first variant's tuned launch choices serve all
        system = system or v["csr"]
        v["csr"] = None
        variants[prec] = v
        print(f"[bench] {prec}: {v['s_per_solve']:.4f} s/solve its={v['iterations']} "
              f"ref={v['refinements']} res={v['residual']:.2e} ok={v['ok']} "
              f"fell_back={v['result'].fell_back_to_fp64}", file=sys.stderr, flush=True)
        prof = {k: (round(ms / PROFILE_STEPS, 3), n // PROFILE_STEPS) for k, (ms, n) in
                ((k, v["profile"][k]) for k in FETIDPSolver.PROFILE_SLOTS)}
        print(f"[bench] {prec} per-step event ms (ms, launches): {prof}", file=sys.stderr,
              flush=True)
        fl = v["profile"].get("flops", {})
        tfs = {k: round(fl[k] / (v["profile"][k][0] / PROFILE_STEPS * 1e-3) / 1e12, 1)
               for k in fl if fl[k] > 0 and v["profile"][k][0] > 0}
        print(f"[bench] {prec} tile classes, achieved TFLOP/s (one stream): {tfs}",
              file=sys.stderr, flush=True)
        print(f"[bench] {prec} refinement residuals: {v['ref_hist']}",
              file=sys.stderr, flush=True)
    if not variants:
        print(json.dumps({"error": "no variant ran", "failed": failed}))
        sys.exit(1)
    # a variant that fell back to FP64 is reported but never stands for its precision
    good = {k: v for k, v in variants.items() if v["ok"] and not v["result"].fell_back_to_fp64}
    best = min(good or variants, key=lambda k: variants[k]["s_per_solve"])
    vb = variants[best]
    import numpy as np
    xs = {k: v["result"].x for k, v in variants.items()}
    agree = None
    if "fp64" in xs:
        ref = xs["fp64"]
        agree = {k: float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
                 for k, x in xs.items() if k != "fp64"}
    ran_fp32 = lambda k: k.startswith("fp32") and not variants[k]["result"].fell_back_to_fp64
    roof = roofline(vb, "fp32" if ran_fp32(best) else "fp64", a.workload)
    t64 = variants["fp64"]["s_per_solve"] if "fp64" in variants else None
    out = {
        "metric": "time-to-1e-8 resid: learned+FP32-refined vs plain FP64 FETI-DP; PCPG iters",
        "value": round(vb["s_per_solve"], 5),
        "unit": "s per solve to ||f-Ku||/||f|| <= 1e-8",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(vb["s_per_solve"] * 1e3, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64" if ran_fp32(best) else "f64",
        "data": "synthetic (seeded Kuhn cube, loop sources)",
        "config": {"workload": a.workload, "description": WORKLOAD_TEXT.get(a.workload, a.workload),
                   "variant": best},
        "variants": {k: dict(s_per_solve=round(v["s_per_solve"], 5),
                             s_per_solve_host_x=round(v["s_per_solve_host_x"], 5),
                             status=v["result"].status.name,
                             factor_ms=round(v["factor_ms"], 2), solve_ms=round(v["solve_ms"], 2),
                             pcpg_iterations=v["iterations"], refinements=v["refinements"],
                             residual=v["residual"], reached_1e8=v["ok"],
                             fell_back_to_fp64=v["result"].fell_back_to_fp64,
                             learned_lambda0=v["result"].learned_guess,
                             precond_fp64=v["result"].precond_fp64,
                             deluxe=v["result"].deluxe,
                             guess_ms=round(v["result"].guess_time_ms, 3),
                             initial_dual_residual=v["result"].initial_dual_residual,
                             launch=list(v["launch"]),
                             autotune=(None if v["tuning"] is None else dict(
                                 chosen=list(v["launch"]),
                                 variants_ms={f"side{a}_graph{b}_waves{c}_group{g}": round(d, 2)
                                              for a, b, c, g, d in v["tuning"]})))
                     for k, v in variants.items()},
        "failed_variants": failed,
        # the metric's own ratio: plain FP64 time over the learned + FP32-refined time
        "speedup_learned_fp32_vs_plain_fp64": (
            round(t64 / variants["fp32+learned"]["s_per_solve"], 3)
            if t64 and "fp32+learned" in variants else None),
        "speedup_vs_plain_fp64": round(t64 / vb["s_per_solve"], 3) if t64 else None,
        "solution_rel_l2_vs_fp64": agree,
        "solves_per_s": round(1.0 / vb["s_per_solve"], 3),
        "roofline": roof,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob, system, a.cpu_sample_iters)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
