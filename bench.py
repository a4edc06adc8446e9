"""Benchmark of the hex8 Neo-Hooke hot path on one MI355X.

Workload (default): 3D hex8 Neo-Hooke block, 64x64x64 elements (823,875 DOFs), unit cube, E=1000, nu=0.3,
seeded +-5% interior jitter (seed 0), bottom face clamped, seeded top-face traction, FP64.

One step = the residual R and tangent K of one Newton state (the pair NewtonRaphson evaluates per state),
FP64, FULL Dirichlet treatment, summed into the 3x3-block pattern (sliced-ELL storage), inputs resident in HBM, through
ik_assemble (the tangent's element kernel and gather, then the residual pass; one error check and one
host synchronisation for the pair).
value = elements assembled per second (Melem/s) over the timed steps. The same state assembled by the two
separate calls (ik_vector, then ik_matrix) is timed beside it.

Untimed before the timed steps: a clock-settling pre-pass (assembly steps for --settle-ms of wall time, reported
as clock_settle), then --warmup steps.

Also reported (same run): the element kernel's roofline figures (HIP events on the library stream; HBM
traffic per launch from the committed rocprofv3 counter summary) and the whole step's FP64 fraction, the gather and
SpMV bandwidths, config c2's 5-step FP64 load path, the Newton comparison (40 load steps to ||R|| <= 1e-8: the
plain FP64 form on the first evaluation seed, the mixed form with the constant, secant and learned predictors, with
the block-Jacobi and with the two-level preconditioner, on 8 evaluation seeds never used for training or tuning) and
the CPU baseline (the oracle restatement, faithful and optimized, on bounded samples on the host cores, plus config
c1 in full).
--tune-eta chooses the forcing-term cap of each mixed predictor on training seeds.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import ikarus_amd as ia  # noqa: E402
from ikarus_amd import _capi as capi  # noqa: E402
from ikarus_amd import problems as pb  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
FLOP_PER_ELEM_KR = 86176.0   # 8 Gauss points x 10,772 (dense B^T C B + kinematics + material + kg + B^T S)
FLOP_PER_ELEM_K = FLOP_PER_ELEM_KR - 8 * 288.0  # the tangent kernel alone: less B^T S (6x24 by 6) per point


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    """Host CPU model, its logical cores and the threads the optimized CPU variant uses (the GPU box's share of
    host cores is 16, whatever the machine reports)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    return model, nproc, max(1, min(threads, nproc, 16))


def _oracle(p, faithful, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    if len(p.face_elem):
        o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction[:p.dim])
    o.set_faithful(faithful)
    o.set_threads(threads)
    return po, o


def cpu_assembly(seconds, faithful, threads, n=16):
    """K+R assembly rate of the CPU oracle on an n^3 sample, repeated until `seconds` have passed."""
    p = pb.hex_block((n, n, n), seed=0)
    _, o = _oracle(p, faithful, threads)
    u = 0.005 * np.random.default_rng(0).standard_normal(p.ndof) * (1 - p.fixed)
    o.pattern()
    reps = 0
    t0 = time.perf_counter()
    while True:
        o.matrix(u, 1.0)
        o.vector(u, 1.0)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return p.ne * reps / dt / 1e6, reps, dt


def cpu_load_path(p, steps, faithful, threads, direct):
    """One LoadControl run of the CPU oracle (Newton to ||R|| <= 1e-8 with the banded direct solve, or Jacobi-CG
    to 1e-10 relative as the plain FP64 device form does)."""
    po, o = _oracle(p, faithful, threads)
    t0 = time.perf_counter()
    r = o.load_control(steps, 0.0, 1.0, tol=1e-8, solver=po.DIRECT if direct else po.JACOBI_CG, cg_tol=1e-10,
                       threads=threads)
    dt = time.perf_counter() - t0
    return dict(seconds=round(dt, 3), seconds_per_step=round(dt / steps, 4), success=bool(r["success"]),
                newton_iterations=int(r["total"]), cg_iterations=int(r["cg_iterations"]))


def cpu_baseline(seconds, c2_sample=12):
    """The CPU oracle (the reference itself cannot be built here: it needs DUNE, Eigen >= 3.4.90 and SuiteSparse,
    none of which is installed), timed on the host cores in two variants:
    * faithful: the reference's cost structure, one thread (material tangent re-evaluated for every node pair of
      the element as nonlinearelastic.hh:390-398 does, serial element loop and scatter, single-thread CG);
    * optimized: material tangent once per Gauss point, elements in a node-conflict coloring with one OpenMP loop
      per color, threaded Jacobi-CG, on `threads` host threads.
    Samples: K+R assembly of a 16^3 hex8 NH mesh (the headline unit, Melem/s); config c1 in full (2D plane-strain
    NH cantilever 64x16, 10 load steps, banded LDL^T direct solve standing in for the reference's sparse direct
    solver); config c2's load path (5 load steps, Jacobi-CG) on a c2_sample^3 mesh."""
    model, nproc, threads = cpu_info()
    log(f"[bench] cpu: {model}, {nproc} logical cores, optimized variant on {threads} threads")
    f_rate, f_reps, f_dt = cpu_assembly(seconds, True, 1)
    o_rate, o_reps, o_dt = cpu_assembly(seconds / 2, False, threads)
    log(f"[bench] cpu assembly: faithful {f_rate:.4f} Melem/s, optimized {o_rate:.4f} Melem/s")
    c1 = pb.cantilever2d(seed=0)
    c2 = pb.hex_block((c2_sample,) * 3, seed=0)
    out = dict(value=round(o_rate, 6), unit="Melem/s", cores=threads, kind="port",
               sample=f"{o_reps}x K+R assembly of a 16^3 hex8 NH mesh (4096 elements), optimized CPU oracle (tangent "
                      f"once per Gauss point, colored elements, OpenMP) on {threads} threads, {o_dt:.1f} s",
               cpu_model=model, logical_cores=nproc,
               reference_build="not available: the reference needs DUNE, dune-localfefunctions, Eigen >= 3.4.90 and "
                               "SuiteSparse, none installed; the oracle restates its arithmetic",
               faithful={"value": round(f_rate, 6), "unit": "Melem/s", "cores": 1,
                         "sample": f"{f_reps}x K+R assembly of the same mesh, reference cost structure (tangent per "
                                   f"node pair), serial, {f_dt:.1f} s"},
               c1={"workload": f"2D plane-strain NH cantilever 64x16 Q1 over 4x1, E=1000, nu=0.3, traction "
                               f"{[round(t, 4) for t in c1.traction[:2]]}, 10 load steps to ||R|| <= 1e-8, banded LDL^T",
                   "faithful_1_thread": cpu_load_path(c1, 10, True, 1, True),
                   "optimized": cpu_load_path(c1, 10, False, threads, True)},
               c2_sample={"workload": f"c2 load path on a {c2_sample}^3 hex8 NH sample (5 load steps, Jacobi-CG to "
                                      "1e-10 per Newton iteration, ||R|| <= 1e-8)",
                          "faithful_1_thread": cpu_load_path(c2, 5, True, 1, False),
                          "optimized": cpu_load_path(c2, 5, False, threads, False)})
    log(f"[bench] cpu c1: {out['c1']}; c2 sample: {out['c2_sample']}")
    return out


NEWTON_MODES = {
    # plain FP64 form: FP64 tangent, Jacobi-PCG to 1e-10 per Newton iteration, constant predictor (LoadControl)
    "fp64": dict(precision=capi.FP64, kind=capi.PCG_JACOBI_FP64, rtol=1e-10, forcing=0.0, pred=capi.PRED_CONSTANT),
    # same precision, inexact Newton + secant predictor (isolates the algorithmic part of the gain)
    "fp64_inexact_secant": dict(precision=capi.FP64, kind=capi.PCG_JACOBI_FP64, rtol=1e-10, forcing=0.9,
                                pred=capi.PRED_LINEAR),
    # mixed form: FP32 tangent storage + FP32 block-Jacobi PCG with FP64 refinement, inexact Newton; constant,
    # secant and learned predictors. eta_max (the forcing term's cap) of each predictor is the best of
    # ETA_CANDIDATES on the training seeds TUNE_SEEDS (bench.py --tune-eta), never on an evaluation seed.
    # (64^3, 40 load steps, seeds 1000-1003: constant 0.03, secant 0.1, learned 0.01)
    "mixed_constant": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.9,
                           pred=capi.PRED_CONSTANT, eta_max=0.03),
    "mixed_secant": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.9,
                         pred=capi.PRED_LINEAR, eta_max=0.1),
    "mixed_mlp": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.9,
                      pred=capi.PRED_MLP, eta_max=0.01),
    # the mixed form with the two-level preconditioner (block Jacobi + rigid-body-mode coarse correction over
    # about 729 box aggregates) in the FP32 PCG; eta_max tuned the same way (64^3, 40 load steps, seeds
    # 1000-1003: constant 0.1, secant 0.01, learned 0.001)
    "tl_constant": dict(precision=capi.FP32, kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, forcing=0.9,
                        pred=capi.PRED_CONSTANT, eta_max=0.1),
    "tl_secant": dict(precision=capi.FP32, kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, forcing=0.9,
                      pred=capi.PRED_LINEAR, eta_max=0.01),
    "tl_mlp": dict(precision=capi.FP32, kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, forcing=0.9,
                   pred=capi.PRED_MLP, eta_max=0.001),
    # precision alone: the plain form's Newton and predictor with the FP32 tangent and refined FP32 PCG
    "mixed_exact": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.0,
                        pred=capi.PRED_CONSTANT),
}


TUNE_SEEDS = (1000, 1001, 1002, 1003)     # training seeds (the network was trained on paths of 1000-1255)
EVAL_SEEDS = tuple(range(2003, 2011))      # evaluation seeds: never used for training or tuning
ETA_CANDIDATES = (0.1, 0.03, 0.01, 1e-3)


def newton_path(p, mode, steps, mlp_params=None, eta_max=None):
    m = dict(NEWTON_MODES[mode])
    if eta_max is not None:
        m["eta_max"] = eta_max
    dev = ia.Device(p)
    if m["pred"] == capi.PRED_MLP:
        dev.check(capi.lib().ik_mlp_set_params(dev.h, capi.ptr(mlp_params), capi.HOST))
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = steps, 0.0, 1.0, m["pred"]
    s = ia.nr_settings(tol=1e-8, precision=m["precision"],
                       linear=ia.solver_opts(m["kind"], rtol=m["rtol"], inner_rtol=1e-3, max_refine=30),
                       forcing=m["forcing"], eta_max=m.get("eta_max", 0.0))
    dev.synchronize()
    t0 = time.perf_counter()
    info, iters, res = dev.load_control(cfg, s)
    dev.synchronize()
    dt = time.perf_counter() - t0
    u, _ = dev.get_requirement()
    dev.close()
    return dict(seconds=round(dt, 4), success=bool(info.success), newton_iterations=int(info.total_iterations),
                per_step=[int(i) for i in iters], linear_iterations=int(info.linear_iterations),
                final_residual=float(res[info.solves - 1]), max_step_residual=float(np.max(res[:info.solves])),
                fallbacks=int(info.fallbacks)), u


MIXED_MODES = ("mixed_constant", "mixed_secant", "mixed_mlp", "tl_constant", "tl_secant", "tl_mlp")


def warm_libraries():
    """One small two-level solve, untimed: rocBLAS / rocSOLVER load their kernels on a process's first calls (about
    0.25 s), which no later Newton path pays (the 16^3 coarse size, 3,072 DOFs, warms what the 64^3 paths' 4,374
    use: their traces show no loading gap)."""
    p = pb.hex_block((16, 16, 16), seed=0)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    dev.matrix(capi.DBC_FULL, capi.FP32)
    t0 = time.perf_counter()
    dev.solve(np.ones(p.ndof) * (1 - p.fixed), ia.solver_opts(capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-6))
    dev.close()
    return time.perf_counter() - t0


def tune_eta(n, steps, mlp, modes=MIXED_MODES):
    """eta_max of each mixed predictor chosen on the training seeds: total seconds over TUNE_SEEDS per candidate."""
    warm_libraries()
    table = {}
    for mode in modes:
        table[mode] = {}
        for eta in ETA_CANDIDATES:
            tot, its = 0.0, 0
            for seed in TUNE_SEEDS:
                r, _ = newton_path(pb.hex_block((n, n, n), seed=seed), mode, steps, mlp, eta_max=eta)
                assert r["success"], (mode, eta, seed)
                tot += r["seconds"]
                its += r["newton_iterations"]
            table[mode][str(eta)] = {"seconds": round(tot, 3), "newton_iterations": its}
            log(f"[bench] tune {mode} eta_max={eta}: {tot:.3f} s, {its} Newton over seeds {TUNE_SEEDS}")
    best = {m: float(min(t, key=lambda k: t[k]["seconds"])) for m, t in table.items()}
    return {"seeds": list(TUNE_SEEDS), "mesh": f"{n}^3", "load_steps": steps, "table": table, "best_eta_max": best}


def evaluate_newton(n, steps, mlp, seeds):
    """The c5 comparison: the plain FP64 form and FP64 inexact + secant on the first evaluation seed, then the mixed
    form with the constant, secant and learned predictors on every evaluation seed; per-seed and median figures."""
    first = pb.hex_block((n, n, n), seed=seeds[0])
    warm_s = warm_libraries()
    log(f"[bench] rocBLAS / rocSOLVER first-call load (untimed warm-up): {warm_s:.3f} s")
    plain, us = {}, {}
    for mode in ("fp64", "fp64_inexact_secant"):
        log(f"[bench] newton {mode} seed {seeds[0]} ...")
        plain[mode], us[mode] = newton_path(first, mode, steps)
        log(f"[bench] newton {mode}: {plain[mode]['seconds']} s, {plain[mode]['newton_iterations']} Newton, "
            f"{plain[mode]['linear_iterations']} PCG")
    per_seed = {}
    for seed in seeds:
        p = pb.hex_block((n, n, n), seed=seed)
        row = {"traction": [round(t, 3) for t in p.traction]}
        for mode in MIXED_MODES:
            r, u = newton_path(p, mode, steps, mlp)
            row[mode] = {k: r[k] for k in ("seconds", "newton_iterations", "linear_iterations", "success",
                                           "max_step_residual", "fallbacks")}
            if seed == seeds[0]:
                row[mode]["rel_diff_u_vs_fp64"] = float(np.linalg.norm(u - us["fp64"]) / np.linalg.norm(us["fp64"]))
        log(f"[bench] seed {seed}: " + ", ".join(f"{m} {row[m]['seconds']} s / {row[m]['newton_iterations']} Newton"
                                                  for m in MIXED_MODES))
        per_seed[str(seed)] = row
    med = {m: {"seconds": float(np.median([per_seed[str(s)][m]["seconds"] for s in seeds])),
               "newton_iterations": float(np.median([per_seed[str(s)][m]["newton_iterations"] for s in seeds]))}
           for m in MIXED_MODES}
    ratio = {f"{a}_over_{b}": float(np.median([per_seed[str(s)][a]["seconds"] / per_seed[str(s)][b]["seconds"]
                                              for s in seeds]))
             for a, b in (("mixed_constant", "mixed_mlp"), ("mixed_secant", "mixed_mlp"), ("tl_constant", "tl_mlp"),
                          ("tl_secant", "tl_mlp"), ("mixed_mlp", "tl_mlp"), ("mixed_secant", "tl_secant"))}
    first_row = per_seed[str(seeds[0])]
    speed = {m: round(plain["fp64"]["seconds"] / first_row[m]["seconds"], 3) for m in MIXED_MODES}
    speed["fp64_inexact_secant"] = round(plain["fp64"]["seconds"] / plain["fp64_inexact_secant"]["seconds"], 3)
    return {
        "workload": f"hex8 Neo-Hooke {n}^3, {steps} load steps to ||R||<=1e-8 at every step, random traction per "
                    f"seed; evaluation seeds {seeds[0]}-{seeds[-1]} (not used for training or tuning); every time "
                    "includes the one-time pattern build of the first matrix call",
        "eta_max": {m: NEWTON_MODES[m]["eta_max"] for m in MIXED_MODES},
        "plain_fp64_seed": seeds[0], "plain": plain,
        "per_seed": per_seed, "median": med, "median_time_ratios": ratio,
        "speedup_vs_fp64_on_first_seed": speed,
        "library_warmup_s": round(warm_s, 3),
        "fastest_median": min(med, key=lambda m: med[m]["seconds"]),
        "headline": f"mixed forms vs fp64 on seed {seeds[0]}; medians over {len(seeds)} seeds for the mixed forms",
    }


def c2_load_path(n, steps):
    """Config c2's load path on the device: plain FP64 form (FP64 tangent, Jacobi-PCG to 1e-10 per Newton
    iteration, constant predictor), seed 0, per-step time."""
    r, _ = newton_path(pb.hex_block((n, n, n), seed=0), "fp64", steps)
    r["seconds_per_step"] = round(r["seconds"] / steps, 4)
    return r


def assembly_step(dev):
    """Residual and tangent of one Newton state (rx = residualFunction_(x), Ax = jacobianFunction_(x),
    newtonraphson.hh:242-243), FULL, FP64, in one ik_assemble call."""
    dev.check(capi.lib().ik_assemble(dev.h, capi.DBC_FULL, capi.FP64, None, capi.DEVICE))


def separate_step(dev):
    """The same state through the two assembler calls in the reference order: ik_vector, then ik_matrix."""
    dev.check(capi.lib().ik_vector(dev.h, capi.DBC_FULL, None, capi.DEVICE))
    dev.check(capi.lib().ik_matrix(dev.h, capi.DBC_FULL, capi.FP64))


def kernel_pass(dev, p, steps):
    """Per-kernel launch durations from HIP events on the library stream: in timing mode the element chunks of
    an assembly run back to back between one pair of events, then the gathers between another, then the
    residual pass."""
    dev.timing(True)
    dev.timing_reset()
    for _ in range(steps):
        assembly_step(dev)
    out = {}
    for name, kid in (("element", capi.KERNEL_ASSEMBLE_KR), ("gather", capi.KERNEL_GATHER_K),
                      ("residual", capi.KERNEL_ASSEMBLE_R)):
        out[name] = dev.timing_read(kid)
    dev.timing(False)
    return out


def spmv_pass(dev, p, precision, iters=50):
    """Average SpMV launch inside the PCG of the given precision (FULL K assembled first, at u = 0 so that K is
    positive definite and the CG runs all its iterations)."""
    dev.set_requirement(np.zeros(p.ndof), 1.0)
    dev.check(capi.lib().ik_matrix(dev.h, capi.DBC_FULL, precision))
    x = np.random.default_rng(1).standard_normal(p.ndof)
    kind = capi.PCG_JACOBI_FP64 if precision == capi.FP64 else capi.PCG_BJACOBI_FP32_IR
    opts = ia.solver_opts(kind, rtol=1e-30, max_iter=iters, inner_rtol=1e-30, max_refine=1)
    dev.timing(True)
    dev.timing_reset()
    dev.solve(x, opts)
    n, ms = dev.timing_read(capi.KERNEL_SPMV)
    dev.timing(False)
    return n, ms


def copy_calibration(nbytes=512 << 20, reps=2):
    """Device-to-device copies of known size (the runtime's blit kernel), the reference for the
    FETCH_SIZE / WRITE_SIZE counter scale."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def read_traffic(path):
    """HBM bytes per launch of the element kernel from the committed counter summary (None if absent)."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return int(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30, help="untimed steps after the clock-settling pre-pass")
    ap.add_argument("--settle-ms", type=float, default=300.0,
                    help="untimed assembly steps until this much wall time has passed, before the warmup steps: the "
                         "GPU clocks ramp over the first ~20-100 ms of load (reported as clock_settle)")
    ap.add_argument("--n", type=int, default=64, help="elements per direction")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--newton-steps", type=int, default=40, help="load steps of the Newton comparison")
    ap.add_argument("--eval-seeds", type=int, default=len(EVAL_SEEDS), help="evaluation seeds of the c5 comparison")
    ap.add_argument("--profile-mode", default="mixed_mlp", help="Newton mode of --profile-newton")
    ap.add_argument("--profile-newton", action="store_true",
                    help="only one mixed + learned load path (the command profiled for the MLP and PCG kernels)")
    ap.add_argument("--tune-modes", default="", help="comma-separated modes for --tune-eta (default: all mixed)")
    ap.add_argument("--tune-eta", action="store_true",
                    help="only choose eta_max per mixed predictor on the training seeds (prints the table)")
    ap.add_argument("--c2-steps", type=int, default=5, help="load steps of config c2's FP64 path on the device")
    ap.add_argument("--mlp", default=os.path.join(ROOT, "ikarus_amd", "data", "predictor_mlp_32cubed.f32"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only the copy calibration and the per-kernel passes (the command profiled by rocprofv3)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "pmc_hex8_assembly_64cubed.txt"))
    ap.add_argument("--config", choices=("c3", "c4"), default=None,
                    help="run the 128x128x64 precision variant (c3) or the 96x96x48 EAS workload (c4) instead")
    ap.add_argument("--c3-steps", type=int, default=5)
    ap.add_argument("--c4-steps", type=int, default=5, help="timed R+K assemblies of the EAS workload")
    ap.add_argument("--c4-load-steps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the hot path is measured on one GPU's "
              "share of the mesh; nothing here uses a second GPU.", file=sys.stderr)
        sys.exit(2)

    if args.config:
        out = config_precision(args) if args.config == "c3" else config_eas(args)
        print(json.dumps(out), flush=True)
        return
    if args.tune_eta:
        mlp = np.fromfile(args.mlp, dtype=np.float32)
        modes = tuple(args.tune_modes.split(",")) if args.tune_modes else MIXED_MODES
        print(json.dumps({"tune_eta": tune_eta(args.n, args.newton_steps, mlp, modes)}), flush=True)
        return
    if args.profile_newton:
        # one mixed + learned load path on the first evaluation seed (the command profiled for the learned and PCG
        # kernels)
        mlp = np.fromfile(args.mlp, dtype=np.float32)
        if NEWTON_MODES[args.profile_mode]["kind"] == capi.PCG_TWOLEVEL_FP32_IR:
            warm_libraries()
        r, _ = newton_path(pb.hex_block((args.n,) * 3, seed=EVAL_SEEDS[0]), args.profile_mode, args.newton_steps, mlp)
        print(json.dumps({"profile_newton": r}), flush=True)
        return

    n = args.n
    p = pb.hex_block((n, n, n), seed=0)
    log(f"[bench] problem {p.name}: {p.ne} elements, {p.ndof} dofs")
    dev = ia.Device(p)
    u = 0.005 * np.random.default_rng(0).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 1.0)
    nbr, nnzb = np_i64(), np_i64()
    t0 = time.perf_counter()  # the first matrix query builds the pattern (createOccupationPattern, once)
    dev.check(capi.lib().ik_matrix_bsr(dev.h, capi.ptr(nbr), capi.ptr(nnzb), None, None, None, None, None))
    pattern_s = time.perf_counter() - t0
    nnzb = int(nnzb[0]) if int(nnzb[0]) > 0 else None
    log(f"[bench] pattern build (host, once per mesh): {pattern_s:.3f} s")

    if args.profile_pass:
        copy_calibration()
        log("[bench] copy calibration done")
        kernel_pass(dev, p, args.warmup)  # warm up in the same (timed, back-to-back) launch pattern
        kp = kernel_pass(dev, p, args.steps)
        s64 = spmv_pass(dev, p, capi.FP64)
        s32 = spmv_pass(dev, p, capi.FP32)
        dev.close()
        print(json.dumps({"profile_pass": True, "steps": args.steps,
                          "ms_per_launch": {k: round(v[1] / max(v[0], 1), 4) for k, v in kp.items()},
                          "launches": {k: v[0] for k, v in kp.items()},
                          "spmv_ms": {"fp64": round(s64[1] / max(s64[0], 1), 4),
                                      "fp32": round(s32[1] / max(s32[0], 1), 4)}}), flush=True)
        return

    # clock-settling pre-pass (untimed, reported): steps until settle_ms of wall time, then the warmup steps
    settle_steps = 0
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < args.settle_ms:
        assembly_step(dev)
        dev.synchronize()
        settle_steps += 1
    settle_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        assembly_step(dev)
    dev.synchronize()
    log(f"[bench] clock settle {settle_steps} steps / {settle_s * 1e3:.0f} ms, then {args.warmup} warmup steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assembly_step(dev)
    dev.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = p.ne * args.steps / dt / 1e6
    step_tflops = FLOP_PER_ELEM_KR * p.ne / (ms_per_step * 1e-3) / 1e12
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} Melem/s")
    for _ in range(args.warmup):
        separate_step(dev)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        separate_step(dev)
    dev.synchronize()
    sep_ms = (time.perf_counter() - t0) / args.steps * 1e3
    log(f"[bench] separate calls (ik_vector, ik_matrix): {sep_ms:.3f} ms/step, {p.ne / sep_ms / 1e3:.1f} Melem/s")

    kp = kernel_pass(dev, p, args.steps)
    ne_, k_ms = kp["element"][0], kp["element"][1] / max(kp["element"][0], 1)
    elems_per_launch = p.ne * args.steps / max(ne_, 1)
    achieved = FLOP_PER_ELEM_K * elems_per_launch / (k_ms * 1e-3) / 1e12
    ng, g_ms_tot = kp["gather"]
    g_ms = g_ms_tot / max(ng, 1)
    r_ms = kp["residual"][1] / max(kp["residual"][0], 1)
    nnzb = nnzb or 0
    staged, gathered, direct = dev.assembly_layout()
    # gather: staged blocks read, its upper blocks and their mirrors written, per-upper-block indices and
    # segment pointers (diagonal blocks have no mirror; counted as if they had, a bound within 7 %)
    gather_bytes = (72 * staged + 2 * 72 * gathered + 18 * gathered) / max(ng / args.steps, 1)
    gather_gbs = gather_bytes / (g_ms * 1e-3) / 1e9
    s64n, s64ms = spmv_pass(dev, p, capi.FP64)
    s32n, s32ms = spmv_pass(dev, p, capi.FP32)
    spmv64_ms, spmv32_ms = s64ms / max(s64n, 1), s32ms / max(s32n, 1)
    spmv64_bytes = 8 * 9 * nnzb + 4 * nnzb + 4 * (p.X.shape[0] + 1) + 16 * p.ndof
    spmv32_bytes = 4 * 9 * nnzb + 4 * nnzb + 4 * (p.X.shape[0] + 1) + 8 * p.ndof
    spmv64_gbs = spmv64_bytes / (spmv64_ms * 1e-3) / 1e9
    spmv32_gbs = spmv32_bytes / (spmv32_ms * 1e-3) / 1e9
    dev.close()
    log(f"[bench] element kernel {k_ms:.4f} ms/launch x {ne_ / args.steps:.0f} launches ({achieved:.1f} TFLOP/s "
        f"algorithmic), gather {g_ms:.4f} ms/launch ({gather_gbs:.0f} GB/s), residual {r_ms:.4f} ms, "
        f"SpMV fp64 {spmv64_ms:.4f} ms ({spmv64_gbs:.0f} GB/s), fp32 {spmv32_ms:.4f} ms ({spmv32_gbs:.0f} GB/s)")
    traffic = read_traffic(args.traffic)

    out = {
        "metric": "Newton solve time to ||R||<=1e-8 (mixed+learned vs FP64); assembly Melem/s",
        "value": round(value, 3),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"hex8 Neo-Hooke R+K assembly (one Newton state), {n}x{n}x{n} elements, unit "
                               "cube, 5% jitter seed 0, FULL Dirichlet, 3x3-block sliced ELL", "elements": p.ne,
                   "dofs": p.ndof, "block_nnz": nnzb, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                     "kernel": "hex8Element<NeoHooke,double,K>", "kernel_ms_per_launch": round(k_ms, 4),
                     "elements_per_launch": int(elems_per_launch), "flop_per_elem": FLOP_PER_ELEM_K,
                     "step_frac": round(step_tflops / PEAK_FP64_TFLOPS, 4), "step_tflops": round(step_tflops, 3),
                     "step_note": "whole R+K step (element, gather and residual passes) at 86,176 flop/element"},
        "clock_settle": {"steps": settle_steps, "ms": round(settle_s * 1e3, 1),
                         "note": "untimed assembly steps before the warmup steps, so the timed loop runs at settled "
                                 "clocks"},
        "assembly_breakdown_ms_per_step": {
            "element_K": round(kp["element"][1] / args.steps, 4), "gather_K": round(g_ms_tot / args.steps, 4),
            "residual": round(r_ms, 4)},
        "separate_calls": {"ms_per_step": round(sep_ms, 4), "Melem_per_s": round(p.ne / sep_ms / 1e3, 3),
                           "residual_kernel_ms": round(r_ms, 4)},
        "gather": {"ms_per_launch": round(g_ms, 4), "GBs": round(gather_gbs, 1),
                   "frac_hbm": round(gather_gbs / PEAK_HBM_GBS, 4), "bytes_per_launch": int(gather_bytes),
                   "staged_slots": staged, "gathered_upper_blocks": gathered, "direct_upper_blocks": direct},
        "pattern_build_s": round(pattern_s, 4),
        "spmv": {"fp64": {"ms": round(spmv64_ms, 4), "GBs": round(spmv64_gbs, 1),
                          "frac_hbm": round(spmv64_gbs / PEAK_HBM_GBS, 4), "bytes": spmv64_bytes},
                 "fp32": {"ms": round(spmv32_ms, 4), "GBs": round(spmv32_gbs, 1),
                          "frac_hbm": round(spmv32_gbs / PEAK_HBM_GBS, 4), "bytes": spmv32_bytes}},
    }
    if not args.no_newton:
        log(f"[bench] c2 load path ({args.c2_steps} steps, plain FP64) ...")
        out["c2_load_path"] = c2_load_path(n, args.c2_steps)
        out["c2_load_path"]["workload"] = (f"hex8 Neo-Hooke {n}^3 seed 0, {args.c2_steps} load steps, FP64 tangent, "
                                           "Jacobi-PCG to 1e-10 per Newton iteration, ||R|| <= 1e-8")
        log(f"[bench] c2: {out['c2_load_path']['seconds_per_step']} s per load step")
        if args.c2_steps > 0:
            r, _ = newton_path(pb.hex_block((12, 12, 12), seed=0), "fp64", args.c2_steps)
            out["c2_load_path"]["device_on_cpu_sample_12cubed"] = {k: r[k] for k in ("seconds", "newton_iterations",
                                                                                     "linear_iterations")}
        # config c1 on the device: the 2D plane-strain cantilever (run as its one-layer hex8 extrusion)
        r, _ = newton_path(pb.cantilever2d(seed=0), "fp64", 10)
        r["seconds_per_step"] = round(r["seconds"] / 10, 4)
        r["workload"] = ("2D plane-strain NH cantilever 64x16 Q1 over 4x1, 10 load steps to ||R|| <= 1e-8, FP64 "
                         "tangent, Jacobi-PCG to 1e-10 per Newton iteration (the reference's c1 uses a sparse direct "
                         "solve; see cpu_baseline.c1)")
        out["c1_load_path"] = r
        log(f"[bench] c1 on the device: {r['seconds']} s, {r['newton_iterations']} Newton")
        mlp = np.fromfile(args.mlp, dtype=np.float32)
        out["newton"] = evaluate_newton(n, args.newton_steps, mlp, EVAL_SEEDS[:max(1, args.eval_seeds)])
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


def config_precision(args):
    """128x128x64 hex8 (3.24 M DOFs), Neo-Hooke and SVK: the FP32-tangent / FP32-PCG + FP64-refinement form
    (block-Jacobi and two-level preconditioners) against the FP64-tangent form on the same seeded mesh and load,
    5 load steps to ||R|| <= 1e-8."""
    warm_libraries()
    res = {}
    for name, mat in (("neohooke", capi.NEOHOOKE), ("svk", capi.SVK)):
        p = pb.hex_block((128, 128, 64), material=mat, seed=0)
        log(f"[bench] c3 {name}: {p.ne} elements, {p.ndof} dofs")
        out, us = {}, {}
        for mode in ("fp64", "mixed_exact", "mixed_secant", "tl_secant"):
            out[mode], us[mode] = newton_path(p, mode, args.c3_steps)
            out[mode]["seconds_per_load_step"] = round(out[mode]["seconds"] / args.c3_steps, 4)
            log(f"[bench] c3 {name} {mode}: {out[mode]}")
        ref = us["fp64"]
        res[name] = {"modes": out,
                     "rel_err_u_vs_fp64": {m: float(np.linalg.norm(us[m] - ref) / np.linalg.norm(ref)) for m in us},
                     "max_abs_err_u_vs_fp64": {m: float(np.max(np.abs(us[m] - ref))) for m in us},
                     "speedup_vs_fp64": {m: round(out["fp64"]["seconds"] / out[m]["seconds"], 3) for m in out}}
    return {"config": "hex8 128x128x64 precision variant (NH and SVK), 5% jitter seed 0, top traction, "
                      f"{args.c3_steps} load steps", "dofs": int(p.ndof), "results": res}


def config_eas(args):
    """96x96x48 hex8 EAS-21 Neo-Hooke, x=0 clamped, bending traction on x=2: FP64 R+K assembly throughput
    (element matrices are checked against the oracle by the GPU tests) and a plain-FP64 load path."""
    p = pb.bending_block((96, 96, 48), seed=0)
    log(f"[bench] c4: {p.ne} elements, {p.ndof} dofs")
    dev = ia.Device(p)
    u = 2e-4 * np.random.default_rng(0).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 1.0)
    for _ in range(2):
        assembly_step(dev)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.c4_steps):
        assembly_step(dev)
    dev.synchronize()
    dt = time.perf_counter() - t0
    dev.close()
    out = {"config": "hex8 EAS-21 Neo-Hooke 96x96x48 over [0,2]x[0,2]x[0,1], 5% jitter seed 0, x=0 clamped, "
                     f"traction {[round(t, 3) for t in p.traction]} on x=2", "elements": p.ne, "dofs": p.ndof,
           "assembly": {"ms_per_step": round(dt / args.c4_steps * 1e3, 3),
                        "Melem_per_s": round(p.ne * args.c4_steps / dt / 1e6, 3)}}
    if args.c4_load_steps > 0:
        r, u64 = newton_path(p, "fp64", args.c4_load_steps)
        r["seconds_per_load_step"] = round(r["seconds"] / args.c4_load_steps, 4)
        out["newton_fp64"] = r
        # the mixed form on the EAS element: FP32 condensed tangent (gathered from the FP64 element pass),
        # FP32 block-Jacobi PCG with FP64 refinement, inexact Newton, secant predictor
        log("[bench] c4 newton mixed_secant ...")
        r, u32 = newton_path(p, "mixed_secant", args.c4_load_steps)
        r["seconds_per_load_step"] = round(r["seconds"] / args.c4_load_steps, 4)
        r["speedup_vs_fp64"] = round(out["newton_fp64"]["seconds"] / r["seconds"], 3)
        r["rel_diff_u_vs_fp64"] = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
        out["newton_mixed_secant"] = r
        # the same with the two-level preconditioner
        log("[bench] c4 newton tl_secant ...")
        warm_libraries()
        r, u2l = newton_path(p, "tl_secant", args.c4_load_steps)
        r["seconds_per_load_step"] = round(r["seconds"] / args.c4_load_steps, 4)
        r["speedup_vs_fp64"] = round(out["newton_fp64"]["seconds"] / r["seconds"], 3)
        r["rel_diff_u_vs_fp64"] = float(np.linalg.norm(u2l - u64) / np.linalg.norm(u64))
        out["newton_tl_secant"] = r
    return out


def np_i64():
    return np.zeros(1, dtype=np.int64)


if __name__ == "__main__":
    main()
