"""Benchmark of the hex8 Neo-Hooke hot path on one MI355X.

Workload (default): 3D hex8 Neo-Hooke block, 64x64x64 elements (823,875 DOFs), unit cube, E=1000, nu=0.3,
seeded +-5% interior jitter (seed 0), bottom face clamped, seeded top-face traction, FP64.

One step = the residual R and tangent K of one Newton state (the pair NewtonRaphson evaluates per state),
FP64, FULL Dirichlet treatment, summed into the 3x3-block CSR pattern, inputs resident in HBM, through
ik_assemble (the tangent's element kernel and gather, then the residual pass; one error check and one
host synchronisation for the pair).
value = elements assembled per second (Melem/s) over the timed steps. The same state assembled by the two
separate calls (ik_vector, then ik_matrix) is timed beside it.

Also reported (same run): the element kernel's roofline figures (HIP events on the library stream; HBM
traffic per launch from the committed rocprofv3 counter summary), the gather and SpMV bandwidths, the CPU
baseline (oracle restatement timed on a bounded sample on the host cores) and the Newton load path
(40 load steps to ||R|| <= 1e-8) in the plain FP64 form and the mixed / learned forms.
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


def cpu_baseline(seconds):
    """Oracle (reference-cost port: tangent per node pair, serial scatter) K+R assembly on a 16^3 sample,
    repeated until `seconds` of CPU time, single thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    p = pb.hex_block((16, 16, 16), seed=0)
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, 0, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    o.set_faithful(True)
    u = 0.005 * np.random.default_rng(0).standard_normal(p.ndof) * (1 - p.fixed)
    o.pattern()
    n = 0
    t0 = time.perf_counter()
    while True:
        o.matrix(u, 1.0)
        o.vector(u, 1.0)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return dict(value=round(n * p.ne / dt / 1e6, 6), unit="Melem/s", cores=1, kind="port",
                sample=f"{n}x K+R assembly of a 16^3 hex8 NH mesh (4096 elements), reference cost structure "
                       f"(tangent per node pair), serial, {dt:.1f} s")


NEWTON_MODES = {
    # plain FP64 form: FP64 tangent, Jacobi-PCG to 1e-10 per Newton iteration, constant predictor (LoadControl)
    "fp64": dict(precision=capi.FP64, kind=capi.PCG_JACOBI_FP64, rtol=1e-10, forcing=0.0, pred=capi.PRED_CONSTANT),
    # same precision, inexact Newton + secant predictor (isolates the algorithmic part of the gain)
    "fp64_inexact_secant": dict(precision=capi.FP64, kind=capi.PCG_JACOBI_FP64, rtol=1e-10, forcing=0.9,
                                pred=capi.PRED_LINEAR),
    # mixed form: FP32 tangent + FP32 block-Jacobi PCG with FP64 refinement, inexact Newton, secant predictor
    "mixed_secant": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.9,
                         pred=capi.PRED_LINEAR),
    # mixed form + learned predictor: Newton starts near the solution, where a loose first linear solve costs a
    # Newton iteration, so the forcing term is capped at 1e-3 (best of 0.1 / 0.01 / 1e-3 on seeds 2000-2002)
    "mixed_mlp": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.9,
                      pred=capi.PRED_MLP, eta_max=1e-3),
    # precision alone: the plain form's Newton and predictor with the FP32 tangent and refined FP32 PCG
    "mixed_exact": dict(precision=capi.FP32, kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, forcing=0.0,
                        pred=capi.PRED_CONSTANT),
}


def newton_path(p, mode, steps, mlp_params=None):
    m = NEWTON_MODES[mode]
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
                final_residual=float(res[info.solves - 1])), u


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
    """Average SpMV launch inside the PCG of the given precision (FULL K assembled first)."""
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
    ap.add_argument("--warmup", type=int, default=30, help="untimed steps (the clocks settle over ~20 ms)")
    ap.add_argument("--n", type=int, default=64, help="elements per direction")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--newton-steps", type=int, default=40, help="load steps of the Newton comparison")
    ap.add_argument("--newton-seed", type=int, default=2000, help="unseen seed of the Newton comparison")
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

    n = args.n
    p = pb.hex_block((n, n, n), seed=0)
    log(f"[bench] problem {p.name}: {p.ne} elements, {p.ndof} dofs")
    dev = ia.Device(p)
    u = 0.005 * np.random.default_rng(0).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 1.0)
    nbr, nnzb = np_i64(), np_i64()
    t0 = time.perf_counter()  # the first matrix query builds the pattern (createOccupationPattern, once)
    dev.check(capi.lib().ik_matrix_bsr(dev.h, capi.ptr(nbr), capi.ptr(nnzb), None, None, None, None))
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

    for _ in range(args.warmup):
        assembly_step(dev)
    dev.synchronize()
    log("[bench] warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assembly_step(dev)
    dev.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = p.ne * args.steps / dt / 1e6
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
                               "cube, 5% jitter seed 0, FULL Dirichlet, 3x3-block CSR", "elements": p.ne,
                   "dofs": p.ndof, "block_nnz": nnzb, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                     "kernel": "hex8Element<NeoHooke,double,K>", "kernel_ms_per_launch": round(k_ms, 4),
                     "elements_per_launch": int(elems_per_launch), "flop_per_elem": FLOP_PER_ELEM_K},
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
        pn = pb.hex_block((n, n, n), seed=args.newton_seed)
        mlp = np.fromfile(args.mlp, dtype=np.float32) if os.path.exists(args.mlp) else None
        res, us = {}, {}
        for mode in ("fp64", "fp64_inexact_secant", "mixed_secant", "mixed_mlp"):
            if mode == "mixed_mlp" and mlp is None:
                continue
            log(f"[bench] newton {mode} ...")
            res[mode], us[mode] = newton_path(pn, mode, args.newton_steps, mlp)
            log(f"[bench] newton {mode}: {res[mode]['seconds']} s, {res[mode]['newton_iterations']} Newton, "
                f"{res[mode]['linear_iterations']} PCG iterations")
        best = "mixed_mlp" if "mixed_mlp" in res else "mixed_secant"
        ref = us["fp64"]
        out["newton"] = {
            "workload": f"hex8 Neo-Hooke {n}^3, {args.newton_steps} load steps to ||R||<=1e-8, unseen seed "
                        f"{args.newton_seed}, traction {[round(t, 3) for t in pn.traction]}; each time includes "
                        "the one-time pattern build of the first matrix call (pattern_build_s)",
            "modes": res,
            "speedup_vs_fp64": {k: round(res["fp64"]["seconds"] / v["seconds"], 3) for k, v in res.items()},
            "rel_diff_u_vs_fp64": {k: float(np.linalg.norm(us[k] - ref) / np.linalg.norm(ref)) for k in res},
            "headline": f"{best} vs fp64",
        }
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


def config_precision(args):
    """128x128x64 hex8 (3.24 M DOFs), Neo-Hooke and SVK: the FP32-tangent / FP32-PCG + FP64-refinement form
    against the FP64-tangent form on the same seeded mesh and load, 5 load steps to ||R|| <= 1e-8."""
    res = {}
    for name, mat in (("neohooke", capi.NEOHOOKE), ("svk", capi.SVK)):
        p = pb.hex_block((128, 128, 64), material=mat, seed=0)
        log(f"[bench] c3 {name}: {p.ne} elements, {p.ndof} dofs")
        out, us = {}, {}
        for mode in ("fp64", "mixed_exact", "mixed_secant"):
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
    return out


def np_i64():
    return np.zeros(1, dtype=np.int64)


if __name__ == "__main__":
    main()
