#!/usr/bin/env python3
"""Benchmark of the multi-domain TDBEM GMRES path on one MI355X.

Workload (default): the layered 2x2x4 block of 4 stacked sub-domains with
seeded materials (E in [1,10], v in [0.2,0.35], rho in [1,3]), constant square
elements with h = 0.1 (6400 elements), Ricker normal traction on the top face,
bottom fixed, dt = 0.5 h / max c1, 400 time steps, plain FP64.

One bench step is one whole transient solve: case setup (geometry, singular
self integrals), step-0 assembly and preconditioner, history assembly
(causality cull + quadrature of every MTS/MGS block) and the 400-step time
loop (history convolution + GMRES per step).  The reported value is
time steps per second of that whole job: K * NStep / wall time.

Inputs are generated on the host and uploaded by the case setup inside the
timed region; the GPU holds everything else.  Prints one JSON line.  Beside
the headline FP64 value the line carries a "batched" block (64 seeded load
cases on the same block solved as one batch) and a "mixed_precision" block
(the 8-domain 8192-element 600-step box in FP64 and in both mixed modes, with
the error of the mixed runs against the FP64 run) and an "adjoint" block
(adjoint gradients/s of a receiver misfit on the 4-domain 3456-element block).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "parallel3.0_omp_multidomin_1node_amd"))
sys.path.insert(0, str(REPO))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per far-convolution launch (k_conv_far_q<16> with 2, 4 or 8
# wavefronts per source row) on the default workload (layered block, h = 0.1,
# 400 steps, FP64), from FETCH_SIZE and WRITE_SIZE counter passes calibrated
# with a 1 GiB device copy (fetch x2.000, write x1.000); the summary with its
# commands is kept in profiles/k_conv_far16_traffic_layered2x2x4_h0.1_fp64.md.
MEASURED_TRAFFIC = {("layered_2x2x4_h0.1", 400, 16, 2): 61703231633.9,
                    ("layered_2x2x4_h0.1", 400, 16, 4): 56938840066.6,
                    ("layered_2x2x4_h0.1", 400, 16, 8): 52872278677.8}
FP64_PEAK_TFLOPS = 78.6
# FP64 lane-operations of one history quadrature (k_history_pw + k_history_reg)
# of the default workload: 64 x (ADD + MUL + TRANS + 2 FMA) F64 VALU wave
# instructions from the SQ_INSTS_VALU_*_F64 counters (every lane of an issued
# instruction counted, so an upper bound of the useful FLOPs); the summary with
# its command is kept in profiles/k_history_fp64_valu_layered2x2x4_h0.1.md.
MEASURED_QUAD_FLOPS = {("layered_2x2x4_h0.1", 400): 1.8535e13}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(tdbem, w):
    m = w.mesh
    return tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                      itf_rows=m.interfaces)


def run_once(tdbem, w, restart=200, tol=1e-5, keep=False, tile=16, gmres_device=1, tune_cache=None, far_parts=0):
    t0 = time.perf_counter()
    c = make_case(tdbem, w)
    c.set_option("conv_tile", tile)
    c.set_option("gmres_device", gmres_device)
    if far_parts:
        c.set_option("far_parts", far_parts)
    c.assemble_step0(10)
    entries = c.assemble_history(w.nstep)
    tuned = c.autotune(tune_cache) if tune_cache else None
    U, T, iters, flags, solved, times = c.solve(w.known, w.nstep, restart=restart, tol=tol)
    tdbem.synchronize()
    wall = time.perf_counter() - t0
    info = dict(wall=wall, entries=entries, iters=iters, flags=flags, solved=solved, timings=c.timings(),
                U=U if keep else None, tuned=tuned)
    if keep:
        info["case"] = c
    else:
        c.close()
    return info


def conv_far_bytes_per_launch(counts, npairs, ne, nhist, nstep, tile):
    """Algorithmic HBM bytes of one far-part convolution launch (tile of up to
    `tile` steps starting at K0 = 2, 2+tile, ...), averaged over launches: each
    band entry with lag j <= K0+Tn-1 read once (9 MTS + 6 MGS doubles), per-pair
    band metadata (kb, len, offset: 16 B), the solved state time series read
    once (two histories, 3 doubles per column and solved step) and Tn
    right-hand-side slices written."""
    cum = np.cumsum(counts)
    total, launches = 0.0, 0
    for K0 in range(2, nstep + 1, tile):
        Tn = min(tile, nstep - K0 + 1)
        j = min(K0 + Tn - 1, len(counts))
        total += cum[j - 1] * 120.0 + npairs * 16.0 + 2.0 * nhist * (K0 - 1) * 24.0 + Tn * ne * 24.0
        launches += 1
    return total / launches


def cpu_baseline(threads):
    """Oracle (CPU FP64 restatement of the reference, OpenMP over the
    reference's independent loops) on a bounded sample of the same workload
    family: the layered block at h = 0.25 (1024 elements), 100 steps."""
    from oracle import oracle as O
    from tdbem import io, mesh
    os.environ["OMP_NUM_THREADS"] = str(threads)
    w = mesh.layered_config2(h=0.25, nstep=100)
    m = w.mesh
    card = io.DataCard(title="x", thread_num=threads, flag_dyna=2, beta=0.5, nstep=w.nstep, gap=w.gap, E=1.0, v=0.0,
                       rho=1.0, amplitude=1, solver=2, domain_count=4,
                       materials=[(int(r[0]), r[1], r[2], r[3]) for r in w.mats.tolist()],
                       interfaces=[tuple(int(x) for x in r) for r in m.interfaces.tolist()])
    mats, itf = io.normalise_model(card, len(m.ele))
    t0 = time.perf_counter()
    c = O.Case(m.points, m.ele, m.bcid, mats, itf, w.dt, w.gap)
    U, T, it, fl, solved = c.solve(w.known, w.nstep, tol=1e-5)
    dt = time.perf_counter() - t0
    return dict(value=w.nstep / dt, wall=dt, w=w, U=U)


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime the library
    already loaded, so a profiler's FETCH_SIZE / WRITE_SIZE of this dispatch
    calibrates the counters against a known byte count."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError("hipMemcpy failed")
    log(f"calibration copy of {nbytes} bytes done")


def mixed_precision_report(tdbem, h=0.125, nstep=600, gmres_device=1):
    """Mixed precision on the 8-domain stacked box (8192 elements, 600 steps):
    whole-job steps/s of the FP64 run and of both mixed modes, each solved to
    GMRES tol 1e-12, and the max relative displacement / traction error of the
    mixed runs against the FP64 run."""
    from tdbem import mesh
    w = mesh.mixed_config3(h=h, nstep=nstep)
    m = w.mesh
    res = {}
    for mixed in (0, 1, 2):
        t0 = time.perf_counter()
        c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                       itf_rows=m.interfaces)
        c.set_option("mixed", mixed)
        c.set_option("gmres_device", gmres_device)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        U, T, it, fl, solved, _ = c.solve(w.known, w.nstep, tol=1e-12)
        tdbem.synchronize()
        wall = time.perf_counter() - t0
        tm = c.timings()
        res[mixed] = dict(U=U, T=T, wall=wall, solved=solved, loop=tm["convolution"] + tm["gmres"])
        c.close()
        log(f"config 3 mixed={mixed}: {wall:.2f} s")
    out = {"workload": w.name, "elements": int(len(m.ele)), "domains": int(m.domain_count), "nstep": w.nstep,
           "gmres_tol": 1e-12}
    U0, T0 = res[0]["U"], res[0]["T"]
    names = {0: "fp64", 1: "fp32_blocks", 2: "fp32_bf16_blocks"}
    for k, r in res.items():
        e = {"steps_per_s": round(w.nstep / r["wall"], 3), "time_loop_steps_per_s": round(w.nstep / r["loop"], 3)}
        if k:
            e["max_rel_err_u"] = float(np.abs(r["U"] - U0).max() / np.abs(U0).max())
            e["max_rel_err_t"] = float(np.abs(r["T"] - T0).max() / np.abs(T0).max())
        out[names[k]] = e
    return out


def batched_mixed_report(tdbem, single_fp64_steps_per_s, h=0.125, nstep=600, nrhs=64, gmres_device=1):
    """The mixed-precision plus batched form on the 8-domain stacked box
    (8192 elements, 600 steps): 64 seeded load cases as one batch over FP32 +
    bf16 bands (mode 2), whole job timed, against the single FP64 whole job of
    the same box (mixed_precision report); displacement error of the batch's
    first case against its own FP64 solve, both at GMRES tol 1e-12."""
    from tdbem import mesh
    w = mesh.mixed_config3(h=h, nstep=nstep)
    m = w.mesh
    cases = mesh.load_cases(w, nrhs)

    def case(mixed):
        c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                       itf_rows=m.interfaces)
        c.set_option("mixed", mixed)
        c.set_option("gmres_device", gmres_device)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        return c

    t0 = time.perf_counter()
    c = case(2)
    BU, BT, it, fl, solved, times = c.solve_batched(cases, w.nstep, tol=1e-5)
    tdbem.synchronize()
    t1 = time.perf_counter()
    c.close()
    v = nrhs * w.nstep / (t1 - t0)
    # accuracy of the mixed batch: case 0 again at tol 1e-12, mixed batch of one vs FP64
    c = case(2)
    U2 = c.solve_batched(cases[:1], w.nstep, tol=1e-12)[0][0]
    c.close()
    c = case(0)
    U0 = c.solve(cases[0], w.nstep, tol=1e-12)[0]
    c.close()
    err = float(np.abs(U2 - U0).max() / np.abs(U0).max())
    log(f"batched mixed {nrhs}: {t1 - t0:.2f} s, error {err:.2e}")
    return {"workload": w.name, "load_cases": nrhs, "mixed": "FP32 + bf16 bands (mode 2)",
            "case_steps_per_s": round(v, 2), "vs_single_fp64_whole_job": round(v / single_fp64_steps_per_s, 2),
            "target_vs_single_fp64_whole_job": 10.0, "max_rel_err_u_vs_fp64": err,
            "all_solved": bool((solved == w.nstep).all()),
            "breakdown_s": {"convolution": round(float(times[0]), 3), "gmres": round(float(times[1]), 3)}}


def batched_report(tdbem, w, single_steps_per_s, nrhs=64):
    """64 seeded load cases on the layered block solved as one batch (shared
    bands, lane-per-case convolution, lockstep GMRES), whole job timed."""
    from tdbem import mesh
    cases = mesh.load_cases(w, nrhs)
    m = w.mesh
    t0 = time.perf_counter()
    c = make_case(tdbem, w)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    t1 = time.perf_counter()
    BU, BT, it, fl, solved, times = c.solve_batched(cases, w.nstep, tol=1e-5)
    tdbem.synchronize()
    t2 = time.perf_counter()
    c.close()
    v = nrhs * w.nstep / (t2 - t0)
    log(f"batched {nrhs}: {t2 - t0:.2f} s")
    return {"load_cases": nrhs, "case_steps_per_s": round(v, 2),
            "time_loop_case_steps_per_s": round(nrhs * w.nstep / (t2 - t1), 2),
            "vs_single_case_whole_job": round(v / single_steps_per_s, 2),
            "all_solved": bool((solved == w.nstep).all()), "gmres_inner_iterations_mean": float(it.mean()),
            "breakdown_s": {"setup_step0_history": round(t1 - t0, 3), "convolution": round(float(times[0]), 3),
                            "gmres": round(float(times[1]), 3)}}


def models_report(tdbem):
    """16 independent single-domain cube models of 500-6000 elements (seeded
    sizes, materials and loads, dt = 0.5 h / c1 each, 200 steps): every
    model's whole job one after another, then all of them through
    tdbem_run_models (one stream and host thread per model, history bands
    held at once within 80 % of the free device memory)."""
    from tdbem import mesh
    ws = mesh.models_config4ii()

    def case(w):
        m = w.mesh
        return tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=1, mat_rows=w.mats,
                          itf_rows=m.interfaces)

    steps = sum(w.nstep for w in ws)
    t0 = time.perf_counter()
    for w in ws:
        c = case(w)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        c.solve(w.known, w.nstep)
        c.close()
    tdbem.synchronize()
    t1 = time.perf_counter()
    cs = [case(w) for w in ws]
    res = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws])
    tdbem.synchronize()
    t2 = time.perf_counter()
    for c in cs:
        c.close()
    log(f"models: one after another {t1 - t0:.2f} s, together {t2 - t1:.2f} s")
    return {"models": len(ws), "elements": [int(len(w.mesh.ele)) for w in ws], "nstep": ws[0].nstep,
            "model_steps_per_s_sequential": round(steps / (t1 - t0), 2),
            "model_steps_per_s_together": round(steps / (t2 - t1), 2),
            "speedup": round((t1 - t0) / (t2 - t1), 3), "all_solved": all(r[4] == w.nstep for r, w in zip(res, ws))}


def adjoint_report(tdbem, gmres_device=1):
    """Reverse-time adjoint on the 1x1x4 block (3456 elements, 4 domains, 300
    steps): misfit over 32 receiver elements against observations from
    materials perturbed by +-10 %, gradient with respect to each domain's
    (E, v, rho).  Timed: one gradient (forward solve, backward sweep, analytic
    history-block derivatives, step-0 differences and contraction) after the
    case's history assembly."""
    from tdbem import mesh
    a = mesh.adjoint_config5()
    w = a.workload
    m = w.mesh

    def case(mats):
        c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=mats, itf_rows=m.interfaces)
        c.set_option("gmres_device", gmres_device)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        return c

    c = case(a.obs_mats)
    U, *_ = c.solve(w.known, w.nstep, tol=1e-12)
    c.close()
    obs = U[:, a.receivers, :]
    t0 = time.perf_counter()
    c = case(w.mats)
    t1 = time.perf_counter()
    J, g, _ = c.adjoint(w.known, w.nstep, a.receivers, obs, tol=1e-12, fd_step=1e-6)
    tdbem.synchronize()
    t2 = time.perf_counter()
    tm = c.timings()
    c.close()
    log(f"adjoint: assembly {t1 - t0:.2f} s, gradient {t2 - t1:.2f} s")
    return {"workload": w.name, "elements": int(len(m.ele)), "domains": 4, "nstep": w.nstep,
            "receivers": int(len(a.receivers)), "gradients_per_s": round(1.0 / (t2 - t1), 4),
            "gradient_s": round(t2 - t1, 3), "assembly_s": round(t1 - t0, 3), "misfit": J,
            "gradient_E_v_rho": np.asarray(g).round(12).tolist(),
            "history_derivatives": "analytic", "step0_difference_step": 1e-6,
            "breakdown_s": {k[8:]: round(float(v), 3) for k, v in tm.items() if k.startswith("adjoint_")}}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed whole-job solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed whole-job solves")
    ap.add_argument("--h", type=float, default=0.1)
    ap.add_argument("--nstep", type=int, default=400)
    ap.add_argument("--conv-tile", type=int, default=16, choices=(8, 16), help="steps per far-part tile")
    ap.add_argument("--gmres-device", type=int, default=1, choices=(0, 1, 2),
                    help="GMRES driver: 0 host loop, 1 device state + hipGraphs (default), 2 device state + direct "
                         "launches (for profiler runs)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--far-parts", type=int, default=0, choices=(0, 1, 2, 4, 8),
                    help="far-convolution wavefronts per row (fixes the shape; 0: autotuned or the default 4)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the config-3 mixed-precision report")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device copy of BYTES (profiler counter calibration)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-load-case batched report")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient report")
    ap.add_argument("--models", action="store_true",
                    help="also run the 16-model report (about 1.3 TB of history bands in total; a few minutes)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only; WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    import tdbem
    from tdbem import mesh

    w = mesh.layered_config2(h=args.h, nstep=args.nstep)
    if args.copy_calibration:
        tdbem.synchronize()
        copy_calibration(args.copy_calibration)
    log(f"workload {w.name}: {len(w.mesh.ele)} elements, {w.nstep} steps, dt={w.dt:.6g}")
    # launch shapes: tuned in the first warm-up run (tdbem_autotune, timed on the device), stored in a cache
    # file of this process, applied from the cache in the timed runs
    tune_cache = None
    if args.warmup >= 1 and not args.no_autotune and not args.far_parts:
        import tempfile
        tune_cache = os.path.join(tempfile.gettempdir(), f"tdbem_tune_{os.getpid()}.txt")
        if os.path.exists(tune_cache):
            os.remove(tune_cache)
    for i in range(args.warmup):
        r = run_once(tdbem, w, tile=args.conv_tile, gmres_device=args.gmres_device, tune_cache=tune_cache,
                     far_parts=args.far_parts)
        log(f"warmup {i}: {r['wall']:.2f} s, entries {r['entries']}, launch shapes {r['tuned']}")
    tdbem.synchronize()
    t0 = time.perf_counter()
    runs = []
    for i in range(args.steps):
        runs.append(run_once(tdbem, w, keep=(i == args.steps - 1), tile=args.conv_tile,
                             gmres_device=args.gmres_device, tune_cache=tune_cache, far_parts=args.far_parts))
        log(f"step {i}: {runs[-1]['wall']:.2f} s")
    tdbem.synchronize()
    elapsed = time.perf_counter() - t0
    last = runs[-1]
    value = args.steps * w.nstep / elapsed
    tm = last["timings"]
    c = last["case"]
    counts, npairs = c.history_stats(w.nstep)
    dof = c.dofmap()["counts"]
    nhist = int(dof[7] + dof[8]) // 2
    conv_avg = tm["conv_far"] / max(tm["conv_far_launches"], 1)
    bytes_launch = conv_far_bytes_per_launch(counts, npairs, c.ne, nhist, w.nstep, args.conv_tile)
    conv_gbs = bytes_launch / conv_avg / 1e9
    loop_time = tm["convolution"] + tm["gmres"]
    far_parts = last["tuned"][0] if last["tuned"] else (args.far_parts or 4)
    far_kernel = f"k_conv_far_q, {far_parts} wavefronts per row" if far_parts else "k_conv_far_p"
    out = {
        "metric": "time steps/s (whole job: setup + step-0 + history assembly + time loop), FP64",
        "value": round(value, 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded layered block, Ricker load)",
        "config": {"workload": w.name, "elements": int(len(w.mesh.ele)), "domains": 4, "nstep": w.nstep,
                   "dt": w.dt, "gmres_restart": 200, "gmres_tol": 1e-5, "history_band_entries": int(last["entries"]),
                   "conv_tile": args.conv_tile,
                   "gmres_driver": ("host loop", "device state, hipGraph replay", "device state, direct launches")[
                       args.gmres_device],
                   "launch_shapes": ({"far_wavefronts_per_row": last["tuned"][0], "gmres_reduction_blocks":
                                      last["tuned"][1], "picked_by": "autotuner (warm-up run)"}
                                     if last["tuned"] else "defaults")},
        "breakdown_s": {k: round(float(v), 4) for k, v in tm.items()
                        if k != "conv_far_launches" and not k.startswith("adjoint_")},
        "time_loop_steps_per_s": round(w.nstep / max(loop_time, 1e-9), 3),
        # every band entry holds one MTS and one MGS block (computed; the
        # 1e-14 drop rule may leave some empty)
        "assembled_blocks_per_s": round(2.0 * float(last["entries"]) / max(tm["history"], 1e-9), 1),
        "gmres_inner_iterations_mean": float(np.mean(last["iters"])),
        "roofline": {"kernel": f"{far_kernel} (time-blocked history convolution, {args.conv_tile} steps per launch)",
                     "bound": "hbm", "achieved": round(conv_gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(conv_gbs / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((w.name, w.nstep, args.conv_tile, far_parts)),
                     "avg_launch_ms": round(1000 * conv_avg, 4),
                     "bytes_per_launch": round(bytes_launch)},
    }
    qflops = MEASURED_QUAD_FLOPS.get((w.name, w.nstep))
    if qflops and tm.get("quadrature", 0.0) > 0.0:
        qrate = qflops / tm["quadrature"] / 1e12
        out["quadrature_roofline"] = {"kernels": "k_history_pw + k_history_reg (history quadrature)",
                                      "bound": "fp64-valu", "achieved": round(qrate, 2), "peak": FP64_PEAK_TFLOPS,
                                      "unit": "TFLOP/s", "frac": round(qrate / FP64_PEAK_TFLOPS, 4),
                                      "flops_per_assembly": qflops, "kernel_s": round(tm["quadrature"], 4)}
    c.close()
    if not args.no_batched:
        out["batched"] = batched_report(tdbem, w, value)
    if not args.no_mixed:
        out["mixed_precision"] = mixed_precision_report(tdbem, gmres_device=args.gmres_device)
        if not args.no_batched:
            out["batched_mixed"] = batched_mixed_report(tdbem, out["mixed_precision"]["fp64"]["steps_per_s"],
                                                        gmres_device=args.gmres_device)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_report(tdbem, gmres_device=args.gmres_device)
    if args.models:
        out["models"] = models_report(tdbem)
    if not args.no_cpu_baseline:
        log(f"cpu baseline on {args.cpu_threads} threads ...")
        cb = cpu_baseline(args.cpu_threads)
        # the GPU path on the same bounded sample, for a like-for-like ratio
        g = run_once(tdbem, cb["w"], keep=True, tile=args.conv_tile)
        err = float(np.abs(g["U"] - cb["U"]).max() / max(np.abs(cb["U"]).max(), 1e-300))
        g["case"].close()
        out["cpu_baseline"] = {"value": round(cb["value"], 4), "unit": "time steps/s", "cores": args.cpu_threads,
                               "kind": "port",
                               "sample": "oracle (C FP64 restatement, OpenMP) whole job on the layered block at "
                                         "h=0.25 (1024 elements), 100 steps, same seeds and load family",
                               "wall_s": round(cb["wall"], 3),
                               "gpu_same_sample_steps_per_s": round(cb["w"].nstep / g["wall"], 3),
                               "gpu_vs_cpu_max_rel_diff_u": err}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
