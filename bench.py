#!/usr/bin/env python3
"""Benchmark of the multi-domain TDBEM GMRES path on one MI355X.

Headline workload: the stacked 2x2x8 box of 8 sub-domains (h = 0.125, 8192
constant square elements, seeded E in [1,20], v in [0.15,0.4], rho in [1,4],
dt = 0.5 h / max c1, 600 time steps), with 64 seeded load cases (Ricker or
Heaviside pulses of random direction and amplitude on random top-face
patches), in the mixed-precision plus batched form: history blocks computed
in FP64 and stored as FP32 plus a bf16 residual, the 64 cases solved as one
batch (shared bands, lockstep GMRES under FP64 arithmetic), GMRES restart 200
and tolerance 1e-10.

One bench step is one whole job: case setup (geometry, singular self
integrals), step-0 assembly and preconditioner, causality cull and history
quadrature of every MTS / MGS block, and the 600-step time loop of the 64
cases (history convolution + GMRES per step).  The load histories are in
device memory before the timed region starts and the results stay there
(tdbem_solve_batched_device).  value = 64 * 600 * steps / elapsed, time
steps per second summed over the load cases.

Beside it the line carries the plain FP64 form of the north-star comparison
(one load case, FP64 bands, same GMRES tolerance, whole job, same device-
resident I/O), the ratio of the two, and the displacement / traction error
of cases of the timed batch against their FP64 single solves; the roofline of
the dominant kernel (the batched far-part convolution, FP64 FMA bound) with
its HIP-event time; the oracle on the host's cores as cpu_baseline; and
secondary blocks (the layered 6400-element block in FP64 with the HBM
roofline of its convolution, the adjoint gradient rate on the 3456-element
block) when the run's time allows.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "parallel3.0_omp_multidomin_1node_amd"))
sys.path.insert(0, str(REPO))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BATCH_TILE = 8  # steps per far-part launch of the batched convolution
# HBM bytes per launch from FETCH_SIZE + WRITE_SIZE counter passes calibrated
# with a 1 GiB device copy (fetch x2.000, write x1.000); summaries with their
# commands under profiles/.
MEASURED_TRAFFIC = {("layered_2x2x4_h0.1", 400, 16, 2): 61703231633.9,
                    ("layered_2x2x4_h0.1", 400, 16, 4): 56938840066.6,
                    ("layered_2x2x4_h0.1", 400, 16, 8): 52872278677.8}
# the spectral products kernel of the 64-case mixed + batched box (same calibration)
HEADLINE_TRAFFIC = {("stacked_2x2x8_h0.125", 600, 64, 2, 16): 452377070181.1}
# timings entries that are counts, flags, flops or bytes rather than seconds
NOT_SECONDS = {"conv_far_launches", "conv_spectral_launches", "conv_spectral_flops", "conv_spectral_bytes",
               "spectral_levels", "spectral_level_stats",
               "spectral", "stream_slabs", "quad_evals_t2", "quad_evals_t1u", "quad_evals_regular", "quad_records",
               "quad_pw_entries", "quad_pw_waves"}
SECONDARY_BUDGET_S = 420.0  # secondary blocks run only while the bench has used less than this


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def physical_cores():
    """Physical cores of the host (lscpu: distinct core, socket pairs)."""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        cores = {ln for ln in out.splitlines() if ln and not ln.startswith("#")}
        return len(cores) or None
    except Exception:
        return None


def make_case(tdbem, w, **opts):
    m = w.mesh
    c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                   itf_rows=m.interfaces)
    for k, v in opts.items():
        c.set_option(k, v)
    return c


def batched_far_flops_per_launch(counts, nstep, maxn, nrhs, tile=BATCH_TILE):
    """Useful FP64 FLOPs of one far-part launch of the batched convolution,
    averaged over the launches: for each band entry of lag j and each step i
    of the tile whose filtered state x = K0 + i - j lies in 0 .. K0 (the
    states solved before the tile), the 3x3 MGS and MTS block products of
    every load case: 18 FMA = 36 FLOPs per (entry, step, case)."""
    counts = np.asarray(counts, np.float64)
    j = np.arange(1, len(counts) + 1)
    total, launches = 0.0, 0
    for K0 in range(2, nstep + 1, tile):
        Tn = min(tile, nstep - K0 + 1)
        jmax = min(K0 + Tn - 1, maxn)
        lo = np.maximum(0, j - K0)
        hi = np.minimum(Tn - 1, j)
        n = np.where(j <= jmax, np.maximum(0, hi - lo + 1), 0)
        total += 36.0 * nrhs * float((counts * n).sum())
        launches += 1
    return total / launches


# Algorithmic FP64 flops of the history quadrature per evaluation (operation
# counts of the closed forms the staged kernels evaluate, divisions and square
# roots counted as one): a patch-wise T2 (entry, point) = the time-window
# coefficients (40) + one 3x3 block of basis products (27 FMA = 54); a
# patch-wise T1 + U (entry, point) = 40 + 54 + the U block (20); a regular
# (entry, point) = both levels' coefficients (80) + their sum (3) + 54 + 20;
# a staged (source, point) record = distance, powers, tensor scalars and the
# basis tensors (114).
QUAD_FLOPS = {"quad_evals_t2": 94, "quad_evals_t1u": 114, "quad_evals_regular": 157, "quad_records": 114}


def quadrature_flops(tm):
    return float(sum(QUAD_FLOPS[k] * tm[k] for k in QUAD_FLOPS))


def conv_far_bytes_per_launch(counts, npairs, ne, nhist, nstep, tile):
    """Algorithmic HBM bytes of one far-part launch of the single-case
    time-blocked convolution, averaged over launches: each band entry with
    lag j <= K0+Tn-1 read once (9 MTS + 6 MGS doubles), the per-pair band
    metadata (16 B), the solved state series read once (two histories, 3
    doubles per column and solved step) and Tn right-hand-side slices
    written."""
    cum = np.cumsum(counts)
    total, launches = 0.0, 0
    for K0 in range(2, nstep + 1, tile):
        Tn = min(tile, nstep - K0 + 1)
        j = min(K0 + Tn - 1, len(counts))
        total += cum[j - 1] * 120.0 + npairs * 16.0 + 2.0 * nhist * (K0 - 1) * 24.0 + Tn * ne * 24.0
        launches += 1
    return total / launches


def batched_job(tdbem, w, d_known, d_lu, d_lt, nrhs, tol, mixed, gmres_device, keep=False, fft_levels=0,
                fft_mfma=None):
    """One whole job of the batched form with device-resident inputs and outputs."""
    t0 = time.perf_counter()
    c = make_case(tdbem, w, mixed=mixed, gmres_device=gmres_device)
    if fft_levels:
        c.set_option("batched_fft_levels", fft_levels)
    if fft_mfma is not None:
        c.set_option("batched_fft_mfma", fft_mfma)
    c.assemble_step0(10)
    entries = c.assemble_history(w.nstep)
    tdbem.synchronize()
    t1 = time.perf_counter()
    it, fl, solved, times = c.solve_batched_device(d_known, nrhs, w.nstep, d_lu, d_lt, tol=tol)
    tdbem.synchronize()
    t2 = time.perf_counter()
    wall = t2 - t0
    info = dict(wall=wall, assembly=t1 - t0, loop=t2 - t1, entries=entries, iters=it, flags=fl, solved=solved,
                timings=c.timings())
    if keep:
        info["case"] = c
    else:
        c.close()
    return info


def batched_job_host(tdbem, w, cases, tol, mixed, gmres_device):
    """One whole job of the batched form through the host-buffer entry point
    (tdbem_solve_batched): the load histories cross PCIe in, the solutions out."""
    t0 = time.perf_counter()
    c = make_case(tdbem, w, mixed=mixed, gmres_device=gmres_device)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    lu, lt, it, fl, solved, times = c.solve_batched(cases, w.nstep, tol=tol)
    wall = time.perf_counter() - t0
    c.close()
    return dict(wall=wall, solved=solved, lu=lu)


def single_fp64_job(tdbem, w, d_known, d_lu, d_lt, tol, gmres_device):
    """One whole job of the plain FP64 form for one load case (device-resident I/O)."""
    t0 = time.perf_counter()
    c = make_case(tdbem, w, gmres_device=gmres_device)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    it, fl, solved, times = c.solve_device(d_known, w.nstep, d_lu, d_lt, tol=tol)
    tdbem.synchronize()
    wall = time.perf_counter() - t0
    tm = c.timings()
    c.close()
    return dict(wall=wall, iters=it, solved=solved, timings=tm)


def cpu_baseline(threads, gpu_tdbem, tol):
    """The oracle (C FP64 restatement of the reference path, OpenMP over the
    reference's independent loops) on a bounded sample of the headline's
    workload family: the stacked 2x2x8 box at h = 0.5 (512 elements, 8
    domains), 60 steps, one load case, whole job; with `threads` threads and
    with one, and the GPU's FP64 whole job on the same sample."""
    from tdbem import mesh
    w = mesh.mixed_config3(h=0.5, nstep=60)
    known = mesh.load_cases(w, 1)[0]
    m = w.mesh
    itf = m.interfaces.copy()
    itf[:, [0, 1, 2, 3]] -= 1
    out = {}
    for nt in (threads, 1):
        env = dict(os.environ, OMP_NUM_THREADS=str(nt))
        code = ("import sys, time, numpy as np; sys.path[:0] = %r; from oracle import oracle as O; "
                "from tdbem import mesh; w = mesh.mixed_config3(h=0.5, nstep=60); m = w.mesh; "
                "k = mesh.load_cases(w, 1)[0]; itf = m.interfaces.copy(); itf[:, :4] -= 1; "
                "t0 = time.perf_counter(); "
                "c = O.Case(m.points, m.ele, m.bcid, w.mats[:, 1:].copy(), itf.astype(np.int32), w.dt, w.gap); "
                "U, *_ = c.solve(k, w.nstep, tol=%r); dt = time.perf_counter() - t0; "
                "np.save(sys.argv[1], U); print(dt)") % ([str(REPO), str(REPO / "parallel3.0_omp_multidomin_1node_amd")], tol)
        upath = str(Path(os.environ.get("TMPDIR", "/tmp")) / f"tdbem_cpu_u_{os.getpid()}_{nt}.npy")
        r = subprocess.run([sys.executable, "-c", code, upath], capture_output=True, text=True, env=env, timeout=400)
        if r.returncode != 0:
            raise RuntimeError("cpu baseline failed: " + r.stderr[-500:])
        out[nt] = (float(r.stdout.strip().splitlines()[-1]), np.load(upath))
        os.remove(upath)
        log(f"cpu baseline {nt} thread(s): {out[nt][0]:.2f} s")
    # the GPU path on the same sample
    c = gpu_tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=8, mat_rows=w.mats, itf_rows=m.interfaces)
    t0 = time.perf_counter()
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    U, *_ = c.solve(known, w.nstep, tol=tol)
    gpu_dt = time.perf_counter() - t0
    c.close()
    Uc = out[threads][1]
    err = float(np.abs(U - Uc).max() / max(np.abs(Uc).max(), 1e-300))
    return {"value": round(w.nstep / out[threads][0], 4), "unit": "time steps/s", "cores": threads,
            "kind": "port",
            "sample": "oracle (C FP64 restatement of the reference path, OpenMP) whole job on the stacked 2x2x8 "
                      "box at h=0.5 (512 elements, 8 domains), 60 steps, one seeded load case, GMRES tol %g" % tol,
            "wall_s": round(out[threads][0], 3), "host_physical_cores": physical_cores(),
            "one_thread": {"value": round(w.nstep / out[1][0], 4), "wall_s": round(out[1][0], 3), "cores": 1},
            "gpu_same_sample_steps_per_s": round(w.nstep / gpu_dt, 3), "gpu_vs_cpu_max_rel_diff_u": err}


def config2_report(tdbem, gmres_device, far_parts=0):
    """The layered 2x2x4 block (6400 elements, 4 domains, 400 steps, Ricker
    load) in FP64, one load case, whole job, device-resident I/O, GMRES tol
    1e-5 (the reference card's): steps/s, and the HBM roofline of the
    single-case far convolution (k_conv_far_q) from its HIP-event time."""
    from tdbem import mesh
    w = mesh.layered_config2(h=0.1, nstep=400)
    d_known = tdbem.DeviceArray.from_host(w.known)
    d_lu, d_lt = tdbem.DeviceArray(w.known.shape), tdbem.DeviceArray(w.known.shape)
    import tempfile
    cache = os.path.join(tempfile.gettempdir(), f"tdbem_tune_{os.getpid()}.txt")
    if os.path.exists(cache):
        os.remove(cache)
    runs = []
    for i in range(3):
        t0 = time.perf_counter()
        c = make_case(tdbem, w, gmres_device=gmres_device)
        if far_parts:
            c.set_option("far_parts", far_parts)
        c.assemble_step0(10)
        entries = c.assemble_history(w.nstep)
        tuned = c.autotune(cache) if not far_parts else None
        c.solve_device(d_known, w.nstep, d_lu, d_lt, tol=1e-5)
        tdbem.synchronize()
        runs.append(time.perf_counter() - t0)
        if i < 2:
            c.close()
    if os.path.exists(cache):
        os.remove(cache)
    tm = c.timings()
    counts, npairs = c.history_stats(w.nstep)
    # the quadrature's evaluations, counted in one more assembly (not timed)
    c.set_option("quad_count", 1)
    c.assemble_history(w.nstep)
    qc = c.timings()
    c.set_option("quad_count", 0)
    qflops = quadrature_flops(qc)
    dof = c.dofmap()["counts"]
    c.close()
    nhist = int(dof[7] + dof[8]) // 2
    avg = tm["conv_far"] / max(tm["conv_far_launches"], 1)
    bpl = conv_far_bytes_per_launch(counts, npairs, len(w.mesh.ele), nhist, w.nstep, 16)
    fp = tuned[0] if tuned else far_parts
    wall = float(np.mean(runs[1:]))
    log(f"config 2 FP64: {wall:.3f} s per job")
    return {"workload": w.name, "elements": int(len(w.mesh.ele)), "domains": 4, "nstep": w.nstep, "gmres_tol": 1e-5,
            "dtype": "f64", "steps_per_s": round(w.nstep / wall, 3),
            "time_loop_steps_per_s": round(w.nstep / (tm["convolution"] + tm["gmres"]), 3),
            "history_band_entries": int(entries),
            "assembled_blocks_per_s": round(2.0 * float(entries) / max(tm["history"], 1e-9), 1),
            "breakdown_s": {k: round(float(v), 4) for k, v in tm.items()
                            if k not in NOT_SECONDS and not k.startswith("adjoint_")},
            "quadrature_roofline": {
                "kernel": "k_history_pw + k_history_reg (history quadrature, one assembly)", "bound": "fp64 valu",
                "achieved": round(qflops / tm["quadrature"] / 1e12, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(qflops / tm["quadrature"] / 1e12 / FP64_PEAK_TFLOPS, 4),
                "flops_per_assembly": qflops, "time_s": round(tm["quadrature"], 4),
                "evaluations": {k: qc[k] for k in ("quad_evals_t2", "quad_evals_t1u", "quad_evals_regular",
                                                   "quad_records", "quad_pw_entries", "quad_pw_waves")},
                "patchwise_lane_occupancy": round(qc["quad_pw_entries"] / max(64.0 * qc["quad_pw_waves"], 1.0), 4)},
            "roofline": {"kernel": f"k_conv_far_q ({fp} wavefronts per row, 16 steps per launch)",
                         "bound": "hbm", "achieved": round(bpl / avg / 1e9, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(bpl / avg / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": MEASURED_TRAFFIC.get((w.name, w.nstep, 16, fp)),
                         "avg_launch_ms": round(1000 * avg, 4), "bytes_per_launch": round(bpl)}}


def streamed_report(tdbem, gmres_device, budget_gb):
    """The layered 2x2x4 block (6400 elements, 400 steps, FP64) with its
    history streamed under a device budget of budget_gb (the bands take 64 GB):
    whole job, slab count, far part with the slabs' regeneration, and the
    largest difference from the resident solve relative to its largest value."""
    from tdbem import mesh
    w = mesh.layered_config2(h=0.1, nstep=400)
    t0 = time.perf_counter()
    c = make_case(tdbem, w, gmres_device=gmres_device, history_stream=1, history_budget_gb=budget_gb)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    U, T, it, fl, solved, _ = c.solve(w.known, w.nstep, tol=1e-5)
    tdbem.synchronize()
    wall = time.perf_counter() - t0
    tm = c.timings()
    c.close()
    r = make_case(tdbem, w, gmres_device=gmres_device)
    r.assemble_step0(10)
    r.assemble_history(w.nstep)
    U0, T0, *_ = r.solve(w.known, w.nstep, tol=1e-5)
    r.close()
    log(f"streamed config 2: {wall:.1f} s, {tm['stream_slabs']} slabs")
    return {"workload": w.name, "budget_gb": budget_gb, "slabs": tm["stream_slabs"], "whole_job_s": round(wall, 3),
            "steps_per_s": round(w.nstep / wall, 3), "history_near_s": round(tm["history"], 3),
            "far_with_regeneration_s": round(tm["stream_far"], 3), "gmres_s": round(tm["gmres"], 3),
            "max_rel_diff_u_vs_resident": float(np.abs(U - U0).max() / np.abs(U0).max()),
            "all_solved": bool(solved == w.nstep)}


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


def models_report(tdbem):
    """16 independent single-domain cube models of 500-6000 elements (seeded
    sizes, materials and loads, dt = 0.5 h / c1 each, 200 steps): every
    model's whole job one after another, then all of them through
    tdbem_run_models (one stream and host thread per model, history bands
    held at once within 80 % of the free device memory), then stacked in
    rounds (tdbem_run_models_grouped: every kernel one launch for all the
    models of a round, rounds sized to the same budget); the grouped results
    are checked against the concurrent ones (equal bit for bit)."""
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
    cs = [case(w) for w in ws]
    tdbem.synchronize()
    t3 = time.perf_counter()
    resg = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], grouped=True)
    tdbem.synchronize()
    t4 = time.perf_counter()
    rounds = tdbem.run_models.rounds
    for c in cs:
        c.close()
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
               for a, b in zip(res, resg))
    log(f"models: one after another {t1 - t0:.2f} s, together {t2 - t1:.2f} s, grouped {t4 - t3:.2f} s "
        f"({rounds} rounds, equal to the concurrent run: {same})")
    return {"models": len(ws), "elements": [int(len(w.mesh.ele)) for w in ws], "nstep": ws[0].nstep,
            "model_steps_per_s_sequential": round(steps / (t1 - t0), 2),
            "model_steps_per_s_together": round(steps / (t2 - t1), 2),
            "model_steps_per_s_grouped": round(steps / (t4 - t3), 2), "grouped_rounds": rounds,
            "grouped_equals_together": bool(same),
            "speedup": round((t1 - t0) / (t2 - t1), 3), "speedup_grouped": round((t1 - t0) / (t4 - t3), 3),
            "all_solved": all(r[4] == w.nstep for r, w in zip(res, ws))}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed whole jobs of the 64-case batch")
    ap.add_argument("--warmup", type=int, default=1, help="untimed whole jobs")
    ap.add_argument("--nrhs", type=int, default=64, help="load cases per batch")
    ap.add_argument("--h", type=float, default=0.125, help="element size of the headline box (0.125: 8192 elements)")
    ap.add_argument("--nstep", type=int, default=600)
    ap.add_argument("--tol", type=float, default=1e-10, help="GMRES tolerance of both forms")
    ap.add_argument("--mixed", type=int, default=2, choices=(0, 1, 2),
                    help="band storage of the batched form: 0 FP64, 1 FP32, 2 FP32 + bf16 residual")
    ap.add_argument("--gmres-device", type=int, default=1, choices=(0, 1, 2),
                    help="GMRES driver of the single solves: 0 host loop, 1 device state + hipGraphs, "
                         "2 device state + direct launches (for profiler runs)")
    ap.add_argument("--error-cases", type=str, default="0,21,42,63",
                    help="cases of the timed batch checked against their FP64 single solves")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the config-2 and adjoint blocks")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint block")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the secondary blocks (profiling them; the JSON line then has no headline value)")
    ap.add_argument("--far-parts", type=int, default=0,
                    help="config-2 block: wavefronts per source row of the far convolution (0: autotuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES device to device once (a known-size dispatch for calibrating the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--streamed", type=float, default=0.0, metavar="GB",
                    help="also run config 2 with its history streamed under this device budget (several minutes)")
    ap.add_argument("--no-longer-levels", dest="longer_levels", action="store_false",
                    help="skip the extra job of the headline batch with three spectral levels")
    ap.add_argument("--fft-levels", type=int, default=0,
                    help="levels of the batch's spectral far part (0: the library's default)")
    ap.add_argument("--fft-mfma", type=int, default=None, choices=(0, 1),
                    help="spectral products: 1 on the matrix cores (the library's default), 0 with vector FMAs")
    ap.add_argument("--host-buffers", action="store_true",
                    help="also time one batched job through the host-buffer entry point (PCIe-inclusive rate)")
    ap.add_argument("--models", action="store_true",
                    help="also run the 16-model report (about 1.3 TB of history bands in total; about a minute)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only; WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)
    t_start = time.perf_counter()

    import tdbem
    from tdbem import mesh

    if args.copy_calibration > 0:
        n = (args.copy_calibration + 7) // 8
        src, dst = tdbem.DeviceArray((n,)), tdbem.DeviceArray((n,))
        if tdbem.lib().tdbem_memcpy(dst.ptr, src.ptr, 8 * n, 2) != 0:
            raise RuntimeError("calibration copy failed")
        tdbem.synchronize()
        del src, dst
        log(f"calibration copy: {8 * n} bytes device to device")
    if args.skip_headline:
        out = {"metric": "secondary blocks only (--skip-headline)", "value": None, "n_gpus": 1}
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline(args.cpu_threads, tdbem, args.tol)
        if not args.no_secondary:
            out["config2_fp64"] = config2_report(tdbem, args.gmres_device, args.far_parts)
            if not args.no_adjoint:
                out["adjoint"] = adjoint_report(tdbem, gmres_device=args.gmres_device)
        if args.streamed > 0:
            out["streamed"] = streamed_report(tdbem, args.gmres_device, args.streamed)
        if args.models:
            out["models"] = models_report(tdbem)
        print(json.dumps(out), flush=True)
        return

    w = mesh.mixed_config3(h=args.h, nstep=args.nstep)
    B = args.nrhs
    cases = mesh.load_cases(w, B)
    log(f"workload {w.name}: {len(w.mesh.ele)} elements, {w.nstep} steps, {B} load cases, dt={w.dt:.6g}")
    # inputs resident in HBM before the timed region; outputs stay there
    d_known = tdbem.DeviceArray.from_host(cases)
    d_lu, d_lt = tdbem.DeviceArray(cases.shape), tdbem.DeviceArray(cases.shape)
    for i in range(args.warmup):
        r = batched_job(tdbem, w, d_known, d_lu, d_lt, B, args.tol, args.mixed, args.gmres_device,
                        fft_levels=args.fft_levels, fft_mfma=args.fft_mfma)
        log(f"warmup {i}: {r['wall']:.2f} s")
    tdbem.synchronize()
    t0 = time.perf_counter()
    runs = []
    for i in range(args.steps):
        runs.append(batched_job(tdbem, w, d_known, d_lu, d_lt, B, args.tol, args.mixed, args.gmres_device,
                                keep=(i == args.steps - 1), fft_levels=args.fft_levels,
                                fft_mfma=args.fft_mfma))
        log(f"step {i}: {runs[-1]['wall']:.2f} s")
    tdbem.synchronize()
    elapsed = time.perf_counter() - t0
    last = runs[-1]
    value = args.steps * B * w.nstep / elapsed
    tm = last["timings"]
    c = last["case"]
    counts, npairs = c.history_stats(w.nstep)
    c.close()
    spectral = bool(tm.get("spectral"))
    # with the spectral far part the lag-by-lag kernel takes the raw lags <= 16
    far_avg = tm["conv_far"] / max(tm["conv_far_launches"], 1)
    flops = batched_far_flops_per_launch(counts, w.nstep, 16 if spectral else w.nstep, B)
    far_tflops = flops / far_avg / 1e12
    all_solved = bool((last["solved"] == w.nstep).all())
    # the timed batch's results, then the plain FP64 form of the same box
    BU, BT = d_lu.download(), d_lt.download()
    # the same batch with sub-filters of 16, 32 and 64 lags (one whole job,
    # timed on its own; its errors against the same FP64 jobs below)
    longer = None
    if args.longer_levels and not args.fft_levels:
        r3 = batched_job(tdbem, w, d_known, d_lu, d_lt, B, args.tol, args.mixed, args.gmres_device, fft_levels=3,
                         fft_mfma=args.fft_mfma)
        longer = {"wall": r3["wall"], "timings": r3["timings"], "U": d_lu.download(), "T": d_lt.download()}
        log(f"three spectral levels: {r3['wall']:.2f} s")
    del d_known, d_lu, d_lt
    ek = [int(x) for x in args.error_cases.split(",") if x.strip() and int(x) < B]
    d1 = tdbem.DeviceArray(cases.shape[1:])
    o1u, o1t = tdbem.DeviceArray(cases.shape[1:]), tdbem.DeviceArray(cases.shape[1:])
    fp64_runs, eu, et, eu3, et3 = [], 0.0, 0.0, 0.0, 0.0
    marks = sorted({k for k in (w.nstep // 4, w.nstep // 2, (3 * w.nstep) // 4, (9 * w.nstep) // 10, w.nstep) if k > 0})
    prof_u = {k: 0.0 for k in marks}
    for j, b in enumerate(ek if ek else [0]):
        d1.upload(cases[b])
        r = single_fp64_job(tdbem, w, d1, o1u, o1t, args.tol, args.gmres_device)
        fp64_runs.append(r)
        if ek:
            U, T = o1u.download(), o1t.download()
            eu = max(eu, float(np.abs(BU[b] - U).max() / np.abs(U).max()))
            et = max(et, float(np.abs(BT[b] - T).max() / np.abs(T).max()))
            for k in marks:  # the largest difference over steps <= k
                prof_u[k] = max(prof_u[k], float(np.abs(BU[b, :k + 1] - U[:k + 1]).max() / np.abs(U).max()))
            if longer:
                eu3 = max(eu3, float(np.abs(longer["U"][b] - U).max() / np.abs(U).max()))
                et3 = max(et3, float(np.abs(longer["T"][b] - T).max() / np.abs(T).max()))
        log(f"fp64 single job (case {b}): {r['wall']:.3f} s")
    del d1, o1u, o1t, BU, BT
    if longer:
        del longer["U"], longer["T"]
    fp64_wall = float(np.mean([r["wall"] for r in fp64_runs[1:]] if len(fp64_runs) > 1 else [fp64_runs[0]["wall"]]))
    fp64_sps = w.nstep / fp64_wall
    ftm = fp64_runs[-1]["timings"]
    out = {
        "metric": "time steps/s of the mixed-precision + batched form (64 load cases, summed) on the 8-domain "
                  "8192-element 600-step box, whole job, vs the plain FP64 form at <=1e-6 relative displacement error",
        "value": round(value, 3),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 arithmetic; bands stored f32 + bf16 residual" if args.mixed == 2 else
                 ("f64 arithmetic; bands stored f32" if args.mixed == 1 else "f64"),
        "data": "synthetic (seeded stacked box, 64 seeded load cases; inputs resident in HBM)",
        "config": {"workload": f"{w.name} x {B} load cases, mixed + batched", "elements": int(len(w.mesh.ele)),
                   "domains": int(w.mesh.domain_count), "nstep": w.nstep, "load_cases": B, "dt": w.dt,
                   "gmres_restart": 200, "gmres_tol": args.tol, "band_storage": ("fp64", "fp32", "fp32+bf16")[args.mixed],
                   "history_band_entries": int(last["entries"])},
        "vs_plain_fp64": {"ratio": round(value / fp64_sps, 3), "target": 10.0,
                          "plain_fp64_steps_per_s": round(fp64_sps, 3), "plain_fp64_whole_job_s": round(fp64_wall, 4),
                          "plain_fp64_breakdown_s": {k: round(float(v), 4) for k, v in ftm.items()
                                                     if k in ("history", "step0", "convolution", "gmres")},
                          "max_rel_err_u": eu, "max_rel_err_t": et, "error_cases": ek,
                          "max_rel_err_u_up_to_step": {str(k): v for k, v in prof_u.items()},
                          "error_bound": 1e-6, "same_gmres_tol": args.tol},
        "time_loop_case_steps_per_s": round(B * w.nstep / last["loop"], 3),
        "assembly_s": round(last["assembly"], 4),
        "assembled_blocks_per_s": round(2.0 * last["entries"] / max(float(tm.get("history", 0.0)), 1e-9), 1),
        "all_solved": all_solved,
        "gmres_inner_iterations_mean": float(np.mean(last["iters"])),
        "longer_sub_filters": None if not longer else {
            "levels": 3, "sub_filter_lags": [16, 32, 64],
            "case_steps_per_s": round(B * w.nstep / longer["wall"], 3),
            "vs_plain_fp64": round(B * w.nstep / longer["wall"] / fp64_sps, 3),
            "max_rel_err_u": eu3, "max_rel_err_t": et3, "error_cases": ek,
            "conv_spectral_products_s": round(float(longer["timings"].get("conv_spectral_products", 0.0)), 4),
            "note": "one whole job of the same batch with option batched_fft_levels 3, timed alone; not the "
                    "headline because its traction error exceeds 1e-6 (the displacement error, the bound, does not)"},
        "breakdown_s": {k: round(float(v), 4) for k, v in tm.items()
                        if k not in NOT_SECONDS and not k.startswith("adjoint_")},
    }
    lag_roof = {"kernel": f"k_conv_far_b (lag-by-lag far part, {BATCH_TILE} steps x {B} load cases per launch"
                          + (", raw lags <= 16)" if spectral else ")"),
                "bound": "mfma", "pipe": "fp64 (vector FMA; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                "achieved": round(far_tflops, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(far_tflops / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "avg_launch_ms": round(1000 * far_avg, 4), "flops_per_launch": flops,
                "launches": int(tm["conv_far_launches"])}
    if spectral:
        # the dominant kernel: the products of the level that takes longest
        lvs = tm["spectral_level_stats"]
        top = max(lvs, key=lambda v: v["products"])
        sl = max(top["launches"], 1)
        prod_avg = top["products"] / sl
        sflops = top["flops"] / sl
        L = top["L"]
        out["roofline"] = {
            "kernel": f"k_conv_far_fft_mf<{L}> (spectral products of the batched far part on the matrix cores, "
                      f"sub-filters of {L} lags, {L} steps x {B} load cases per launch; flops counted without the zero "
                      f"rows and bins)",
            "bound": "mfma",
            "pipe": "fp64 matrix cores (v_mfma_f64_16x16x4f64; FP64 matrix and vector peaks are both 78.6 TFLOP/s)",
            "achieved": round(sflops / prod_avg / 1e12, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(sflops / prod_avg / 1e12 / FP64_PEAK_TFLOPS, 4),
            "traffic": HEADLINE_TRAFFIC.get((w.name, w.nstep, B, args.mixed, L)),
            "avg_launch_ms": round(1000 * prod_avg, 4), "flops_per_launch": round(sflops),
            "algorithmic_bytes_per_launch": round(top["bytes"] / sl), "launches": sl}
        out["spectral_levels"] = [
            {"L": v["L"], "launches": v["launches"], "products_s": round(v["products"], 4),
             "avg_launch_ms": round(1000 * v["products"] / max(v["launches"], 1), 4),
             "tflops": round(v["flops"] / max(v["products"], 1e-12) / 1e12, 3)} for v in lvs]
        out["lag_by_lag_far"] = lag_roof
    else:
        lag_roof["traffic"] = HEADLINE_TRAFFIC.get((w.name, w.nstep, B, args.mixed))
        out["roofline"] = lag_roof
    if args.host_buffers:
        hb = batched_job_host(tdbem, w, cases, args.tol, args.mixed, args.gmres_device)
        log(f"host-buffer job: {hb['wall']:.2f} s")
        out["host_buffers"] = {"entry": "tdbem_solve_batched (load histories in and solutions out over PCIe)",
                               "whole_job_s": round(hb["wall"], 3),
                               "case_steps_per_s": round(B * w.nstep / hb["wall"], 3),
                               "bytes_in": int(cases.nbytes), "bytes_out": int(2 * cases.nbytes),
                               "all_solved": bool((hb["solved"] == w.nstep).all())}
        del hb
    if not args.no_cpu_baseline:
        log(f"cpu baseline on {args.cpu_threads} threads and on 1 ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_threads, tdbem, args.tol)
    if not args.no_secondary and time.perf_counter() - t_start < SECONDARY_BUDGET_S:
        out["config2_fp64"] = config2_report(tdbem, args.gmres_device, args.far_parts)
    if not args.no_secondary and not args.no_adjoint and time.perf_counter() - t_start < SECONDARY_BUDGET_S:
        out["adjoint"] = adjoint_report(tdbem, gmres_device=args.gmres_device)
    if args.streamed > 0:
        out["streamed"] = streamed_report(tdbem, args.gmres_device, args.streamed)
    if args.models:
        out["models"] = models_report(tdbem)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
