"""Benchmark of the plain FP64 time step on one MI355X.

Workload (default): one synthetic 4D volume 256x256x128x16 (x,y,z,t), six
ellipsoids with time-periodic radii and wedge gaps, noise sigma 25; edge
indicator with K = 1000 after Gaussian presmoothing (sigma = 1 voxel); then
30 semi-implicit time steps (eps = 1e-3, tau = 0.5, h = 1), each solved by
red-black SOR (omega = 1.8) to relative residual 1e-6, in FP64.

One bench step = one time step (coefficient assembly + SOR to tolerance +
epilogue) with every array resident in HBM.  The timed region covers
exactly --steps time steps from the initial function, after --warmup
untimed steps (the state is reset to u0 after the warm-up).

Reported:
  value          SOR voxel-updates/s = interior voxels x SOR iterations / time
  roofline       the red SOR pass with the fused residual (k_sor_sweep<1>):
                 96 algorithmic bytes per red point per launch, divided by its
                 average launch duration from HIP events on the library's
                 stream over the timed region; peak 8 TB/s (MI355X HBM3E)
  cpu_baseline   the oracle's OpenMP restatement of the reference path
                 (OpenMP over the t axis, as the reference's OpenMP program),
                 timed on a bounded sample on this host: one coefficient pass
                 plus a fixed number of SOR iterations of step 1, extrapolated
                 with the GPU run's mean sweeps per step.

The same time steps are then run in the FP32-refined form (FP32 SOR inside
FP64 defect correction, same FP64 stopping test) from u^n, from the linear and
the cubic extrapolation of the previous solutions, and from the learned
network's prediction, each reported with its time, sweeps, max|u - u_FP64| and
mask IoU; then three non-parity FP64 controls (SOR factor 1.2, the cubic
start, both) that separate what the factor, the start and FP32 contribute;
then a few steps in mirror mode (the drop-in default, PCIe-inclusive).  The
edge indicator is timed both PCIe-inclusive (mirror mode) and resident.

Usage: python bench.py [--steps 30] [--warmup 1] [--no-cpu] [--cpu-iters 10]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "4d-image-segmentation-algorithm_amd")
sys.path.insert(0, PKG)

METRIC = "SOR voxel-updates/s + HBM GB/s; time-to-tolerance plain FP64 vs mixed+learned"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 96.0  # 8 coefficients + prev + u in/out + other colour, FP64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_traffic(dims):
    """HBM bytes per launch of k_sor_sweep<1> measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), as committed in
    profiles/sor_sweep_pmc.md; None if that summary is absent or was measured
    on another volume than dims = (nx, ny, nz, nt)."""
    path = os.path.join(ROOT, "profiles", "sor_sweep_pmc.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", text)
    v = re.search(r"traffic_volume\s*=\s*(\d+)x(\d+)x(\d+)x(\d+)", text)
    if not m or not v or tuple(int(x) for x in v.groups()) != tuple(dims):
        return None
    return float(m.group(1))


def params(lib):
    lib.set(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
            theta_img=1.0, rho_thr=0.0, gsTol=0.0, maxNoGSIteration=100000,
            maxNoGSIteration_sigma=100000.0)
    lib.set_relative_tolerance(1e-6)


def host_cores():
    """(physical cores, logical CPUs) this process may run on, from the
    kernel's CPU topology (distinct core_id per package)."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    phys = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            pk = open(f"{base}/physical_package_id").read().strip()
            co = open(f"{base}/core_id").read().strip()
            phys.add((pk, co))
        except OSError:
            phys.add(("?", str(c)))
    return len(phys), len(cpus)


def cpu_baseline(lib, shape, u0, iters, mean_sweeps, nint):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harness import Oracle, params as oparams, zeros

    O = Oracle()
    phys, logical = host_cores()
    # the reference OpenMP program parallelises the t loop: at most lMax - 2
    # threads do work; one thread per physical core, at most 16
    threads = max(1, min(phys, 16, shape[0] - 2))
    p = oparams(shape, h=1.0, tau=0.5, eps2=1e-6, omega_c=1.8, gsTol=-1.0,
                maxNoGSIteration=iters, nthreads=threads, par_reduce=1)
    g = [np.ascontiguousarray(lib.arr("g_" + f)) for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
    u = u0.copy()
    prev = u0.copy()
    coef = zeros(shape, 8)
    t0 = O.L.orc_wtime()
    O.coefficients(p, u, g, coef)
    t1 = O.L.orc_wtime()
    O.sor(p, u, prev, coef)
    t2 = O.L.orc_wtime()
    t_coef, t_iter = t1 - t0, (t2 - t1) / iters
    step_s = t_coef + mean_sweeps * t_iter
    return {
        "value": nint * mean_sweeps / step_s,
        "unit": "voxel-updates/s",
        "cores": threads,
        "physical_cores_available": phys,
        "logical_cpus_available": logical,
        "kind": "port",
        "sample": (f"oracle OpenMP restatement of the reference path: 1 coefficient pass "
                   f"({t_coef:.2f} s) + {iters} SOR iterations ({t_iter:.3f} s each) of "
                   f"time step 1 on the same volume; per-step time extrapolated with "
                   f"{mean_sweeps:.1f} sweeps/step"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--preset", type=int, default=1, help="synthetic volume preset 0-4 (the generator's workloads, 64x64x32x8 to 512x512x256x8)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--variant", type=int, default=None, help="SOR pass kernel variant")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the FP32-refined run of the same time steps")
    ap.add_argument("--net", default=None,
                    help="weights of the learned warm start (default: the package's "
                         "weights/warmstart.bin when present)")
    ap.add_argument("--train-net", action="store_true",
                    help="train the learned warm start first (solver runs + Adam) and "
                         "save it to --net-out")
    ap.add_argument("--net-out", default=None)
    ap.add_argument("--train-iters", type=int, default=2500)
    ap.add_argument("--train-runs", type=int, default=6)
    ap.add_argument("--train-dims", type=int, nargs=4, default=None,
                    help="interior size nx ny nz nt of the training runs (default: warmstart.py's)")
    ap.add_argument("--net-layers", type=int, default=1,
                    help="layers of a network trained with --train-net")
    ap.add_argument("--net-base", type=int, default=3, choices=(0, 1, 3),
                    help="base a network trained with --train-net corrects (0: u^n itself, "
                         "inputs u^n and g; 1: linear, 3: cubic extrapolation)")
    ap.add_argument("--inner-omega", type=float, default=None,
                    help="SOR factor of the FP32 correction solves (default: omega_c)")
    ap.add_argument("--batch", action="store_true",
                    help="the batched workload instead: 32 volumes of uneven size")
    ap.add_argument("--batch-volumes", type=int, default=32)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-streams", default="2,4",
                    help="batched workload: group counts of the concurrent form to time")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched workload: skip the one-volume-at-a-time comparison")
    ap.add_argument("--no-controls", action="store_true",
                    help="skip the non-parity control runs (FP64 at omega 1.2 / from the "
                         "cubic start) that attribute the refined forms' speed-up")
    ap.add_argument("--mirror-steps", type=int, default=2,
                    help="time steps timed in mirror mode (host arrays, PCIe) after the runs")
    ap.add_argument("--fused", action="store_true",
                    help="after the timed run, time the same FP64 steps again with fused sweeps "
                         "(each solve one cooperative launch) and check them bitwise")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, one device copy of u (padded volume, "
                         "FP64) as the known-size reference for rocprofv3 --pmc byte counts")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import seg4d
    import warmstart

    lib = seg4d.SegLib()
    lib.set_verbose(False)
    if args.batch:
        run_batch_bench(args, lib)
        return
    net_path = args.net or warmstart.DEFAULT_WEIGHTS
    training = None
    if args.train_net and not args.no_mixed:
        log("[bench] training the learned warm start ...")
        kw = {"dims": tuple(args.train_dims)} if args.train_dims else {}
        training = warmstart.train(lib, runs=args.train_runs, iters=args.train_iters,
                                   out=args.net_out or net_path, layers=args.net_layers,
                                   base=args.net_base, log=log, **kw)
        net_path = args.net_out or net_path
        log(f"[bench] trained: {training}")
    spec = lib.gen_spec(args.preset, args.preset)
    nx, ny, nz, nt = spec.nx, spec.ny, spec.nz, spec.nt
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    nint = nx * ny * nz * nt
    ta = time.perf_counter()
    lib.allocate(nx, ny, nz, nt)  # includes page-locking and zero-filling the host arrays
    alloc_s = time.perf_counter() - ta
    params(lib)
    if args.variant is not None:
        lib.set_sweep_variant(args.variant)
    log(f"[bench] volume {nx}x{ny}x{nz}x{nt} on {lib.device_info()}; allocated in {alloc_s:.2f} s")
    t = time.perf_counter()
    u0 = np.zeros(shape)
    lib.generate(spec, lib.arr("i0"), u0)
    lib.arr("u")[...] = 0.0
    lib.arr("Mmask")[...] = 1.0
    log(f"[bench] generated in {time.perf_counter() - t:.2f} s")

    # edge indicator, as a reference driver gets it (mirror mode: the host
    # arrays in and the presmoothed images, prev_u and the 8 g arrays out
    # over PCIe), then again with the arrays resident (device time)
    img0 = lib.arr("i0").copy()
    t = time.perf_counter()
    lib.gf("i0", "u")
    edge_s = time.perf_counter() - t
    edge_iters = lib.lib.seg4dLastSweeps()
    lib.set_resident(True)
    lib.arr("i0")[...] = img0
    lib.arr("u")[...] = 0.0
    lib.upload("i0")
    lib.upload("u")
    lib.synchronize()
    t = time.perf_counter()
    lib.gf("i0", "u")
    lib.synchronize()
    edge_dev_s = time.perf_counter() - t
    del img0
    log(f"[bench] edge indicator {edge_s:.2f} s PCIe-inclusive, {1e3 * edge_dev_s:.1f} ms resident")

    lib.arr("u")[...] = u0
    lib.upload("u")
    lib.upload("Mmask")
    for w in range(args.warmup):
        lib.set(numberOfTimeStep=w + 1)
        lib.seg_time_step()
    lib.arr("u")[...] = u0
    lib.upload("u")
    lib.synchronize()
    log("[bench] warm-up done")

    lib.profile(True)
    lib.profile_reset()
    sweeps = []
    lib.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        lib.set(numberOfTimeStep=s + 1)
        z, _ = lib.seg_time_step()
        sweeps.append(z)
    lib.synchronize()
    t1 = time.perf_counter()
    prof = lib.profile_read()
    lib.profile(False)
    nred, nblack = prof["nred"], prof["nblack"]
    total = t1 - t0
    log(f"[bench] {args.steps} steps in {total:.3f} s, sweeps/step {sweeps}")
    if args.calib_copy:
        lib.copy("u", "prev_u")  # resident mode: one device-to-device copy
        lib.synchronize()
        log(f"[bench] calibration copy of {8 * np.prod(shape)} bytes done")
    # same-box streaming reference: the library's FP64 copy of the padded
    # volume (one 8-byte load and store per lane, as the passes), HIP events
    lib.profile(True)
    lib.profile_reset()
    for _ in range(6):
        lib.copy("u", "prev_u")
    cprof = lib.profile_read()
    lib.profile(False)
    copy_gbps = (2 * 8 * float(np.prod(shape)) * cprof["copy_launches"] /
                 (cprof["copy_ms"] / 1e3) / 1e9) if cprof["copy_ms"] > 0 else None

    lib.download("u")
    u = lib.arr("u")
    inner = u[1:-1, 1:-1, 1:-1, 1:-1]
    thr = 0.5 * (inner.max() + inner.min())
    mask64 = inner >= thr
    mask_frac = float(mask64.mean())
    u64 = inner.copy()

    fused = None
    if args.fused:
        # the same FP64 steps with every solve in one cooperative launch (not
        # profiled: per-pass events would split the launch)
        lib.set_fused_sweeps(1)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.set(numberOfTimeStep=1)
        lib.seg_time_step()
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.synchronize()
        fsweeps, fflags = [], []
        tf0 = time.perf_counter()
        for s in range(args.steps):
            lib.set(numberOfTimeStep=s + 1)
            z, _ = lib.seg_time_step()
            fsweeps.append(z)
            fflags.append(lib.last_solve_fused())
        lib.synchronize()
        tf = time.perf_counter() - tf0
        lib.set_fused_sweeps(0)
        lib.download("u")
        same = bool(np.array_equal(lib.arr("u")[1:-1, 1:-1, 1:-1, 1:-1].view(np.uint64),
                                   u64.view(np.uint64))) and fsweeps == sweeps
        fused = {"time_s": tf, "ms_per_step": tf / args.steps * 1e3,
                 "speed_vs_pass_per_launch": total / tf, "all_solves_fused": all(fflags),
                 "bitwise_equal_to_pass_per_launch": same}
        log(f"[bench] fused sweeps: {tf:.3f} s ({total / tf:.2f}x the pass-per-launch solve), "
            f"bitwise equal {same}")

    def run_form(precision, warm, omega=1.8, label=None):
        """The same time steps from the same u0 in another form: precision 1 =
        FP32-refined (FP32 SOR inside FP64 defect correction, FP64 stopping
        test); warm 1 / 4 = linear / cubic extrapolation start, 2 = the
        learned network's prediction; omega = the FP64 solve's SOR factor
        (1.8 is the workload's; other values are non-parity controls).
        Timed like the plain run; error and mask against its final u."""
        lib.set_precision(precision)
        lib.set_warm_start(warm)
        lib.set(omega_c=omega)
        om_keep = lib.lib.seg4dSetInnerOmega(
            args.inner_omega if args.inner_omega is not None else -1.0)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.set(numberOfTimeStep=1)
        lib.seg_time_step()  # warm-up of the form's buffers and kernels
        lib.arr("u")[...] = u0
        lib.upload("u")  # also discards the time history
        lib.synchronize()
        lib.profile(True)
        lib.profile_reset()
        msweeps, mrounds = [], []
        pms0, pn0 = lib.net_predict_stats()
        t2 = time.perf_counter()
        for s in range(args.steps):
            lib.set(numberOfTimeStep=s + 1)
            z, _ = lib.seg_time_step()
            msweeps.append(z)
            mrounds.append(lib.last_rounds())
        lib.synchronize()
        t3 = time.perf_counter()
        mprof = lib.profile_read()
        lib.profile(False)
        pms1, pn1 = lib.net_predict_stats()
        lib.set_precision(0)
        lib.set_warm_start(0)
        lib.set(omega_c=1.8)
        inner_om = lib.lib.seg4dSetInnerOmega(om_keep)
        lib.download("u")
        um = lib.arr("u")[1:-1, 1:-1, 1:-1, 1:-1]
        thm = 0.5 * (um.max() + um.min())
        maskm = um >= thm
        inter = int(np.logical_and(mask64, maskm).sum())
        union = int(np.logical_or(mask64, maskm).sum())
        r32 = mprof["red1_f32_ms"] / max(mprof["red1_f32_launches"], 1)
        starts = {0: "u^n", 1: "2u^n - u^(n-1)", 2: "learned",
                  4: "4u^n - 6u^(n-1) + 4u^(n-2) - u^(n-3)"}
        res = {
            "precision": "fp32-refined" if precision else "f64",
            "warm_start": starts[warm],
            "sor_omega": (inner_om if inner_om > 0 else omega) if precision else omega,
            "parity_form": precision == 0 and warm == 0 and omega == 1.8,
            "time_to_tolerance_s": t3 - t2,
            "ms_per_step": (t3 - t2) / args.steps * 1e3,
            "speedup_vs_fp64": total / (t3 - t2),
            "sweeps_per_step": msweeps,
            "sweeps_total": int(sum(msweeps)),
            "rounds_per_step": mrounds if precision else None,
            "setup_ms_per_step": mprof["setup_ms"] / max(mprof["setup_launches"], 1),
            "red_pass_f32_avg_ms": r32 if precision else None,
            "red_pass_f32_gbps": (nred * 48.0 / (r32 / 1e3) / 1e9) if precision and r32 > 0
            else None,
            "defect_ms_per_step": mprof["defect_ms"] / args.steps if precision else None,
            "predict_ms_per_step": (pms1 - pms0) / max(pn1 - pn0, 1) if warm == 2 else None,
            "max_abs_du_vs_fp64_final": float(np.abs(um - u64).max()),
            "max_abs_u_fp64_final": float(np.abs(u64).max()),
            "mask_iou_vs_fp64_final": inter / union if union else 1.0,
            "mask_voxels_differing": union - inter,
            "reference": ("final u of the plain FP64 GPU path, bitwise equal to the FP64 oracle "
                          "(tests: this workload's edge indicator and first steps, all 20 steps "
                          "at 64x64x32x8)"),
        }
        name = label or f"{res['precision']} start {res['warm_start']} omega {res['sor_omega']}"
        log(f"[bench] {name}: {t3 - t2:.3f} s ({total / (t3 - t2):.2f}x), sweeps {msweeps}, "
            f"IoU {res['mask_iou_vs_fp64_final']:.5f}")
        return res

    mixed = mixed_warm = mixed_cubic = learned = None
    controls = None
    if not args.no_mixed:
        mixed = run_form(1, 0)
        mixed_warm = run_form(1, 1)
        mixed_cubic = run_form(1, 4)
        if os.path.exists(net_path):
            lib.net_load(net_path)
            learned = run_form(1, 2)
            layers, base = lib.net_shape()
            learned["network"] = {"layers": layers, "channels": 16,
                                  "corrects": {0: "u^n (inputs u^n, g)", 1: "2u^n - u^(n-1)",
                                               3: "4u^n - 6u^(n-1) + 4u^(n-2) - u^(n-3)"}[base]}
            learned["weights"] = os.path.relpath(net_path, ROOT)
            learned["training"] = training
            learned["sweeps_saved_vs_start_un"] = mixed["sweeps_total"] - learned["sweeps_total"]
            learned["sweeps_saved_vs_cubic_start"] = (mixed_cubic["sweeps_total"] -
                                                      learned["sweeps_total"])
        else:
            log(f"[bench] no learned warm start weights at {net_path}: skipped")
        if not args.no_controls:
            # attribution of the refined forms' speed-up (non-parity: the
            # workload fixes omega 1.8 and the start u^n for the FP64 path)
            controls = {
                "fp64_omega_1.2": run_form(0, 0, 1.2),
                "fp64_cubic_start": run_form(0, 4, 1.8),
                "fp64_omega_1.2_cubic_start": run_form(0, 4, 1.2),
            }

    # mirror mode (the drop-in default: host arrays authoritative, u and Mmask
    # up and u, prev_u down every step) on a few steps
    mirror = None
    if args.mirror_steps > 0:
        lib.set_resident(False)
        lib.arr("u")[...] = u0
        ms, l2ms = [], []
        for s in range(args.mirror_steps):
            lib.set(numberOfTimeStep=s + 1)
            t = time.perf_counter()
            lib.seg_time_step()
            ms.append((time.perf_counter() - t) * 1e3)
            # the reference driver's stopping test after every step
            # (segmentation4D.c:1408): both host arrays go up again
            t = time.perf_counter()
            lib.l2norm("prev_u", "u")
            l2ms.append((time.perf_counter() - t) * 1e3)
        mirror = {"ms_per_step": ms, "resident_ms_per_step_same_steps":
                  total / args.steps * 1e3,
                  "pcie_bytes_per_step": 3 * 8 * int(np.prod(shape)),  # u up, Mmask up, u down
                  "l2norm_ms_after_step": l2ms,
                  "driver_loop_ms_per_step": [a + b for a, b in zip(ms, l2ms)]}
        log(f"[bench] mirror mode: {[round(x, 1) for x in ms]} ms per step, "
            f"_l2norm_(prev_u, u) after it {[round(x, 1) for x in l2ms]} ms")
        # the same loop with host-write tracking: after the first step only u
        # comes down (u, prev_u and Mmask stay equal to their device copies)
        u_plain = lib.arr("u").copy()
        lib.set_write_tracking(True)
        lib.arr("u")[...] = u0
        tms, tl2 = [], []
        for s in range(args.mirror_steps):
            lib.set(numberOfTimeStep=s + 1)
            t = time.perf_counter()
            lib.seg_time_step()
            tms.append((time.perf_counter() - t) * 1e3)
            t = time.perf_counter()
            lib.l2norm("prev_u", "u")
            tl2.append((time.perf_counter() - t) * 1e3)
        mirror["write_tracking"] = {
            "active": bool(lib.write_tracking_state() & 1),
            "ms_per_step": tms, "l2norm_ms_after_step": tl2,
            "driver_loop_ms_per_step": [a + b for a, b in zip(tms, tl2)],
            "bitwise_equal_to_untracked": bool(np.array_equal(lib.arr("u"), u_plain))}
        lib.set_write_tracking(False)
        log(f"[bench] mirror mode with write tracking: {[round(x, 1) for x in tms]} ms per "
            f"step, _l2norm_ after it {[round(x, 1) for x in tl2]} ms, bitwise "
            f"{mirror['write_tracking']['bitwise_equal_to_untracked']}")
        lib.set_resident(True)

    red1_avg_s = prof["red1_ms"] / max(prof["red1_launches"], 1) / 1e3
    achieved = nred * BYTES_PER_UPDATE / red1_avg_s / 1e9
    sweep_ms = prof["red0_ms"] + prof["red1_ms"] + prof["black_ms"]
    it = prof["iterations"]
    sweep_bytes = BYTES_PER_UPDATE * ((it + args.steps) * nred + it * nblack)
    traffic = load_traffic((nx, ny, nz, nt))
    value = prof["voxel_updates"] / total

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline sample ...")
        cpu = cpu_baseline(lib, shape, u0, args.cpu_iters, float(np.mean(sweeps)), nint)
        log(f"[bench] CPU baseline {cpu['value']:.3e} voxel-updates/s on {cpu['cores']} threads "
            f"({cpu['physical_cores_available']} physical cores available)")
    lib.deallocate()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"plain FP64 subjective-surface time steps, {nx}x{ny}x{nz}x{nt} "
                         f"volume, red-black SOR omega 1.8 to relative residual 1e-6"),
            "volume": [nx, ny, nz, nt],
            "time_steps": args.steps,
            "solver": "colour-layout red-black SOR, operand blocks, residual fused into the passes",
            "sweep_variant": args.variant if args.variant is not None else 35,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_sor_sweep<1> (red pass + fused residual)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": nred * BYTES_PER_UPDATE,
            "avg_launch_ms": red1_avg_s * 1e3,
            "launches": int(prof["red1_launches"]),
            "copy_gbps_same_run": copy_gbps,
            "frac_of_copy": achieved / copy_gbps if copy_gbps else None,
        },
        "hbm_gbps_all_sweeps": sweep_bytes / (sweep_ms / 1e3) / 1e9 if sweep_ms > 0 else None,
        # average HIP-event time per launch of each FP64 pass kind (launches
        # queued after a solve's stop are not counted)
        "pass_avg_ms": {k: prof[k + "_ms"] / prof[k + "_launches"] if prof[k + "_launches"] else None
                        for k in ("red0", "red1", "black")},
        "time_to_tolerance_s": total,
        "sweeps_per_step": sweeps,
        "setup_ms_per_step": prof["setup_ms"] / max(prof["setup_launches"], 1),
        "allocate_s": alloc_s,
        "edge_indicator_s_pcie_inclusive": edge_s,
        "edge_indicator_ms_resident": edge_dev_s * 1e3,
        "edge_indicator_presmoothing_sweeps": edge_iters,
        "mirror_mode": mirror,
        "mask_fraction": mask_frac,
        "mask_voxels": int(mask64.sum()),
        "fp32_refined": mixed,
        "fp32_refined_extrapolated_start": mixed_warm,
        "fp32_refined_cubic_start": mixed_cubic,
        "fp32_refined_learned_start": learned,
        "controls_non_parity": controls,
        "fused_sweeps": fused,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_batch_bench(args, lib):
    """Batched workload: `--batch-volumes` volumes of uneven size (generator
    preset 3: x, y in {48..192 step 16}, z in {32..128 step 16}, t in
    {6..16}, 2-6 objects each), `--batch-steps` time steps, relative residual
    1e-6, launch tiles from the autotuner.  Times the batch in FP64 (one
    launch per pass with eager and graph-replayed iterations, and the
    concurrent form with `--batch-streams` volume groups on their own
    streams, and every step after the first as one captured graph) and
    FP32-refined (start u^n and cubic start, with graph-replayed iterations
    and with captured steps), and the same FP64
    steps one volume at a time on the single-volume path; `value` is the
    fastest FP64 batch form.  Prints one JSON line."""
    nv, steps = args.batch_volumes, args.batch_steps
    specs = [lib.gen_spec(3, 300 + v) for v in range(nv)]
    dims = [(sp.nx, sp.ny, sp.nz, sp.nt) for sp in specs]
    nint = [nx * ny * nz * nt for nx, ny, nz, nt in dims]
    log(f"[batch] {nv} volumes, {sum(nint) / 1e6:.1f} M interior voxels, on {lib.device_info()}")
    vols = []
    for sp, d in zip(specs, dims):
        shape = (d[3] + 2, d[2] + 2, d[1] + 2, d[0] + 2)
        img, u0 = np.zeros(shape), np.zeros(shape)
        lib.generate(sp, img, u0)
        vols.append((img, u0))
    params(lib)

    def batch_run(precision, warm, graph, depth=2, streams=0):
        lib.batch_create(dims)
        try:
            lib.batch_configure(precision, warm, graph, depth)
            lib.batch_streams(streams)
            t = time.perf_counter()
            for v, (img, u0) in enumerate(vols):
                lib.batch_set_volume(v, img, u0, np.ones_like(img))
            edge_s = time.perf_counter() - t
            tiles, _ = lib.batch_autotune(3)
            lib.synchronize()
            sweeps, dev_ms = [], 0.0
            t0 = time.perf_counter()
            for s in range(steps):
                lib.set(numberOfTimeStep=s + 1)
                dev_ms += lib.batch_step()
                sweeps.append([lib.batch_stats(v)[0] for v in range(nv)])
            t1 = time.perf_counter()
            us = [lib.batch_u(v)[1:-1, 1:-1, 1:-1, 1:-1].copy() for v in range(nv)]
            git, grd, gst, gre, ghm = lib.batch_graph_info()
        finally:
            lib.batch_destroy()
        total = t1 - t0
        upd = sum(sum(nint[v] * sw[v] for v in range(nv)) for sw in sweeps)
        tl = {}
        for t_ in tiles:
            tl[str(t_)] = tl.get(str(t_), 0) + 1
        return {"time_s": total, "device_ms": dev_ms, "iterations": int(np.sum(sweeps)),
                "voxel_updates_per_s": upd / total, "edge_indicator_s_pcie_inclusive": edge_s,
                "tiles": tl, "max_sweeps_per_step": [int(max(sw)) for sw in sweeps],
                "graph": ({"captured_steps": gst, "redone_eagerly": gre,
                           "last_plan_iterations": git, "last_plan_rounds": grd,
                           "host_capture_ms": ghm}
                          if graph == 2 else None)}, us

    log("[batch] FP64, eager iterations ...")
    fp64, u64 = batch_run(0, 0, 0)
    log(f"[batch] FP64: {fp64['time_s']:.3f} s, {fp64['voxel_updates_per_s']:.3e} voxel-updates/s")
    log("[batch] FP64, graph-replayed iterations ...")
    fp64g, _ = batch_run(0, 0, 1)
    log(f"[batch] FP64 graph: {fp64g['time_s']:.3f} s")
    log("[batch] FP64, every step after the first as one captured graph ...")
    fp64s, us_ = batch_run(0, 0, 2)
    fp64s["bitwise_equal_to_eager_form"] = all(
        np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(u64, us_))
    log(f"[batch] FP64 step graph: {fp64s['time_s']:.3f} s, {fp64s['graph']}, "
        f"bitwise equal {fp64s['bitwise_equal_to_eager_form']}")
    conc = {}
    for k in [int(x) for x in args.batch_streams.split(",") if x.strip()]:
        log(f"[batch] FP64, concurrent form, {k} groups ...")
        r, uc = batch_run(0, 0, 0, streams=k)
        r["groups"] = k
        r["bitwise_equal_to_one_launch_form"] = all(
            np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(u64, uc))
        conc[f"fp64_batch_streams_{k}"] = r
        log(f"[batch] FP64 {k} groups: {r['time_s']:.3f} s, "
            f"bitwise equal {r['bitwise_equal_to_one_launch_form']}")
    forms = {"fp64_batch": fp64, "fp64_batch_graph": fp64g, "fp64_batch_step_graph": fp64s,
             **conc}
    best = min(forms, key=lambda f: forms[f]["time_s"])
    ref = {}
    for name, (prec, warm, graph) in (("fp32_refined", (1, 0, 1)),
                                      ("fp32_refined_cubic_start", (1, 4, 1)),
                                      ("fp32_refined_step_graph", (1, 0, 2)),
                                      ("fp32_refined_cubic_start_step_graph", (1, 4, 2))):
        log(f"[batch] {name} ...")
        r, um = batch_run(prec, warm, graph)
        ious, dus = [], []
        for a, b in zip(u64, um):
            ma = a >= 0.5 * (a.max() + a.min())
            mb = b >= 0.5 * (b.max() + b.min())
            un = np.logical_or(ma, mb).sum()
            ious.append(float(np.logical_and(ma, mb).sum() / un) if un else 1.0)
            dus.append(float(np.abs(a - b).max()))
        r["speedup_vs_fp64_batch"] = fp64["time_s"] / r["time_s"]
        r["min_mask_iou_vs_fp64"] = min(ious)
        r["max_abs_du_vs_fp64"] = max(dus)
        ref[name] = r
        log(f"[batch] {name}: {r['time_s']:.3f} s ({r['speedup_vs_fp64_batch']:.2f}x), "
            f"min IoU {min(ious):.5f}")
    seq = None
    if not args.no_sequential:
        log("[batch] FP64 one volume at a time ...")
        tot, upd, ok = 0.0, 0, True
        for v, (img, u0) in enumerate(vols):
            lib.allocate(*dims[v])
            params(lib)
            lib.arr("i0")[...] = img
            lib.arr("u")[...] = 0.0
            lib.arr("Mmask")[...] = 1.0
            lib.gf("i0", "u")
            lib.set_resident(True)
            lib.arr("u")[...] = u0
            lib.upload("u")
            lib.upload("Mmask")
            lib.autotune(3)
            lib.synchronize()
            t0 = time.perf_counter()
            for s in range(steps):
                lib.set(numberOfTimeStep=s + 1)
                z, _ = lib.seg_time_step()
                upd += nint[v] * z
            lib.synchronize()
            tot += time.perf_counter() - t0
            lib.download("u")
            ok = ok and bool(np.array_equal(lib.arr("u")[1:-1, 1:-1, 1:-1, 1:-1], u64[v]))
            lib.set_resident(False)
            lib.deallocate()
        seq = {"time_s": tot, "voxel_updates_per_s": upd / tot,
               "batch_speedup": tot / fp64["time_s"],
               "batch_speedup_by_form": {f: tot / forms[f]["time_s"] for f in forms},
               "bitwise_equal_to_batch": ok}
        log(f"[batch] sequential: {tot:.3f} s; batch {tot / fp64['time_s']:.2f}x, "
            f"{best} {tot / forms[best]['time_s']:.2f}x; bitwise equal {ok}")
    out = {
        "metric": "SOR voxel-updates/s, batched volumes",
        "value": forms[best]["voxel_updates_per_s"],
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"{nv} volumes of uneven size (x, y in 48..192, z in 32..128, "
                                f"t in 6..16), {steps} time steps each, relative residual 1e-6"),
                   "volumes": [list(d) for d in dims], "interior_voxels": int(sum(nint)),
                   "form": best},
        "fp64_batch": fp64,
        "fp64_batch_graph": fp64g,
        "fp64_batch_step_graph": fp64s,
        **conc,
        "fp32_refined_batch": ref["fp32_refined"],
        "fp32_refined_cubic_start_batch": ref["fp32_refined_cubic_start"],
        "fp32_refined_step_graph_batch": ref["fp32_refined_step_graph"],
        "fp32_refined_cubic_start_step_graph_batch": ref["fp32_refined_cubic_start_step_graph"],
        "fp64_sequential": seq,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
