"""Training pipeline of the learned warm start (host plumbing over the C ABI;
every computation runs in libseg4d_amd.so on the GPU).

Runs of the plain solver on synthetic volumes of the 256x256x128x16
workload's style (six ellipsoids with semi-axes drawn in [12, 40] voxels,
time-periodic radii, wedge gaps, noise sigma 25) fill the network's pool
with the solution after every time step; the network is then fitted with
Adam on patches of those runs and its weights saved.

    python warmstart.py [--runs 6] [--steps 30] [--dims 160 160 96 16]
                        [--iters 2500] [--batch 8] [--layers 1] [--base 3]
                        [--out weights.bin]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(PKG_DIR, "weights", "warmstart.bin")


def solver_params(lib):
    """The workload's scheme: eps = 1e-3, tau = 0.5, h = 1, K = 1000,
    presmoothing sigma = 1 voxel, SOR omega 1.8 to relative residual 1e-6."""
    lib.set(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
            theta_img=1.0, rho_thr=0.0, gsTol=0.0, maxNoGSIteration=100000,
            maxNoGSIteration_sigma=100000.0)
    lib.set_relative_tolerance(1e-6)


def record_runs(lib, runs, steps, dims, seed0=1000, log=print):
    """Appends `runs` solver runs of `steps` time steps on volumes of
    interior size dims = (nx, ny, nz, nt) to the network's pool."""
    nx, ny, nz, nt = dims
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    for r in range(runs):
        t0 = time.perf_counter()
        spec = lib.gen_spec(1, seed0 + r)
        spec.nx, spec.ny, spec.nz, spec.nt = nx, ny, nz, nt
        lib.allocate(nx, ny, nz, nt)
        try:
            solver_params(lib)
            u0 = np.zeros(shape)
            lib.generate(spec, lib.arr("i0"), u0)
            lib.arr("u")[...] = 0.0
            lib.arr("Mmask")[...] = 1.0
            lib.gf("i0", "u")
            lib.set_resident(True)
            lib.arr("u")[...] = u0
            lib.upload("u")
            lib.upload("Mmask")
            lib.set_warm_start(4)  # the plain solver's system, fewer sweeps
            lib.net_record(new_run=True)
            for s in range(steps):
                lib.set(numberOfTimeStep=s + 1)
                lib.seg_time_step()
                lib.net_record()
        finally:
            lib.set_warm_start(0)
            lib.set_resident(False)
            lib.set_relative_tolerance(0.0)
            lib.deallocate()
        log(f"[warmstart] run {r + 1}/{runs}: {steps} steps in {time.perf_counter() - t0:.1f} s")


def train(lib, runs=6, steps=30, dims=(160, 160, 96, 16), iters=2500, batch=8, lr=1e-3,
          seed=1, out=DEFAULT_WEIGHTS, layers=1, base=3, log=print):
    """Records the runs, trains a network of `layers` layers correcting the
    extrapolation of order `base`, saves; returns a summary dict."""
    lib.net_init(seed, layers=layers, base=base)
    t0 = time.perf_counter()
    record_runs(lib, runs, steps, dims, log=log)
    t1 = time.perf_counter()
    loss = lib.net_train(iters, batch=batch, lr=lr, seed=seed, log_every=max(1, iters // 10))
    t2 = time.perf_counter()
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        lib.net_save(out)
    s_d, s_t = lib.net_scales()
    return {
        "runs": runs, "steps_per_run": steps, "run_volume": list(dims),
        "pairs": int(lib.net_pool_pairs()), "iterations": iters, "batch": batch,
        "patches": iters * batch, "lr": lr, "final_loss": loss,
        "layers": layers, "base": base,
        "record_s": t1 - t0, "train_s": t2 - t1, "s_d": s_d, "s_t": s_t,
        "weights": out,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=6)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--dims", type=int, nargs=4, default=[160, 160, 96, 16])
    ap.add_argument("--iters", type=int, default=2500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--layers", type=int, default=1)
    ap.add_argument("--base", type=int, default=3, choices=(0, 1, 3))
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args()
    sys.path.insert(0, PKG_DIR)
    import seg4d

    lib = seg4d.SegLib()
    lib.set_verbose(False)
    res = train(lib, a.runs, a.steps, tuple(a.dims), a.iters, a.batch, a.lr, a.seed, a.out,
                layers=a.layers, base=a.base, log=lambda m: print(m, flush=True))
    print(res, flush=True)


if __name__ == "__main__":
    main()
