"""MCF verification at growing resolution (mcf_4D.c:344-524): the analytic
solution u = (|x|^2 - 1)/6 + t of 4D level-set mean-curvature flow pins the
coefficient and SOR kernels at sizes the CPU oracle cannot reach in a test.

n = 10, 20, 40 reproduce the compiled reference's final-time L2 errors
(tests/golden/mcf_eoc.npz; equal up to the order of the device's stopping
sum, rel 1e-9); n = 80 (64 steps on 82^4 points) must continue the
convergence: experimental order log2(e40 / e80) >= 1.85, after 1.35 and
1.82 on the coarser pairs."""
import os

import numpy as np
import pytest

from harness import GOLDEN

pytestmark = pytest.mark.gpu


def _final_error(lib, n, T):
    hh = 2.5 / n
    steps = int(round(T / (hh * hh)))
    lib.set(maxNoGSIteration=1000)
    lib.allocate(n)
    lib.set(h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8)
    lib.set_resident(True)
    sweeps = 0
    try:
        lib.exact(0.0)
        lib.copy("exact_u", "u")
        for st in range(1, steps + 1):
            lib.set(numberOfTimeStep=st)
            z, _ = lib.time_step()
            sweeps += z
        lib.exact(steps * hh * hh)
        return lib.l2norm("exact_u", "u"), steps, sweeps
    finally:
        lib.set_resident(False)
        lib.deallocate()


def test_mcf_convergence(mcflib):
    fx = np.load(os.path.join(GOLDEN, "mcf_eoc.npz"), allow_pickle=False)
    T = float(fx["T"])
    errs = []
    for n, ref in zip(fx["n"], fx["err"]):
        print(f"[gpu] mcf n = {n}", flush=True)
        e, steps, sweeps = _final_error(mcflib, int(n), T)
        print(f"  {steps} steps, {sweeps} sweeps, error {e:.6e} (reference {ref:.6e})",
              flush=True)
        assert e == pytest.approx(ref, rel=1e-9)
        errs.append(e)
    print("[gpu] mcf n = 80", flush=True)
    e80, steps, sweeps = _final_error(mcflib, 80, T)
    eoc = np.log2(errs[-1] / e80)
    print(f"  {steps} steps, {sweeps} sweeps, error {e80:.6e}, EOC {eoc:.3f}", flush=True)
    assert eoc >= 1.85
