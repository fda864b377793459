"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the CPU
oracle (every threaded branch, the ball loops with centres on and beyond
the volume's edges), the file formats and the synthetic generator, built
with gcc -fsanitize=address,undefined by tests/sanitize/Makefile.  The
reference indexes its ball loops unchecked (segmentation4D.c:907-930,
991-1015, 1033-1090, 1119-1145); the restatement and the library clamp
them, and these runs show no access leaves its array.

The `io` run also checks what the sanitized readers produce against the
golden fixture the compiled reference wrote (tests/golden/vtk.npz)."""
import os
import subprocess

import numpy as np
import pytest

from harness import GOLDEN, ROOT

SAN = os.path.join(ROOT, "tests", "sanitize")
DRIVER = os.path.join(SAN, "_build", "san_driver")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="3")


@pytest.fixture(scope="module")
def driver():
    r = subprocess.run(["make", "-C", SAN], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(f"sanitizer build failed:\n{r.stderr[-2000:]}")
    return DRIVER


def run(args, **kw):
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, env=ENV, timeout=300,
                       **kw)
    clean = "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    return r, clean


def test_sanitizer_is_active(driver):
    r, clean = run(["selfcheck"])
    assert r.returncode != 0 and "AddressSanitizer" in r.stderr


def test_oracle_under_sanitizers(driver):
    r, clean = run(["oracle"])
    assert r.returncode == 0 and clean, r.stderr[-3000:]
    lines = [x for x in r.stdout.splitlines() if x.startswith("oracle variant")]
    assert len(lines) == 4
    # serial, threaded with serial sums, OpenMP reduction, parallel terms:
    # everything but the residual order is identical, so the summary is too
    assert len({x.split(":", 1)[1] for x in lines}) == 1


def test_generator_under_sanitizers(driver):
    r, clean = run(["gen"])
    assert r.returncode == 0 and clean, r.stderr[-3000:]
    assert r.stdout.count("gen preset") == 5


def test_file_formats_under_sanitizers(driver, tmp_path):
    fx = np.load(os.path.join(GOLDEN, "vtk.npz"), allow_pickle=False)
    L, K, J, I = fx["i0"].shape
    counts = [int(c) for c in fx["counts"]]
    out = tmp_path / "read.raw"
    base = tmp_path / "w"
    args = ["io", os.path.join(GOLDEN, "vtk"), str(I), str(J), str(K), str(L)]
    args += [str(c) for c in counts] + [repr(float(fx["spacing"])), str(out), str(base)]
    r, clean = run(args)
    assert r.returncode == 0 and clean, r.stderr[-3000:]
    raw = np.fromfile(out, dtype=np.float64)
    n = I * J * K * L
    assert np.array_equal(raw[:n].reshape(fx["i0"].shape), fx["i0"])
    assert np.array_equal(raw[n:].reshape(-1, 3), fx["centres"])
    for l in range(1, L - 1):
        assert (tmp_path / f"w_at_time_{l:03d}.vtk").stat().st_size > (I - 2) * (J - 2) * (K - 2)
        assert (tmp_path / f"w_mask_at_time_{l:03d}.vtk").exists()
