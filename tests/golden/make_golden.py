"""Writes the golden fixtures of this directory by running the reference's own
C programs (segmentation4D.c, mcf_4D.c), compiled from /root/reference into
oracle/_ref/ by `make -C oracle ref`, on small seeded inputs.

Each .npz holds inputs and the reference's outputs (float64, bitwise):
  seg_h1.npz   h = 1 (power of two), theta = rho = 0.5, balls in the rescale
               epilogue, a partial Mmask, 3 tolerance-driven time steps
  seg_h001.npz the reference main()'s own parameters (h = tau = 0.01,
               eps2 = 1, K = 1, sigma = h^2/8, 5 sweeps cap, 10 presmoothing
               sweeps cap), 2 time steps
  mcf.npz      the MCF program's parameters (h = 2.5/n, tau = eps2 = h^2),
               3 time steps, L2 error per step
  init.npz     local threshold, seed function, rescale and the ball / max /
               mass helpers on a volume with bright blobs at the centres
  main_flow.npz the reference main()'s flow and parameters end to end
  vtk.npz      _readData_ of the files in vtk/ and _writeData_ output bytes
  mcf_eoc.npz  the MCF program's L2 error at T = 0.0625 for n = 10, 20, 40

Run:  make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import os
import sys
from ctypes import c_double, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from harness import DP, RefMcf, RefSeg  # noqa: E402

FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")


def zero_ghosts(a):
    a = a.copy()
    a[0] = a[-1] = 0
    a[:, 0] = a[:, -1] = 0
    a[:, :, 0] = a[:, :, -1] = 0
    a[..., 0] = a[..., -1] = 0
    return a


def seg_case(name, shape, seed, dparams, iparams, steps, counts=None, centres=None,
             radius=0.0, partial_mask=False):
    rng = np.random.default_rng(seed)
    i0 = zero_ghosts(np.round(rng.uniform(0, 255, size=shape)) / 255.0)
    thr = zero_ghosts(np.round(rng.uniform(0, 255, size=shape)) / 255.0)
    u0 = zero_ghosts(rng.uniform(0.05, 1.0, size=shape))
    mask = np.ones(shape)
    if partial_mask:
        mask = (rng.uniform(size=shape) > 0.1).astype(np.float64)
    L_ = shape[0]
    counts = counts or [0] * L_
    ref = RefSeg()
    ref.allocate(shape, counts)
    if centres:
        ref.set_centres(centres)
    ref.setd(ballRadius=radius, **dparams)
    ref.seti(**iparams)
    ref.arr("i0")[...] = i0
    ref.arr("u")[...] = thr
    ref.arr("Mmask")[...] = mask
    gf = ref.fn("_gf_inputImage_plus_thresholdedImage_", argtypes=[DP, DP])
    assert gf(ref.ptr("i0"), ref.ptr("u"))
    out = dict(i0=i0, thr=thr, u0=u0, mask=mask, counts=np.array(counts, np.int32),
               centres=np.array([c for lst in (centres or []) for c in lst] or
                                np.zeros((0, 3))).reshape(-1, 3),
               radius=np.float64(radius),
               i0_s=ref.arr("i0").copy(), thr_s=ref.arr("u").copy(),
               prev_gf=ref.arr("prev_u").copy())
    for f in FACES:
        out["g_" + f] = ref.arr("g_" + f).copy()
    ref.arr("u")[...] = u0
    step = ref.fn("_segTimeStep_")
    l2 = ref.fn("_l2norm_", restype=c_double, argtypes=[DP, DP])
    norms = []
    for st in range(steps):
        c_int.in_dll(ref.L, "numberOfTimeStep").value = st + 1
        assert step()
        out[f"u{st + 1}"] = ref.arr("u").copy()
        norms.append(l2(ref.ptr("prev_u"), ref.ptr("u")))
        if st == 0:
            for f in FACES:
                out["coef1_" + f] = ref.arr(f).copy()
    out["prev_last"] = ref.arr("prev_u").copy()
    out["l2"] = np.array(norms)
    out["steps"] = np.int32(steps)
    for k, v in dparams.items():
        out["p_" + k] = np.float64(v)
    for k, v in iparams.items():
        out["p_" + k] = np.int64(v)
    ref.deallocate()
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name)


def init_case(name, shape, seed):
    """The reference's initialisation chain (segmentation4D.c:1386-1395):
    _locallyThreshold4Ddata_(i0, u), _generate4DsegmFunct_(u),
    _rescale4DdataToZeroOne_(u); centres at least ballRadius + 2 voxels from
    the edges so that the reference's unchecked ball loops stay in bounds."""
    rng = np.random.default_rng(seed)
    L_, K_, J_, I_ = shape
    img = zero_ghosts(np.round(rng.uniform(0, 120, size=shape)) / 255.0)
    counts = [0] + [int(rng.integers(1, 4)) for _ in range(L_ - 2)] + [0]
    R = 3.0
    centres = [[]]
    for l in range(1, L_ - 1):
        lst = []
        for c in range(counts[l]):
            cx = float(rng.integers(int(R) + 2, I_ - int(R) - 3))
            cy = float(rng.integers(int(R) + 2, J_ - int(R) - 3))
            cz = float(rng.integers(int(R) + 2, K_ - int(R) - 3))
            lst.append((cx, cy, cz))
            # a bright blob around the centre so that thresholding selects it
            kk, jj, ii = np.meshgrid(np.arange(K_), np.arange(J_), np.arange(I_), indexing="ij")
            blob = ((ii - cx) ** 2 + (jj - cy) ** 2 + (kk - cz) ** 2) <= 4.0
            img[l][blob] = np.round(rng.uniform(180, 255, size=blob.sum())) / 255.0
        centres.append(lst)
    centres.append([])
    ref = RefSeg()
    ref.allocate(shape, counts)
    ref.set_centres(centres)
    ref.setd(ballRadius=R, alpha=0.85, beta=0.15)
    ref.arr("i0")[...] = img
    thr = ref.fn("_locallyThreshold4Ddata_", argtypes=[DP, DP])
    gen = ref.fn("_generate4DsegmFunct_", argtypes=[DP])
    resc = ref.fn("_rescale4DdataToZeroOne_", argtypes=[DP])
    gmax = ref.fn("_get4DMax_", restype=c_double, argtypes=[DP])
    bmax = ref.fn("_get3DMaxWithinR_", restype=c_double,
                  argtypes=[DP, c_int, c_double, c_double, c_double, c_double])
    bmin = ref.fn("_get3DMinWithinR_", restype=c_double,
                  argtypes=[DP, c_int, c_double, c_double, c_double, c_double])
    mass = ref.fn("_mass4D_", restype=c_double, argtypes=[DP])
    out = dict(i0=img, counts=np.array(counts, np.int32),
               centres=np.array([c for lst in centres for c in lst]).reshape(-1, 3),
               radius=np.float64(R), alpha=np.float64(0.85), beta=np.float64(0.15))
    l1, c1 = 1, centres[1][0]
    out["gmax"] = np.float64(gmax(ref.ptr("i0")))
    out["bmax"] = np.float64(bmax(ref.ptr("i0"), l1, *c1, R))
    out["bmin"] = np.float64(bmin(ref.ptr("i0"), l1, *c1, R))
    ref.setd(h=0.5)
    out["mass"] = np.float64(mass(ref.ptr("i0")))
    assert thr(ref.ptr("i0"), ref.ptr("u"))
    out["thr"] = ref.arr("u").copy()
    assert gen(ref.ptr("u"))
    out["seed"] = ref.arr("u").copy()
    out["mask"] = ref.arr("Mmask").copy()
    assert resc(ref.ptr("u"))
    out["rescaled"] = ref.arr("u").copy()
    ref.deallocate()
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name)


def main_case(name, shape, seed, steps):
    """The reference main()'s whole flow (segmentation4D.c:1383-1410) with its
    own parameters (segmentation4D.c:1348-1366) except a smaller ballRadius:
    local threshold, edge indicator, seed function, rescale, time steps."""
    rng = np.random.default_rng(seed)
    L_, K_, J_, I_ = shape
    R = 3.0
    img = zero_ghosts(np.round(rng.uniform(0, 90, size=shape)) / 255.0)
    counts = [0] + [int(rng.integers(1, 3)) for _ in range(L_ - 2)] + [0]
    centres = [[]]
    kk, jj, ii = np.meshgrid(np.arange(K_), np.arange(J_), np.arange(I_), indexing="ij")
    for l in range(1, L_ - 1):
        lst = []
        for _ in range(counts[l]):
            c = tuple(float(rng.integers(5, n - 6)) for n in (I_, J_, K_))
            lst.append(c)
            blob = ((ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2) <= 5.0
            img[l][blob] = np.round(rng.uniform(170, 255, size=blob.sum())) / 255.0
        centres.append(lst)
    centres.append([])
    hh = 0.01
    dparams = dict(h=hh, tau=0.01, eps2=1.0, K=1.0, omega_c=1.8, omega_sigma=1.3,
                   sigma=hh * hh / 8, gsTol=1e-12, theta_img=0.5, rho_thr=0.5,
                   maxNoGSIteration_sigma=10.0, alpha=0.85, beta=0.15, ballRadius=R)
    ref = RefSeg()
    ref.allocate(shape, counts)
    ref.set_centres(centres)
    ref.setd(**dparams)
    ref.seti(maxNoGSIteration=5)
    ref.arr("i0")[...] = img
    P = [ref.ptr("i0"), ref.ptr("u")]
    assert ref.fn("_locallyThreshold4Ddata_", argtypes=[DP, DP])(*P)
    assert ref.fn("_gf_inputImage_plus_thresholdedImage_", argtypes=[DP, DP])(*P)
    out = dict(i0=img, counts=np.array(counts, np.int32),
               centres=np.array([c for lst in centres for c in lst]).reshape(-1, 3),
               thr_s=ref.arr("u").copy(), i0_s=ref.arr("i0").copy())
    for f in FACES:
        out["g_" + f] = ref.arr("g_" + f).copy()
    assert ref.fn("_generate4DsegmFunct_", argtypes=[DP])(ref.ptr("u"))
    assert ref.fn("_rescale4DdataToZeroOne_", argtypes=[DP])(ref.ptr("u"))
    out["u0"] = ref.arr("u").copy()
    out["mask"] = ref.arr("Mmask").copy()
    step = ref.fn("_segTimeStep_")
    l2 = ref.fn("_l2norm_", restype=c_double, argtypes=[DP, DP])
    norms = []
    for st in range(steps):
        c_int.in_dll(ref.L, "numberOfTimeStep").value = st + 1
        assert step()
        out[f"u{st + 1}"] = ref.arr("u").copy()
        norms.append(l2(ref.ptr("prev_u"), ref.ptr("u")))
    out["l2"] = np.array(norms)
    out["steps"] = np.int32(steps)
    for k, v in dparams.items():
        out["p_" + k] = np.float64(v)
    out["p_maxNoGSIteration"] = np.int64(5)
    ref.deallocate()
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name)


VTK_HEADER = ("# vtk DataFile Version 3.0\nfile in binary format\nBINARY\n"
              "DATASET STRUCTURED_POINTS\nDIMENSIONS {nx} {ny} {nz}\nORIGIN 0 0 0\n"
              "SPACING 0.5 0.5 0.5\nPOINT_DATA {n}\nSCALARS scalars unsigned_char\n"
              "LOOKUP_TABLE default\n")


def vtk_case(name, dims, seed):
    """Input files in the reference's formats under vtk/ (image frames and
    centre files, one frame truncated by 5 bytes, centre files with and
    without extra comment lines), what the reference's _readData_ makes of
    them, and the bytes its _writeData_ writes for values inside and outside
    [0, 1]."""
    rng = np.random.default_rng(seed)
    nx, ny, nz, nt = dims
    d = os.path.join(HERE, "vtk")
    os.makedirs(d, exist_ok=True)
    counts = [0] + [2, 3, 1][:nt] + [0]
    sp = 0.5
    centres_in = []
    for l in range(1, nt + 1):
        data = rng.integers(0, 256, size=nx * ny * nz, dtype=np.uint8)
        data[:3] = (0, 255, 128)
        if l == nt:
            data = data[:-5]
        with open(os.path.join(d, f"movingBalls_457_at_time_{l:03d}.vtk"), "wb") as f:
            f.write(VTK_HEADER.format(nx=nx, ny=ny, nz=nz, n=nx * ny * nz).encode())
            f.write(data.tobytes())
        pts = [tuple(float(np.round(rng.uniform(0, n * sp), 3)) for n in (nx, ny, nz))
               for _ in range(counts[l])]
        centres_in.append(pts)
        lines = ["# vtk DataFile Version 3.0"]
        if l == 2:
            lines += ["# an extra comment line"]
        lines += ["centres at one time step", "ASCII", "DATASET POLYDATA",
                  f"POINTS {len(pts)} float"]
        if l == 1:
            lines += [" ".join(f"{v}" for p in pts for v in p)]
        else:
            lines += [f"{p[0]} {p[1]} {p[2]}" for p in pts]
        with open(os.path.join(d, f"movingBalls_457_centers_at_timestep_{l:03d}.vtk"), "w") as f:
            f.write("\n".join(lines) + "\n")
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    ref = RefSeg()
    ref.allocate(shape, counts)
    ref.setd(spacing_x=sp, spacing_y=sp, spacing_z=sp, origin_x=0.25, origin_y=0.0,
             origin_z=-1.5)
    read = ref.fn("_readData_", restype=None, argtypes=[DP, ctypes.c_char_p])
    read(ref.ptr("i0"), d.encode())
    out = dict(i0=ref.arr("i0").copy(), counts=np.array(counts, np.int32))
    cx, cy, cz = (ctypes.POINTER(ctypes.POINTER(c_double)).in_dll(ref.L, n)
                  for n in ("center_x", "center_y", "center_z"))
    out["centres"] = np.array([(cx[l][c], cy[l][c], cz[l][c]) for l in range(nt + 2)
                               for c in range(counts[l])]).reshape(-1, 3)
    vals = rng.uniform(-0.3, 1.3, size=shape)
    vals.flat[:6] = (0.5 / 255, 1.5 / 255, 1.0, 0.0, -1.0 / 255, 256.5 / 255)
    out["w_values"] = vals
    ref.arr("u")[...] = vals
    write = ref.fn("_writeData_", argtypes=[DP, ctypes.c_char_p])
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        base = os.path.join(td, "res")
        assert write(ref.ptr("u"), base.encode())
        for l in range(1, nt + 1):
            with open(f"{base}_at_time_{l:03d}.vtk", "rb") as f:
                out[f"w_file{l}"] = np.frombuffer(f.read(), dtype=np.uint8)
    out["spacing"] = np.float64(sp)
    out["origin"] = np.array([0.25, 0.0, -1.5])
    ref.deallocate()
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name, "and vtk/")


def mcf_case(name, m, steps):
    ref = RefMcf()
    ref.seti(iMax=m, jMax=m, kMax=m, lMax=m, maxNoGSIteration=1000)
    hh = 2.5 / (m - 2)
    ref.setd(h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8, mcfTol=1e-12)
    assert ref.fn("_allocateMem_mcf")()
    exact = ref.fn("_exactSolution_mcf", restype=None, argtypes=[c_double])
    copy = ref.fn("_copy4DdataFromSrcToDest_mcf", restype=None, argtypes=[DP, DP])
    step = ref.fn("_mcfTimeStep")
    l2 = ref.fn("_l2norm_mcf", restype=c_double, argtypes=[DP, DP])
    exact(0.0)
    copy(ref.ptr("exact_u"), ref.ptr("u"))
    out = dict(u0=ref.arr("u").copy(), h=np.float64(hh), m=np.int32(m))
    errs = []
    for st in range(1, steps + 1):
        c_int.in_dll(ref.L, "numberOfTimeStep").value = st
        assert step()
        out[f"u{st}"] = ref.arr("u").copy()
        exact(st * hh * hh)
        errs.append(l2(ref.ptr("exact_u"), ref.ptr("u")))
    out["exact_last"] = ref.arr("exact_u").copy()
    out["err"] = np.array(errs)
    out["steps"] = np.int32(steps)
    ref.fn("_deallocateMem_mcf")()
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name)


def mcf_eoc_case(name, ns, T):
    """The MCF program's loop (mcf_4D.c:473-524) at n = ns to the final time
    T: the L2 error after the last step and the SOR iterations per level."""
    errs, sweeps = [], []
    for m in ns:
        hh = 2.5 / m
        steps = int(round(T / (hh * hh)))
        ref = RefMcf()
        ref.seti(iMax=m + 2, jMax=m + 2, kMax=m + 2, lMax=m + 2, maxNoGSIteration=1000)
        ref.setd(h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8, mcfTol=1e-12)
        assert ref.fn("_allocateMem_mcf")()
        exact = ref.fn("_exactSolution_mcf", restype=None, argtypes=[c_double])
        exact(0.0)
        ref.arr("u")[...] = ref.arr("exact_u")
        step = ref.fn("_mcfTimeStep")
        for st in range(1, steps + 1):
            c_int.in_dll(ref.L, "numberOfTimeStep").value = st
            assert step()
        exact(steps * hh * hh)
        l2 = ref.fn("_l2norm_mcf", restype=c_double, argtypes=[DP, DP])
        errs.append(l2(ref.ptr("exact_u"), ref.ptr("u")))
        ref.fn("_deallocateMem_mcf")()
        print(f"  n = {m}: {steps} steps, error {errs[-1]:.6e}", flush=True)
    np.savez_compressed(os.path.join(HERE, name), n=np.array(ns, np.int32),
                        T=np.float64(T), err=np.array(errs))
    print("wrote", name)


def main():
    shape = (5, 6, 7, 8)  # padded (lMax, kMax, jMax, iMax)
    counts = [0, 1, 2, 1, 0]
    centres = [[], [(3.0, 3.0, 2.0)], [(2.0, 4.0, 3.0), (5.0, 3.0, 3.0)], [(4.0, 3.0, 2.5)], []]
    seg_case("seg_h1.npz", shape, 11,
             dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3,
                  sigma=0.5, gsTol=1e-14, theta_img=0.5, rho_thr=0.5,
                  maxNoGSIteration_sigma=50.0),
             dict(maxNoGSIteration=500), steps=3, counts=counts, centres=centres,
             radius=1.6, partial_mask=True)
    h = 0.01
    seg_case("seg_h001.npz", (5, 5, 6, 7), 12,
             dict(h=h, tau=0.01, eps2=1.0, K=1.0, omega_c=1.8, omega_sigma=1.3,
                  sigma=h * h / 8, gsTol=1e-12, theta_img=0.5, rho_thr=0.5,
                  maxNoGSIteration_sigma=10.0),
             dict(maxNoGSIteration=5), steps=2)
    mcf_case("mcf.npz", 7, 3)
    init_case("init.npz", (5, 16, 18, 20), 13)
    main_case("main_flow.npz", (6, 18, 20, 22), 17, 3)
    vtk_case("vtk.npz", (6, 5, 4, 3), 19)
    mcf_eoc_case("mcf_eoc.npz", (10, 20, 40), 0.0625)


if __name__ == "__main__":
    main()
