"""Python host mirror of the C ABI in include/seg4d.h and include/mcf4d.h.

The product is the pair of C-ABI shared libraries built from csrc/
(lib/libseg4d_amd.so, lib/libmcf4d_amd.so); this module only binds them with
ctypes so that tests, the smoke check and the benchmark can drive them the
way the reference's C drivers do (segmentation4D.c:1326-1424,
mcf_4D.c:473-524): set the globals, allocate, fill the host arrays through
the global pointers, call the entry points.

Nothing here computes: there is no CPU fallback.  Loading fails loudly if a
library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_bool, c_char_p, c_double, c_int, c_uint64

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
SEG_LIB = os.path.join(LIB_DIR, "libseg4d_amd.so")
MCF_LIB = os.path.join(LIB_DIR, "libmcf4d_amd.so")

# order of the 8 faces everywhere: e(+x) w(-x) n(+y) s(-y) t(+z) b(-z) fw(+t) bw(-t)
FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")

GEN_MAX_OBJECTS = 64


class GenSpec(ctypes.Structure):
    """Mirror of seg4d_gen_spec (include/seg4d_gen.h)."""

    _fields_ = [
        ("nx", c_int), ("ny", c_int), ("nz", c_int), ("nt", c_int),
        ("n_objects", c_int),
        ("axes", (c_double * 3) * GEN_MAX_OBJECTS),
        ("semi_min", c_double), ("semi_max", c_double),
        ("margin", c_double), ("amp", c_double),
        ("wedges", c_int), ("wedge_deg", c_double), ("shell", c_double),
        ("noise_sigma", c_double),
        ("background", c_double), ("foreground", c_double),
        ("seed", c_uint64),
    ]


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run `make -C 4d-image-segmentation-algorithm_amd` "
            "(or __graft_entry__.build()) first")
    return ctypes.CDLL(path, mode=os.RTLD_LOCAL)


class _Lib:
    """Common global-variable plumbing of both libraries."""

    _int_globals: tuple = ()
    _dbl_globals: tuple = ()
    _arrays: tuple = ()

    def __init__(self, path: str):
        self.lib = _load(path)

    def set(self, **kw):
        for k, v in kw.items():
            if k in self._int_globals:
                c_int.in_dll(self.lib, k).value = int(v)
            elif k in self._dbl_globals:
                c_double.in_dll(self.lib, k).value = float(v)
            else:
                raise KeyError(k)

    def get(self, name):
        if name in self._int_globals:
            return c_int.in_dll(self.lib, name).value
        if name in self._dbl_globals:
            return c_double.in_dll(self.lib, name).value
        raise KeyError(name)

    @property
    def shape(self):
        return (self.get("lMax"), self.get("kMax"), self.get("jMax"), self.get("iMax"))

    def ptr(self, name):
        """The global pointer `name` (a host array owned by the library)."""
        return POINTER(c_double).in_dll(self.lib, name)

    def arr(self, name) -> np.ndarray:
        """Writable numpy view (l, k, j, i) of the host array `name`."""
        p = self.ptr(name)
        if not p:
            raise RuntimeError(f"{name} is not allocated")
        n = int(np.prod(self.shape))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.shape)


class SegLib(_Lib):
    """libseg4d_amd.so: the segmentation program's ABI."""

    _int_globals = ("iMax", "jMax", "kMax", "lMax", "maxNoGSIteration",
                    "numberOfTimeStep", "maxNoOfTimeSteps", "mod", "timeStepsNum")
    _dbl_globals = ("eps2", "K", "segTol", "tau", "h", "omega_c", "gsTol",
                    "omega_sigma", "theta_img", "rho_thr", "alpha", "beta",
                    "sigma", "maxNoGSIteration_sigma", "ballRadius",
                    "spacing_x", "spacing_y", "spacing_z",
                    "origin_x", "origin_y", "origin_z")
    _arrays = ("u", "prev_u", "i0", "Mmask") + FACES + tuple("g_" + f for f in FACES)

    def __init__(self, path: str = SEG_LIB):
        super().__init__(path)
        L = self.lib
        for fn in ("_allocateMem_", "_deallocateMem_", "_gs_coefficients_",
                   "_segTimeStep_", "seg4dSynchronize"):
            getattr(L, fn).restype = c_bool
            getattr(L, fn).argtypes = []
        for fn in ("_presmoothStep_", "_rescale4DdataToZeroOne_", "_reflection4Ddata_",
                   "_generate4DsegmFunct_", "seg4dUpload", "seg4dDownload"):
            getattr(L, fn).restype = c_bool
            getattr(L, fn).argtypes = [POINTER(c_double)]
        L._gf_inputImage_plus_thresholdedImage_.restype = c_bool
        L._gf_inputImage_plus_thresholdedImage_.argtypes = [POINTER(c_double)] * 2
        L._l2norm_.restype = c_double
        L._l2norm_.argtypes = [POINTER(c_double)] * 2
        L._locallyThreshold4Ddata_.restype = c_bool
        L._locallyThreshold4Ddata_.argtypes = [POINTER(c_double)] * 2
        for fn in ("_get4DMax_", "_mass4D_"):
            getattr(L, fn).restype = c_double
            getattr(L, fn).argtypes = [POINTER(c_double)]
        for fn in ("_get3DMaxWithinR_", "_get3DMinWithinR_"):
            getattr(L, fn).restype = c_double
            getattr(L, fn).argtypes = [POINTER(c_double), c_int] + [c_double] * 4
        L._copy4DdataFromSrcToDest_.restype = None
        L._copy4DdataFromSrcToDest_.argtypes = [POINTER(c_double)] * 2
        L.ijkl.restype = c_int
        L.ijkl.argtypes = [c_int] * 4
        L._gFunction_.restype = c_double
        L._gFunction_.argtypes = [c_double, c_double]
        L.seg4dSetResident.restype = c_int
        L.seg4dSetResident.argtypes = [c_int]
        L.seg4dSetRelativeTolerance.argtypes = [c_double]
        L.seg4dHostWrote.argtypes = [POINTER(c_double)]
        L.seg4dSetWriteTracking.restype = c_int
        L.seg4dSetWriteTracking.argtypes = [c_int]
        L.seg4dWriteTrackingState.restype = c_int
        L.seg4dWriteTrackingState.argtypes = []
        L.seg4dSetVerbose.argtypes = [c_int]
        L.seg4dLastError.restype = c_char_p
        L.seg4dLastSweeps.restype = c_int
        L.seg4dLastResidual.restype = c_double
        L.seg4dLastTolerance.restype = c_double
        L.seg4dProfile.argtypes = [c_int]
        L.seg4dProfileRead.restype = c_int
        L.seg4dProfileRead.argtypes = [POINTER(c_double), c_int]
        L.seg4dSetSweepVariant.restype = c_int
        L.seg4dSetSweepVariant.argtypes = [c_int]
        L.seg4dSetSweepTile.restype = c_int
        L.seg4dSetSweepTile.argtypes = [c_int, c_int]
        L.seg4dSetInnerOmega.restype = c_double
        L.seg4dSetInnerOmega.argtypes = [c_double]
        L.seg4dSetFusedSweeps.restype = c_int
        L.seg4dSetFusedSweeps.argtypes = [c_int]
        L.seg4dLastSolveFused.restype = c_int
        for fn in ("seg4dSetSweepChunk", "seg4dSetPipelineDepth", "seg4dSetSweepBands"):
            getattr(L, fn).restype = c_int
            getattr(L, fn).argtypes = [c_int]
        L.seg4dSetPrecision.restype = c_int
        L.seg4dSetPrecision.argtypes = [c_int]
        L.seg4dSetRefinement.restype = None
        L.seg4dSetRefinement.argtypes = [c_int, c_double]
        L.seg4dLastRounds.restype = c_int
        L.seg4dSetWarmStart.restype = c_int
        L.seg4dSetWarmStart.argtypes = [c_int]
        L.seg4dNetInit.restype = c_int
        L.seg4dNetInit.argtypes = [c_uint64]
        L.seg4dNetInitShape.restype = c_int
        L.seg4dNetInitShape.argtypes = [c_uint64, c_int, c_int]
        L.seg4dNetShape.restype = c_int
        L.seg4dNetShape.argtypes = [POINTER(c_int), POINTER(c_int)]
        for fn in ("seg4dNetLoad", "seg4dNetSave"):
            getattr(L, fn).restype = c_int
            getattr(L, fn).argtypes = [c_char_p]
        L.seg4dNetRecord.restype = c_int
        L.seg4dNetRecord.argtypes = [c_int]
        L.seg4dNetPoolPairs.restype = ctypes.c_long
        L.seg4dNetTrain.restype = c_double
        L.seg4dNetTrain.argtypes = [c_int, c_int, c_double, c_uint64, c_int]
        L.seg4dNetParams.restype = c_int
        L.seg4dNetParams.argtypes = [POINTER(ctypes.c_float), c_int]
        L.seg4dNetSetParams.restype = c_int
        L.seg4dNetSetParams.argtypes = [POINTER(ctypes.c_float), c_int, ctypes.c_float,
                                        ctypes.c_float]
        L.seg4dNetScales.restype = None
        L.seg4dNetScales.argtypes = [POINTER(ctypes.c_float)] * 2
        L.seg4dNetForward.restype = c_int
        L.seg4dNetForward.argtypes = [c_int] * 5 + [POINTER(ctypes.c_uint16),
                                                    POINTER(ctypes.c_float)]
        L.seg4dNetGrad.restype = c_int
        L.seg4dNetGrad.argtypes = [c_int] * 5 + [POINTER(ctypes.c_uint16),
                                                 POINTER(ctypes.c_float),
                                                 POINTER(ctypes.c_ubyte),
                                                 POINTER(ctypes.c_float), POINTER(c_double)]
        L.seg4dNetPreview.restype = c_int
        L.seg4dNetPreview.argtypes = [POINTER(ctypes.c_float), POINTER(c_double)]
        L.seg4dNetPredictMs.restype = c_double
        L.seg4dNetPredicts.restype = ctypes.c_long
        L.seg4dDeviceInfo.restype = c_int
        L.seg4dDeviceInfo.argtypes = [c_char_p, c_int]
        L.seg4dAutotune.restype = c_int
        L.seg4dAutotune.argtypes = [c_int, POINTER(c_int), POINTER(c_int),
                                    POINTER(ctypes.c_float)]
        L.seg4dBatchCreate.restype = c_int
        L.seg4dBatchCreate.argtypes = [c_int, POINTER(c_int)]
        L.seg4dBatchDestroy.restype = None
        L.seg4dBatchSetVolume.restype = c_int
        L.seg4dBatchSetVolume.argtypes = [c_int] + [POINTER(c_double)] * 3
        L.seg4dBatchStep.restype = c_int
        L.seg4dBatchGetU.restype = c_int
        L.seg4dBatchGetU.argtypes = [c_int, POINTER(c_double)]
        L.seg4dBatchStats.restype = c_int
        L.seg4dBatchStats.argtypes = [c_int, POINTER(c_int), POINTER(c_int), POINTER(c_double),
                                      POINTER(c_double)]
        L.seg4dBatchLastStepMs.restype = c_double
        L.seg4dBatchConfigure.restype = c_int
        L.seg4dBatchConfigure.argtypes = [c_int] * 4
        L.seg4dBatchSetStreams.restype = c_int
        L.seg4dBatchSetStreams.argtypes = [c_int]
        L.seg4dBatchAutotune.restype = c_int
        L.seg4dBatchAutotune.argtypes = [c_int, POINTER(ctypes.c_float)]
        L.seg4dBatchTile.restype = c_int
        L.seg4dBatchTile.argtypes = [c_int, POINTER(c_int), POINTER(c_int)]
        L.seg4dBatchSetGraphPlan.restype = c_int
        L.seg4dBatchSetGraphPlan.argtypes = [c_int, c_int]
        L.seg4dBatchGraphInfo.restype = c_int
        L.seg4dBatchGraphInfo.argtypes = [POINTER(c_int), POINTER(c_int),
                                          POINTER(ctypes.c_long), POINTER(ctypes.c_long),
                                          POINTER(c_double)]
        L.seg4dInjectStepFailure.restype = c_int
        L.seg4dInjectStepFailure.argtypes = [c_int]
        L.seg4d_gen_preset.restype = c_int
        L.seg4d_gen_preset.argtypes = [c_int, c_uint64, POINTER(GenSpec)]
        L.seg4d_generate.restype = c_int
        L.seg4d_generate.argtypes = [POINTER(GenSpec), POINTER(c_double),
                                     POINTER(c_double), POINTER(c_double)]
        self._centres = None  # keeps no_of_centers alive

    # -- configuration --------------------------------------------------------
    def set_centres(self, counts):
        """no_of_centers[l] for l = 0..lMax-1 (caller-owned, as the reference)."""
        arr = (c_int * len(counts))(*[int(c) for c in counts])
        self._centres = arr
        ctypes.c_void_p.in_dll(self.lib, "no_of_centers").value = ctypes.addressof(arr)

    def centre_arrays(self):
        """(center_x, center_y, center_z) as ctypes double** globals."""
        return tuple(POINTER(POINTER(c_double)).in_dll(self.lib, n)
                     for n in ("center_x", "center_y", "center_z"))

    def allocate(self, nx, ny, nz, nt, centres=None):
        self.set(iMax=nx + 2, jMax=ny + 2, kMax=nz + 2, lMax=nt + 2)
        if centres is not None:
            self.set_centres(centres)
        else:
            self._centres = None
            ctypes.c_void_p.in_dll(self.lib, "no_of_centers").value = None
        if not self.lib._allocateMem_():
            raise RuntimeError(f"_allocateMem_ failed: {self.last_error()}")

    def deallocate(self):
        self.lib._deallocateMem_()

    def last_error(self) -> str:
        return (self.lib.seg4dLastError() or b"").decode()

    def check(self, ok, what):
        if not ok:
            raise RuntimeError(f"{what} failed: {self.last_error()}")

    # -- entry points (by global array name) ----------------------------------
    def seg_time_step(self):
        self.check(self.lib._segTimeStep_(), "_segTimeStep_")
        return self.lib.seg4dLastSweeps(), self.lib.seg4dLastResidual()

    def gs_coefficients(self):
        self.check(self.lib._gs_coefficients_(), "_gs_coefficients_")

    def presmooth(self, name):
        self.check(self.lib._presmoothStep_(self.ptr(name)), "_presmoothStep_")
        return self.lib.seg4dLastSweeps(), self.lib.seg4dLastResidual()

    def gf(self, a="i0", b="u"):
        self.check(self.lib._gf_inputImage_plus_thresholdedImage_(self.ptr(a), self.ptr(b)),
                   "_gf_inputImage_plus_thresholdedImage_")

    def reflect(self, name):
        self.check(self.lib._reflection4Ddata_(self.ptr(name)), "_reflection4Ddata_")

    def rescale(self, name="u"):
        self.check(self.lib._rescale4DdataToZeroOne_(self.ptr(name)), "_rescale4DdataToZeroOne_")

    def l2norm(self, a, b):
        return self.lib._l2norm_(self.ptr(a), self.ptr(b))

    def set_balls(self, centres, radius):
        """centres[l] = list of (cx, cy, cz) for l = 0..lMax-1, written into
        the centre arrays _allocateMem_ sized from no_of_centers."""
        cx, cy, cz = self.centre_arrays()
        for l, lst in enumerate(centres):
            for c, (x, y, z) in enumerate(lst):
                cx[l][c], cy[l][c], cz[l][c] = x, y, z
        self.set(ballRadius=float(radius))

    def locally_threshold(self, src="i0", dst="u"):
        self.check(self.lib._locallyThreshold4Ddata_(self.ptr(src), self.ptr(dst)),
                   "_locallyThreshold4Ddata_")

    def generate_seed(self, name="u"):
        self.check(self.lib._generate4DsegmFunct_(self.ptr(name)), "_generate4DsegmFunct_")

    def get4dmax(self, name):
        return self.lib._get4DMax_(self.ptr(name))

    def ball_max(self, name, l, c, R):
        return self.lib._get3DMaxWithinR_(self.ptr(name), l, *c, R)

    def ball_min(self, name, l, c, R):
        return self.lib._get3DMinWithinR_(self.ptr(name), l, *c, R)

    def mass(self, name):
        return self.lib._mass4D_(self.ptr(name))

    def copy(self, src, dst):
        self.lib._copy4DdataFromSrcToDest_(self.ptr(src), self.ptr(dst))

    def upload(self, name):
        self.check(self.lib.seg4dUpload(self.ptr(name)), "seg4dUpload")

    def download(self, name):
        self.check(self.lib.seg4dDownload(self.ptr(name)), "seg4dDownload")

    def host_wrote(self, name):
        self.lib.seg4dHostWrote(self.ptr(name))

    def set_resident(self, on):
        return self.lib.seg4dSetResident(1 if on else 0)

    def set_write_tracking(self, on):
        """Mirror-mode host-write tracking of u, prev_u and Mmask; returns the
        previous setting."""
        return self.lib.seg4dSetWriteTracking(1 if on else 0)

    def write_tracking_state(self):
        """Bit 0: tracking active; bits 1-3: u, prev_u, Mmask clean."""
        return self.lib.seg4dWriteTrackingState()

    def set_relative_tolerance(self, rel):
        self.lib.seg4dSetRelativeTolerance(float(rel))

    def set_verbose(self, on):
        self.lib.seg4dSetVerbose(1 if on else 0)

    def profile(self, on=True):
        self.lib.seg4dProfile(1 if on else 0)

    def profile_reset(self):
        self.lib.seg4dProfileReset()

    def profile_read(self):
        out = (c_double * 25)()
        self.lib.seg4dProfileRead(out, 25)
        keys = ("red0_ms", "red0_launches", "red1_ms", "red1_launches",
                "black_ms", "black_launches", "setup_ms", "setup_launches",
                "step_ms", "steps", "voxel_updates", "iterations", "nred", "nblack",
                "red0_f32_ms", "red0_f32_launches", "red1_f32_ms", "red1_f32_launches",
                "black_f32_ms", "black_f32_launches", "defect_ms", "defect_launches",
                "rounds", "copy_ms", "copy_launches")
        return dict(zip(keys, list(out)))

    def set_precision(self, mode):
        """0 = FP64 (reference arithmetic), 1 = FP32-refined defect correction."""
        return self.lib.seg4dSetPrecision(int(mode))

    def set_refinement(self, max_rounds=0, inner_floor=0.0):
        self.lib.seg4dSetRefinement(int(max_rounds), float(inner_floor))

    def last_rounds(self):
        return self.lib.seg4dLastRounds()

    def set_warm_start(self, mode):
        """0 = start SOR from u^n (reference), 1 = from 2 u^n - u^(n-1),
        2 = from the learned network's prediction."""
        return self.lib.seg4dSetWarmStart(int(mode))

    # -- learned warm start ---------------------------------------------------
    def net_init(self, seed=0, layers=None, base=None):
        """Fresh network weights; `layers` (1..4, default 4) layers correcting
        the base of order `base` (0: u^n itself; 1 linear, default; 3 cubic)."""
        if layers is None and base is None:
            rc = self.lib.seg4dNetInit(int(seed))
        else:
            rc = self.lib.seg4dNetInitShape(int(seed), 4 if layers is None else int(layers),
                                            1 if base is None else int(base))
        if rc != 0:
            raise RuntimeError(f"seg4dNetInit failed: {self.last_error()}")

    def net_shape(self):
        """(layers, base order) of the current network."""
        a, b = c_int(), c_int()
        if self.lib.seg4dNetShape(ctypes.byref(a), ctypes.byref(b)) != 0:
            raise RuntimeError(f"seg4dNetShape failed: {self.last_error()}")
        return a.value, b.value

    def net_load(self, path):
        if self.lib.seg4dNetLoad(str(path).encode()) != 0:
            raise RuntimeError(f"seg4dNetLoad failed: {self.last_error()}")

    def net_save(self, path):
        if self.lib.seg4dNetSave(str(path).encode()) != 0:
            raise RuntimeError(f"seg4dNetSave failed: {self.last_error()}")

    def net_record(self, new_run=False):
        if self.lib.seg4dNetRecord(1 if new_run else 0) != 0:
            raise RuntimeError(f"seg4dNetRecord failed: {self.last_error()}")

    def net_pool_pairs(self):
        return self.lib.seg4dNetPoolPairs()

    def net_train(self, iters, batch=4, lr=1e-3, seed=0, log_every=0):
        loss = self.lib.seg4dNetTrain(int(iters), int(batch), float(lr), int(seed),
                                      int(log_every))
        if loss < 0:
            raise RuntimeError(f"seg4dNetTrain failed: {self.last_error()}")
        return loss

    def net_params(self) -> np.ndarray:
        n = self.lib.seg4dNetParams(None, 0)
        if n < 0:
            raise RuntimeError(f"seg4dNetParams failed: {self.last_error()}")
        out = np.zeros(n, dtype=np.float32)
        self.lib.seg4dNetParams(out.ctypes.data_as(POINTER(ctypes.c_float)), n)
        return out

    def net_set_params(self, w, s_d=1.0, s_t=1.0):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if self.lib.seg4dNetSetParams(w.ctypes.data_as(POINTER(ctypes.c_float)), w.size,
                                      float(s_d), float(s_t)) != 0:
            raise RuntimeError(f"seg4dNetSetParams failed: {self.last_error()}")

    def net_scales(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        self.lib.seg4dNetScales(ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def net_forward(self, feat_bits: np.ndarray, P: int) -> np.ndarray:
        """feat_bits: uint16 BF16 patterns of shape (L, K, J, I, 8)."""
        L_, K_, J_, I_, _ = feat_bits.shape
        f = np.ascontiguousarray(feat_bits, dtype=np.uint16)
        out = np.zeros((L_, K_, J_, I_), dtype=np.float32)
        if self.lib.seg4dNetForward(I_, J_, K_, L_, P, f.ctypes.data_as(POINTER(ctypes.c_uint16)),
                                    out.ctypes.data_as(POINTER(ctypes.c_float))) != 0:
            raise RuntimeError(f"seg4dNetForward failed: {self.last_error()}")
        return out

    def net_grad(self, feat_bits, target, lmask, P):
        L_, K_, J_, I_, _ = feat_bits.shape
        f = np.ascontiguousarray(feat_bits, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(lmask, dtype=np.uint8)
        g = np.zeros(self.lib.seg4dNetParams(None, 0), dtype=np.float32)
        loss = c_double()
        if self.lib.seg4dNetGrad(I_, J_, K_, L_, P, f.ctypes.data_as(POINTER(ctypes.c_uint16)),
                                 t.ctypes.data_as(POINTER(ctypes.c_float)),
                                 m.ctypes.data_as(POINTER(ctypes.c_ubyte)),
                                 g.ctypes.data_as(POINTER(ctypes.c_float)),
                                 ctypes.byref(loss)) != 0:
            raise RuntimeError(f"seg4dNetGrad failed: {self.last_error()}")
        return g, loss.value

    def net_preview(self, shape):
        """(c, q): the network's correction of the next start (FP32, padded
        shape) and its scale, from the current resident state."""
        c = np.zeros(shape, dtype=np.float32)
        q = c_double()
        if self.lib.seg4dNetPreview(c.ctypes.data_as(POINTER(ctypes.c_float)), ctypes.byref(q)) != 0:
            raise RuntimeError(f"seg4dNetPreview failed: {self.last_error()}")
        return c, q.value

    def net_predict_stats(self):
        return self.lib.seg4dNetPredictMs(), self.lib.seg4dNetPredicts()

    def set_sweep_variant(self, v):
        return self.lib.seg4dSetSweepVariant(int(v))

    def set_fused_sweeps(self, mode):
        """0 one launch per pass (default), 1 the whole solve in one
        cooperative launch whenever it fits."""
        return self.lib.seg4dSetFusedSweeps(int(mode))

    def last_solve_fused(self):
        return self.lib.seg4dLastSolveFused()

    def autotune(self, reps=3):
        """Tile of the fastest SOR launch order on the current volume:
        (runs, planes, ms); (0, 0, ms) is the linear order."""
        tb, tk, ms = c_int(), c_int(), ctypes.c_float()
        if self.lib.seg4dAutotune(int(reps), ctypes.byref(tb), ctypes.byref(tk),
                                  ctypes.byref(ms)) != 0:
            raise RuntimeError(f"seg4dAutotune failed: {self.last_error()}")
        return tb.value, tk.value, ms.value

    # -- batched form -------------------------------------------------------------
    def batch_create(self, dims):
        """dims: list of interior (nx, ny, nz, nt)."""
        flat = (c_int * (4 * len(dims)))(*[int(x) for d in dims for x in d])
        if self.lib.seg4dBatchCreate(len(dims), flat) != 0:
            raise RuntimeError(f"seg4dBatchCreate failed: {self.last_error()}")
        self._batch_dims = [tuple(int(x) for x in d) for d in dims]

    def batch_destroy(self):
        self.lib.seg4dBatchDestroy()
        self._batch_dims = []

    def batch_shape(self, v):
        nx, ny, nz, nt = self._batch_dims[v]
        return (nt + 2, nz + 2, ny + 2, nx + 2)

    def batch_set_volume(self, v, img, u0, mask):
        shp = self.batch_shape(v)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (img, u0, mask)]
        for a in arrs:
            if a.shape != shp:
                raise ValueError(f"volume {v}: array shape {a.shape}, expected {shp}")
        if self.lib.seg4dBatchSetVolume(v, *[a.ctypes.data_as(POINTER(c_double))
                                              for a in arrs]) != 0:
            raise RuntimeError(f"seg4dBatchSetVolume failed: {self.last_error()}")

    def batch_configure(self, precision=0, warm=0, graph=0, depth=2):
        if self.lib.seg4dBatchConfigure(int(precision), int(warm), int(graph), int(depth)) != 0:
            raise RuntimeError(f"seg4dBatchConfigure failed: {self.last_error()}")

    def batch_streams(self, nstreams):
        """Concurrent form: the volumes in `nstreams` groups on their own streams
        (0: one launch per pass for all volumes)."""
        if self.lib.seg4dBatchSetStreams(int(nstreams)) != 0:
            raise RuntimeError(f"seg4dBatchSetStreams failed: {self.last_error()}")

    def batch_step(self):
        if self.lib.seg4dBatchStep() != 0:
            raise RuntimeError(f"seg4dBatchStep failed: {self.last_error()}")
        return self.lib.seg4dBatchLastStepMs()

    def batch_u(self, v):
        out = np.zeros(self.batch_shape(v))
        if self.lib.seg4dBatchGetU(v, out.ctypes.data_as(POINTER(c_double))) != 0:
            raise RuntimeError(f"seg4dBatchGetU failed: {self.last_error()}")
        return out

    def batch_stats(self, v):
        z, r, R, tol = c_int(), c_int(), c_double(), c_double()
        if self.lib.seg4dBatchStats(v, ctypes.byref(z), ctypes.byref(r), ctypes.byref(R),
                                    ctypes.byref(tol)) != 0:
            raise RuntimeError(f"seg4dBatchStats failed: {self.last_error()}")
        return z.value, r.value, R.value, tol.value

    def batch_graph_info(self):
        """Graph mode 2: (iterations per solve and rounds the next captured step
        holds, steps run as one graph, steps redone eagerly, host ms spent
        capturing and instantiating)."""
        it, rd, st, re, hm = c_int(), c_int(), ctypes.c_long(), ctypes.c_long(), c_double()
        if self.lib.seg4dBatchGraphInfo(ctypes.byref(it), ctypes.byref(rd), ctypes.byref(st),
                                        ctypes.byref(re), ctypes.byref(hm)) != 0:
            raise RuntimeError(f"seg4dBatchGraphInfo failed: {self.last_error()}")
        return it.value, rd.value, st.value, re.value, hm.value

    def batch_set_graph_plan(self, iters, rounds):
        if self.lib.seg4dBatchSetGraphPlan(int(iters), int(rounds)) != 0:
            raise RuntimeError(f"seg4dBatchSetGraphPlan failed: {self.last_error()}")

    def batch_autotune(self, reps=3):
        n = len(self._batch_dims)
        ms = (ctypes.c_float * n)()
        if self.lib.seg4dBatchAutotune(int(reps), ms) != 0:
            raise RuntimeError(f"seg4dBatchAutotune failed: {self.last_error()}")
        tiles = []
        for v in range(n):
            tb, tk = c_int(), c_int()
            self.lib.seg4dBatchTile(v, ctypes.byref(tb), ctypes.byref(tk))
            tiles.append((tb.value, tk.value))
        return tiles, list(ms)

    def inject_step_failure(self, where):
        """Test hook: the next _segTimeStep_ fails after its set-up (1) or
        at its final wait (2), forced from the host."""
        if self.lib.seg4dInjectStepFailure(int(where)) != 0:
            raise RuntimeError("seg4dInjectStepFailure: no volume allocated")

    def synchronize(self):
        self.check(self.lib.seg4dSynchronize(), "seg4dSynchronize")

    def device_info(self) -> str:
        buf = ctypes.create_string_buffer(256)
        if self.lib.seg4dDeviceInfo(buf, 256) != 0:
            return "unknown"
        return buf.value.decode()

    # -- synthetic inputs -------------------------------------------------------
    def gen_spec(self, cfg: int, seed: int) -> GenSpec:
        spec = GenSpec()
        if self.lib.seg4d_gen_preset(int(cfg), int(seed), ctypes.byref(spec)) != 0:
            raise ValueError(f"unknown preset {cfg}")
        return spec

    def generate(self, spec: GenSpec, image: np.ndarray, u0: np.ndarray | None):
        """Fills padded float64 arrays in place; returns the seed centre."""
        centre = (c_double * 3)()
        assert image.dtype == np.float64 and image.flags.c_contiguous
        up = None
        if u0 is not None:
            assert u0.dtype == np.float64 and u0.flags.c_contiguous
            up = u0.ctypes.data_as(POINTER(c_double))
        rc = self.lib.seg4d_generate(ctypes.byref(spec), image.ctypes.data_as(POINTER(c_double)),
                                     up, centre)
        if rc != 0:
            raise ValueError("seg4d_generate rejected the spec")
        return tuple(centre)


class McfLib(_Lib):
    """libmcf4d_amd.so: the MCF verification program's ABI."""

    _int_globals = ("iMax", "jMax", "kMax", "lMax", "maxNoGSIteration",
                    "numberOfTimeStep", "maxNoOfTimeSteps")
    _dbl_globals = ("eps2", "mcfTol", "tau", "h", "omega_c", "gsTol")
    _arrays = ("u", "prev_u", "exact_u") + FACES

    def __init__(self, path: str = MCF_LIB):
        super().__init__(path)
        L = self.lib
        for fn in ("_allocateMem_mcf", "_deallocateMem_mcf", "_gs_coefficients_mcf",
                   "_mcfTimeStep", "mcf4dSynchronize"):
            getattr(L, fn).restype = c_bool
            getattr(L, fn).argtypes = []
        L._set4DBoundaryToExactValues_mcf.restype = c_bool
        L._set4DBoundaryToExactValues_mcf.argtypes = [c_double]
        L._exactSolution_mcf.restype = None
        L._exactSolution_mcf.argtypes = [c_double]
        L._l2norm_mcf.restype = c_double
        L._l2norm_mcf.argtypes = [POINTER(c_double)] * 2
        L._copy4DdataFromSrcToDest_mcf.restype = None
        L._copy4DdataFromSrcToDest_mcf.argtypes = [POINTER(c_double)] * 2
        for fn in ("mcf4dUpload", "mcf4dDownload"):
            getattr(L, fn).restype = c_bool
            getattr(L, fn).argtypes = [POINTER(c_double)]
        L.mcf4dSetWriteTracking.restype = c_int
        L.mcf4dSetWriteTracking.argtypes = [c_int]
        L.mcf4dWriteTrackingState.restype = c_int
        L.mcf4dWriteTrackingState.argtypes = []
        L.mcf4dSetResident.restype = c_int
        L.mcf4dSetResident.argtypes = [c_int]
        L.mcf4dSetRelativeTolerance.argtypes = [c_double]
        L.mcf4dSetVerbose.argtypes = [c_int]
        L.mcf4dLastError.restype = c_char_p
        L.mcf4dLastSweeps.restype = c_int
        L.mcf4dLastResidual.restype = c_double
        L.ijkl.restype = c_int
        L.ijkl.argtypes = [c_int] * 4

    def last_error(self) -> str:
        return (self.lib.mcf4dLastError() or b"").decode()

    def check(self, ok, what):
        if not ok:
            raise RuntimeError(f"{what} failed: {self.last_error()}")

    def allocate(self, n_interior):
        m = n_interior + 2
        self.set(iMax=m, jMax=m, kMax=m, lMax=m)
        self.check(self.lib._allocateMem_mcf(), "_allocateMem_mcf")

    def deallocate(self):
        self.lib._deallocateMem_mcf()

    def time_step(self):
        self.check(self.lib._mcfTimeStep(), "_mcfTimeStep")
        return self.lib.mcf4dLastSweeps(), self.lib.mcf4dLastResidual()

    def gs_coefficients(self):
        self.check(self.lib._gs_coefficients_mcf(), "_gs_coefficients_mcf")

    def boundary(self, t):
        self.check(self.lib._set4DBoundaryToExactValues_mcf(float(t)), "_set4DBoundary")

    def exact(self, t):
        self.lib._exactSolution_mcf(float(t))

    def l2norm(self, a, b):
        return self.lib._l2norm_mcf(self.ptr(a), self.ptr(b))

    def copy(self, src, dst):
        self.lib._copy4DdataFromSrcToDest_mcf(self.ptr(src), self.ptr(dst))

    def set_verbose(self, on):
        self.lib.mcf4dSetVerbose(1 if on else 0)

    def set_resident(self, on):
        return self.lib.mcf4dSetResident(1 if on else 0)

    def set_write_tracking(self, on):
        return self.lib.mcf4dSetWriteTracking(1 if on else 0)

    def write_tracking_state(self):
        """Bit 0: tracking active; bits 1-3: u, prev_u, exact_u clean."""
        return self.lib.mcf4dWriteTrackingState()

    def upload(self, name):
        self.check(self.lib.mcf4dUpload(self.ptr(name)), "mcf4dUpload")

    def download(self, name):
        self.check(self.lib.mcf4dDownload(self.ptr(name)), "mcf4dDownload")
