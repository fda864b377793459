"""CPU checks of the product libraries that need no GPU: the synthetic
generator (known-answer tests of SplitMix64, properties of the volumes), the
exported C ABI (every symbol the headers declare), and the failure path
without a device (the product must refuse, not fall back to the CPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import seg4d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    L = seg4d.SegLib()
    L.set_verbose(False)
    return L


def test_splitmix64_known_answers(lib):
    """Vigna's reference splitmix64.c with state 0 produces these first three
    outputs."""
    f = lib.lib.seg4d_splitmix64_next
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    st = ctypes.c_uint64(0)
    got = [f(ctypes.byref(st)) for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_counter_rng_and_normals(lib):
    u = lib.lib.seg4d_rng_u64
    u.restype = ctypes.c_uint64
    u.argtypes = [ctypes.c_uint64] * 3
    # counter-based: draw n of stream s does not depend on earlier draws
    assert u(5, 1, 1000) == u(5, 1, 1000)
    assert u(5, 1, 1000) != u(5, 2, 1000)
    nrm = lib.lib.seg4d_rng_normal
    nrm.restype = ctypes.c_double
    nrm.argtypes = [ctypes.c_uint64] * 3
    x = np.array([nrm(7, 1, n) for n in range(20000)])
    assert abs(x.mean()) < 0.03
    assert abs(x.std() - 1.0) < 0.03
    assert abs(np.mean(x ** 3)) < 0.08  # symmetric


def test_generated_volume_properties(lib):
    spec = lib.gen_spec(0, 0)
    assert (spec.nx, spec.ny, spec.nz, spec.nt, spec.n_objects) == (64, 64, 32, 8, 2)
    shape = (spec.nt + 2, spec.nz + 2, spec.ny + 2, spec.nx + 2)
    img, u0 = np.zeros(shape), np.zeros(shape)
    c = lib.generate(spec, img, u0)
    inner = img[1:-1, 1:-1, 1:-1, 1:-1]
    # ghosts zero, values k/255
    assert img.sum() == inner.sum()
    assert np.all(inner >= 0) and np.all(inner <= 1)
    assert np.allclose(np.round(inner * 255), inner * 255, atol=1e-9)
    # two intensity populations around 40 and 200 (noise sigma 20)
    q = inner * 255
    assert 0.02 < np.mean(q > 120) < 0.6
    assert abs(np.median(q[q < 120]) - 40) < 3
    assert abs(np.median(q[q > 120]) - 200) < 3
    # seed function: 1/(|x-s|+1), maximal next to the centre, ghosts zero
    assert np.all(u0[0] == 0) and np.all(u0[..., 0] == 0)
    i, j, k = int(round(c[0])), int(round(c[1])), int(round(c[2]))
    d = np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2)
    assert u0[1, k, j, i] == pytest.approx(1.0 / (d + 1.0))
    # deterministic
    img2 = np.zeros(shape)
    lib.generate(spec, img2, None)
    assert np.array_equal(img, img2)
    # radius modulation: foreground volume varies over t
    fg = (q > 120).reshape(spec.nt, -1).mean(axis=1)
    assert fg.max() > fg.min()


def test_presets(lib):
    s1 = lib.gen_spec(1, 1)
    assert (s1.nx, s1.ny, s1.nz, s1.nt, s1.n_objects) == (256, 256, 128, 16, 6)
    s4 = lib.gen_spec(4, 4)
    assert (s4.nx, s4.ny, s4.nz, s4.nt, s4.n_objects) == (512, 512, 256, 8, 12)
    for seed in range(8):
        s3 = lib.gen_spec(3, seed)
        assert 48 <= s3.nx <= 192 and s3.nx % 16 == 0
        assert 32 <= s3.nz <= 128 and s3.nz % 16 == 0
        assert 6 <= s3.nt <= 16 and 2 <= s3.n_objects <= 6
    with pytest.raises(ValueError):
        lib.gen_spec(9, 0)


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    funcs = re.findall(r"^\s*(?:bool|double|int|void|const char \*|uint64_t)\s*\*?\s*(\w+)\s*\(",
                       text, flags=re.M)
    globs = []
    for line in text.splitlines():
        m = re.match(r"^extern\s+(int|double)\s+([^;(]+);", line)
        if m:
            globs += [x.strip().lstrip("*").strip() for x in m.group(2).split(",")]
    return funcs, globs


@pytest.mark.parametrize("header,libname", [("seg4d.h", "libseg4d_amd.so"),
                                            ("mcf4d.h", "libmcf4d_amd.so"),
                                            ("seg4d_gen.h", "libseg4d_amd.so")])
def test_every_declared_symbol_is_exported(header, libname):
    L = ctypes.CDLL(os.path.join(seg4d.LIB_DIR, libname), mode=os.RTLD_LOCAL)
    funcs, globs = _declared(header)
    assert funcs
    assert len(globs) >= 20 if header != "seg4d_gen.h" else True
    for name in funcs + globs:
        # in_dll raises ValueError for a symbol the library does not export
        ctypes.c_char.in_dll(L, name)


def test_host_helpers_without_device(lib):
    """ijkl and _gFunction_ are host functions of the ABI."""
    lib.set(iMax=9, jMax=8, kMax=7, lMax=6)
    assert lib.lib.ijkl(1, 2, 3, 4) == ((4 * 7 + 3) * 8 + 2) * 9 + 1
    assert lib.lib._gFunction_(3.0, 2.0) == 1.0 / 7.0


def test_allocate_fails_loudly_without_gpu(lib):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    with pytest.raises(RuntimeError, match="device"):
        lib.allocate(4, 4, 4, 4)
    lib.deallocate()
    assert lib.lib._segTimeStep_() is False
    assert "not allocated" in lib.last_error()


@pytest.mark.parametrize("which", ["seg", "mcf"])
def test_oversize_volume_fails_before_any_allocation(which):
    """A padded volume of 2^31 or more elements (the reference's int index,
    segmentation4D.c:61-65, mcf_4D.c:36-40) is refused before any host or
    device array exists: the global pointers stay NULL and the error names
    the limit.  Runs with or without a GPU (nothing reaches the device)."""
    L = seg4d.SegLib() if which == "seg" else seg4d.McfLib()
    L.set_verbose(False)
    alloc = L.lib._allocateMem_ if which == "seg" else L.lib._allocateMem_mcf
    for dims in ((1291, 1291, 1291, 3), (2002, 2002, 1002, 5)):
        # 1291^3 * 3 = 6,454,995,513 >= 2^31; the first fits a long, not an int
        L.set(iMax=dims[0], jMax=dims[1], kMax=dims[2], lMax=dims[3])
        assert alloc() is False
        assert "2^31" in L.last_error()
        assert not L.ptr("u") and not L.ptr("prev_u")
    L.set(iMax=0, jMax=0, kMax=0, lMax=0)
