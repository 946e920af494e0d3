"""GPU tests of the exported-frame entries (dsmvc_hip_*_exported): planes of
GPU frames that another API exports as POSIX file descriptors, as
VapourSynth's GPU core does for the reference's CUDA executor (reference
include/dsmvc/engine.hpp:306-326, 414-442).  The producer here is
tests/native/libexportable.so: device memory made with the HIP virtual
memory API and exported anew for every call.

  * 2D F32 and u16, rows and columns through imported memory equal the
    device-pointer entries on the same data bit for bit;
  * an allocation is imported once per memory_id and later exports of it
    hit the cache, whose handles the library closes;
  * a plane outside its allocation is rejected and its handle closed;
  * a semaphore handle HIP cannot import is reported as not imported and
    closed;
  * the reference-side binding (reference_binding/hip_executor.cpp) takes
    the reference planner's plans and exported planes and equals its own
    host-buffer entry bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from dsmvc_hip import _lib  # noqa: E402
from support import (BICUBIC, ROOT, Conversion, make_request,
                     reference_available)  # noqa: E402
from workloads import integer_conversion  # noqa: E402

DEV = torch.device("cuda:0")
HELPER = ROOT / "tests" / "native" / "libexportable.so"
OPAQUE_FD = 0x1


class Exportable:
    """One exportable device allocation (tests/native/exportable.cpp)."""

    lib = None

    def __init__(self, nbytes, memory_id):
        if Exportable.lib is None:
            lib = ctypes.CDLL(str(HELPER))
            vp, sz = ctypes.c_void_p, ctypes.c_size_t
            lib.xp_create.argtypes = [ctypes.c_int, sz, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(vp), ctypes.POINTER(sz)]
            lib.xp_export.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
            lib.xp_copy_in.argtypes = [vp, vp, sz]
            lib.xp_copy_out.argtypes = [vp, vp, sz]
            lib.xp_destroy.argtypes = [ctypes.c_uint64, vp, sz]
            Exportable.lib = lib
        self.handle, self.ptr, self.size = ctypes.c_uint64(), ctypes.c_void_p(), ctypes.c_size_t()
        status = self.lib.xp_create(0, nbytes, ctypes.byref(self.handle), ctypes.byref(self.ptr),
                                    ctypes.byref(self.size))
        if status != 0:
            pytest.skip(f"exportable device memory unavailable (hipError {status})")
        self.memory_id = memory_id

    def write(self, offset, array):
        a = np.ascontiguousarray(array)
        assert self.lib.xp_copy_in(self.ptr.value + offset, a.ctypes.data, a.nbytes) == 0

    def read(self, offset, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.lib.xp_copy_out(out.ctypes.data, self.ptr.value + offset, out.nbytes) == 0
        return out

    def plane(self, offset, pitch_bytes, rows):
        fd = ctypes.c_int(-1)
        assert self.lib.xp_export(self.handle.value, ctypes.byref(fd)) == 0 and fd.value >= 0
        p = _lib.ExportedPlaneC()
        p.memory_id = self.memory_id
        p.memory_size = self.size.value
        p.offset = offset
        p.plane_bytes = pitch_bytes * rows
        p.handle_type = OPAQUE_FD
        p.handle = fd.value
        p.pitch_bytes = pitch_bytes
        return p

    def free(self):
        assert self.lib.xp_destroy(self.handle.value, self.ptr, self.size.value) == 0


def _closed(fd):
    try:
        os.fstat(fd)
    except OSError:
        return True
    return False


def _plans(kernel=dh.BICUBIC):
    return (dh.Plan(dh.AxisRequest(400, 256, 256.0, kernel=kernel)),
            dh.Plan(dh.AxisRequest(226, 144, 144.0, kernel=kernel)))


@pytest.fixture(scope="module")
def memory():
    if not HELPER.exists():
        pytest.skip("tests/native/libexportable.so is not built")
    src, dst = Exportable(8 << 20, 101), Exportable(8 << 20, 202)
    yield src, dst
    src.free()
    dst.free()


def test_2d_float32_equals_device_entry_and_reuses_imports(memory):
    src, dst = memory
    ex = dh.Executor(0, "auto")
    hp, vp = _plans()
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, (226, 400)).astype(np.float32)
    in_pitch, out_pitch, in_off, out_off = 448 * 4, 320 * 4, 4096, 12288
    padded = np.zeros((226, 448), np.float32)
    padded[:, :400] = x
    src.write(in_off, padded)
    expected = ex.inverse_2d(hp, vp, torch.from_numpy(x).to(DEV)).cpu().numpy()
    for call in range(3):
        pin, pout = src.plane(in_off, in_pitch, 226), dst.plane(out_off, out_pitch, 144)
        ex.inverse_2d_exported(hp, vp, pin, pout)
        got = dst.read(out_off, (144, 320), np.float32)[:, :256]
        assert np.array_equal(got, expected), call
        if call:
            # Cache hits: the library closed the new descriptors.
            assert _closed(pin.handle) and _closed(pout.handle)
    stats = ex.import_stats()
    assert stats["imports"] == 2 and stats["hits"] == 4 and stats["entries"] == 2


def test_2d_u16_equals_device_entry(memory):
    src, dst = memory
    ex = dh.Executor(0, "auto")
    hp, vp = _plans(dh.LANCZOS)
    conv = dh.integer_conversion(16, False, False)
    rng = np.random.default_rng(8)
    x = rng.integers(4096, 60000, (226, 400)).astype(np.uint16)
    src.write(0, x)
    expected = ex.inverse_2d(hp, vp, torch.from_numpy(x.astype(np.int32)).to(DEV)
                             .to(torch.uint16), conversion=conv).cpu().numpy()
    ex.inverse_2d_exported(hp, vp, src.plane(0, 800, 226), dst.plane(256, 512, 144),
                           dh.SAMPLE_KINDS["uint16"], conv)
    got = dst.read(256, (144, 256), np.uint16)
    assert np.array_equal(got, expected)


def test_rows_and_columns_equal_device_entries(memory):
    src, dst = memory
    ex = dh.Executor(0, "auto")
    hp, vp = _plans(dh.SPLINE36)
    rng = np.random.default_rng(9)
    rows = rng.uniform(0, 1, (70, 400)).astype(np.float32)
    src.write(0, rows)
    ex.inverse_rows_exported(hp, src.plane(0, 1600, 70), dst.plane(0, 1024, 70), 70)
    want = ex.inverse_rows(hp, torch.from_numpy(rows).to(DEV)).cpu().numpy()
    assert np.array_equal(dst.read(0, (70, 256), np.float32), want)
    cols = rng.uniform(0, 1, (226, 96)).astype(np.float32)
    src.write(0, cols)
    ex.inverse_columns_exported(vp, src.plane(0, 384, 226), dst.plane(0, 384, 144), 96)
    want = ex.inverse_columns(vp, torch.from_numpy(cols).to(DEV)).cpu().numpy()
    assert np.array_equal(dst.read(0, (144, 96), np.float32), want)


def test_rejected_planes_and_semaphores_consume_their_handles(memory):
    src, dst = memory
    ex = dh.Executor(0, "auto")
    hp, vp = _plans()
    pin = src.plane(src.size.value - 1024, 1600, 226)      # runs past the allocation
    pout = dst.plane(0, 1024, 144)
    with pytest.raises(ValueError, match="outside its allocation"):
        ex.inverse_2d_exported(hp, vp, pin, pout)
    assert _closed(pin.handle) and _closed(pout.handle)
    pin, pout = src.plane(0, 1000, 226), dst.plane(0, 1024, 144)   # pitch shorter than a row
    with pytest.raises(ValueError, match="geometry"):
        ex.inverse_2d_exported(hp, vp, pin, pout)
    assert _closed(pin.handle) and _closed(pout.handle)
    fd = os.open(os.devnull, os.O_RDONLY)
    assert not ex.import_exported_semaphore(77, OPAQUE_FD, fd)
    assert _closed(fd) and not ex.has_exported_semaphore(77)


BINDING_SO = ROOT / "oracle" / "_ref" / "libdsmvc_binding.so"


def test_reference_binding_exported_equals_its_host_entry(memory):
    if not reference_available() or not BINDING_SO.exists():
        pytest.skip("oracle/_ref (reference engine and binding) not built")
    from support import Request
    lib = ctypes.CDLL(str(BINDING_SO))
    vp, pd, i32 = ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_int32
    lib.bnd_last_error.restype = ctypes.c_char_p
    lib.bnd_create.argtypes = [i32, ctypes.POINTER(vp)]
    lib.bnd_free.argtypes = [vp]
    lib.bnd_inverse_2d.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(Request), i32,
                                   vp, pd, vp, pd, ctypes.POINTER(Conversion)]
    lib.bnd_inverse_2d_exported.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(Request),
                                            i32, ctypes.POINTER(_lib.ExportedPlaneC),
                                            ctypes.POINTER(_lib.ExportedPlaneC),
                                            ctypes.POINTER(Conversion)]
    src, dst = memory
    ex = vp()
    assert lib.bnd_create(0, ctypes.byref(ex)) == 0, lib.bnd_last_error()
    try:
        h, v = make_request(400, 256, kernel=BICUBIC), make_request(226, 144, kernel=BICUBIC)
        conv = integer_conversion(16, False, False)
        rng = np.random.default_rng(11)
        x = rng.integers(4096, 60000, (226, 400)).astype(np.uint16)
        want = np.empty((144, 256), np.uint16)
        assert lib.bnd_inverse_2d(ex, h, v, 2, x.ctypes.data, 400, want.ctypes.data, 256,
                                  conv) == 0, lib.bnd_last_error()
        src.write(0, x)
        pin, pout = src.plane(0, 800, 226), dst.plane(0, 512, 144)
        assert lib.bnd_inverse_2d_exported(ex, h, v, 2, pin, pout, conv) == 0, \
            lib.bnd_last_error()
        assert np.array_equal(dst.read(0, (144, 256), np.uint16), want)
    finally:
        lib.bnd_free(ex)
