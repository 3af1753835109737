"""AMReX patch data on the device through the C ABI (uclbs_import_fab /
uclbs_export_fab): a level whose state is imported from caller-owned device
FArrayBoxes (FP64, component-major, ghost width 2, padded component stride,
ghost cells NaN) steps bit for bit as one uploaded from the host, and the
export fills exactly the FABs' valid regions: D3Q27 FP32-shifted Taylor-Green
box, D3Q39 FP64 box with faces (f, g and the 13 multipliers).  Host pointers
and strides too small for the grown box are refused."""
import ctypes
import ctypes.util
import os

import numpy as np
import pytest

from test_gpu_d3q39_faces import fields

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", os.path.join("/opt/rocm/lib", "libamdhip64.so"),
                 ctypes.util.find_library("amdhip64") or ""):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.skip("libamdhip64 not found")


class Fab:
    """A device FArrayBox of ncomp components over `shape` grown by `ghost`."""

    def __init__(self, hip, shape, ncomp, ghost, pad=7):
        self.hip, self.shape, self.ncomp, self.ghost = hip, shape, ncomp, ghost
        nx, ny, nz = shape
        self.grown = (nz + 2 * ghost, ny + 2 * ghost, nx + 2 * ghost)
        self.cstride = int(np.prod(self.grown)) + pad
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.ncomp * self.cstride * 8)) == 0
        self.put(np.full((self.ncomp, self.cstride), np.nan))

    def put(self, host):
        h = np.ascontiguousarray(host, np.float64)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(h.nbytes), 1) == 0      # host to device

    def get(self):
        h = np.empty((self.ncomp, self.cstride))
        assert self.hip.hipMemcpy(h.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                  ctypes.c_size_t(h.nbytes), 2) == 0      # device to host
        return h

    def with_valid(self, dense):
        """FAB contents (ghosts NaN) holding dense [ncomp][n] in its valid region."""
        g = self.ghost
        nx, ny, nz = self.shape
        h = np.full((self.ncomp, self.cstride), np.nan)
        box = h[:, :np.prod(self.grown)].reshape((self.ncomp,) + self.grown)
        box[:, g:g + nz, g:g + ny, g:g + nx] = dense.reshape(self.ncomp, nz, ny, nx)
        return h

    def valid(self, h):
        g = self.ghost
        nx, ny, nz = self.shape
        box = h[:, :np.prod(self.grown)].reshape((self.ncomp,) + self.grown)
        return box[:, g:g + nz, g:g + ny, g:g + nx].reshape(self.ncomp, -1)

    def free(self):
        self.hip.hipFree(self.ptr)


def test_d3q27_fab_roundtrip_steps_like_upload(pkg, orc):
    hip = _hip()
    shape = (24, 20, 16)
    f0 = orc.init_tgv27(shape, 0.05, 1e-3, 5)
    a = pkg.Solver(precision="fp32s", nu=0.01)
    a.define_level(0, shape)
    a.upload_f(0, f0)
    a.step(0, 10)
    fa = a.download_f(0)
    b = pkg.Solver(precision="fp32s", nu=0.01)
    b.define_level(0, shape)
    src, dst = Fab(hip, shape, 27, 2), Fab(hip, shape, 27, 2)
    src.put(src.with_valid(f0))
    b.import_fab(0, "f", src.ptr.value, 2, src.cstride)
    b.step(0, 10)
    b.export_fab(0, "f", dst.ptr.value, 2, dst.cstride)
    out = dst.get()
    assert np.array_equal(dst.valid(out), fa)
    ghost = np.isnan(out).sum()
    assert ghost == out.size - fa.size          # ghost cells and padding untouched
    src.free(), dst.free()


def test_d3q39_fab_roundtrip_steps_like_upload(pkg, orc):
    hip = _hip()
    shape = (12, 10, 8)
    faces = [-1000, -10, -11, -6, -10, -11]
    c = orc.consts39(mach=0.3, shifted=False)
    f0, g0, lm0 = orc.init39(shape, *fields(shape, 0.3), c)

    def solver():
        s = pkg.Solver(lattice="d3q39", precision="fp64")
        s.set_gas(mach=0.3, shifted=0, turbulent=True)
        s.define_level(0, shape, periodic=(0, 0, 0))
        s.set_faces(0, faces)
        return s
    a = solver()
    a.upload_state(0, f0, g0, lm0)
    a.step(0, 10)
    fa, ga, la = a.download_state(0)
    b = solver()
    fabs = {"f": Fab(hip, shape, 39, 2), "g": Fab(hip, shape, 39, 2), "lm": Fab(hip, shape, 13, 2)}
    for name, dense in (("f", f0), ("g", g0), ("lm", lm0)):
        fabs[name].put(fabs[name].with_valid(dense))
        b.import_fab(0, name, fabs[name].ptr.value, 2, fabs[name].cstride)
    b.step(0, 10)
    for name, ref in (("f", fa), ("g", ga), ("lm", la)):
        out = Fab(hip, shape, ref.shape[0], 2)
        b.export_fab(0, name, out.ptr.value, 2, out.cstride)
        h = out.get()
        assert np.array_equal(out.valid(h), ref), name
        assert np.isnan(h).sum() == h.size - ref.size
        out.free()
    for fab in fabs.values():
        fab.free()


def test_fab_refusals(pkg):
    hip = _hip()
    shape = (8, 8, 8)
    s = pkg.Solver(precision="fp32s")
    s.define_level(0, shape)
    host = np.zeros(27 * 12 ** 3)
    with pytest.raises(pkg.UclbsError, match="device memory"):
        s.import_fab(0, "f", host.ctypes.data, 2, 12 ** 3)
    fab = Fab(hip, shape, 27, 2)
    with pytest.raises(pkg.UclbsError, match="stride"):
        s.import_fab(0, "f", fab.ptr.value, 2, 12 ** 3 - 1)
    with pytest.raises(pkg.UclbsError, match="field"):
        s.import_fab(0, "g", fab.ptr.value, 2, fab.cstride)
    fab.free()
