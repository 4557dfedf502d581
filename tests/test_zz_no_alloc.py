"""lsmp_mass_vol_noalloc, the allocation-free entry point, against
lsmp_mass_vol on the same fields, called through the C ABI.

Both run the same block partial sums and the same fixed-order compensated
final reduction, so the results must agree bit for bit.  Two calls in a row,
the second on a smaller grid (the static scratch keeps the first call's
partial sums: the second must not read them), then a 2D field, then the first
grid again on a stream of its own.  Device memory and the stream come from
the HIP runtime the library itself links (through ctypes).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import load_lsmp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class Hip:
    """The few HIP runtime calls the test needs."""

    def __init__(self):
        h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        self.h = h

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), nbytes))
        return p

    def to_device(self, arr):
        p = self.malloc(arr.nbytes)
        self.check(self.h.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1))   # host to device
        return p

    def to_host(self, p, arr):
        self.check(self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), p, arr.nbytes, 2))   # device to host


def sphere_interior(nx, ny, nz, dim, c=(0.45, 0.5, 0.55), r=0.27):
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    if dim == 3:
        z = (np.arange(nz) + 0.5) / nz
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    else:
        Y, X = np.meshgrid(y, x, indexing="ij")
        X = X[None]
        d = np.sqrt((X[0] - c[0]) ** 2 + (Y - c[1]) ** 2)[None]
    return np.ascontiguousarray(r - d * (1 + 0.1 * np.sin(9 * X)))


def box_from_interior(inner):
    """Zero-gradient ghosts of 3 cells (7 z planes in 2D)."""
    return np.ascontiguousarray(np.pad(inner, 3, mode="edge"))


def handle_mass_vol(lsmp, inner, nx, ny, nz, dim, rho_12, inverse):
    p = lsmp.make_params(nx, dim=dim, ny=ny, nz=nz, rho_12=rho_12, inverse=inverse)
    with lsmp.Solver(p) as s:
        s.upload_phi(box_from_interior(inner))
        m, v = s.mass_vol()
    return p, m, v


def test_noalloc_matches_mass_vol():
    lsmp = load_lsmp()
    lsmp.load_library()
    hip = Hip()
    out = hip.malloc(16)
    side = C.c_void_p()
    hip.check(hip.h.hipStreamCreate(C.byref(side)))
    cases = [(64, 60, 56, 3, 1.0, 0), (20, 16, 12, 3, 0.25, 1), (96, 80, 1, 2, 0.5, 0), (64, 60, 56, 3, 1.0, 0)]
    try:
        for k, (nx, ny, nz, dim, r12, inv) in enumerate(cases):
            inner = sphere_interior(nx, ny, nz, dim)
            p, m_ref, v_ref = handle_mass_vol(lsmp, inner, nx, ny, nz, dim, r12, inv)
            phi = hip.to_device(inner)
            stream = side if k == len(cases) - 1 else None
            lsmp.mass_vol_noalloc(phi, nx, ny, nz, dim, p.dx, p.dy, p.dz, p.ls_wid, p.rho_12, p.inverse, out,
                                  stream)
            hip.check(hip.h.hipStreamSynchronize(stream) if stream is not None else hip.h.hipDeviceSynchronize())
            res = np.zeros(2)
            hip.to_host(out, res)
            hip.check(hip.h.hipFree(phi))
            m, v = res.tolist()
            say(f"call {k + 1}: {nx}x{ny}x{nz} dim {dim}: mass {m!r} / {m_ref!r}, vol {v!r} / {v_ref!r}"
                f"{' (own stream)' if stream is not None else ''}")
            assert m == m_ref and v == v_ref
    finally:
        hip.h.hipStreamDestroy(side)
        hip.h.hipFree(out)


def test_noalloc_rejects_bad_arguments():
    lsmp = load_lsmp()
    with pytest.raises(lsmp.LsmpError, match="-1"):
        lsmp.mass_vol_noalloc(None, 8, 8, 8, 3, 0.1, 0.1, 0.1, 0.15, 1.0, 0, None)
    with pytest.raises(lsmp.LsmpError, match="-1"):
        lsmp.mass_vol_noalloc(1, 8, 8, 2, 2, 0.1, 0.1, 1.0, 0.15, 1.0, 0, 1)
