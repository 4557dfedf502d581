"""lsmp_mass_vol_noalloc, the allocation-free entry point, against
lsmp_mass_vol on the same fields, called through the C ABI.

Both run the same block partial sums and the same fixed-order final
reduction, so the results must agree bit for bit.  Two calls in a row, the
second on a smaller grid (the static scratch keeps the first call's partial
sums: the second must not read them), then a 2D field, then the first grid
again from a non-default stream.  torch only provides device memory and
streams here.
"""
import numpy as np
import pytest

from conftest import load_lsmp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def sphere_interior(nx, ny, nz, dim, c=(0.45, 0.5, 0.55), r=0.27):
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    if dim == 3:
        z = (np.arange(nz) + 0.5) / nz
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    else:
        Y, X = np.meshgrid(y, x, indexing="ij")
        d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2)[None]
    return np.ascontiguousarray(r - d * (1 + 0.1 * np.sin(9 * X)))


def box_from_interior(inner, dim):
    """Zero-gradient ghosts of 3 cells (7 z planes in 2D)."""
    if dim == 3:
        return np.ascontiguousarray(np.pad(inner, 3, mode="edge"))
    return np.ascontiguousarray(np.pad(inner, ((3, 3), (3, 3), (3, 3)), mode="edge"))


def handle_mass_vol(lsmp, inner, nx, ny, nz, dim, rho_12, inverse):
    p = lsmp.make_params(nx, dim=dim, ny=ny, nz=nz, rho_12=rho_12, inverse=inverse)
    with lsmp.Solver(p) as s:
        s.upload_phi(box_from_interior(inner, dim))
        m, v = s.mass_vol()
    return p, m, v


def test_noalloc_matches_mass_vol():
    torch = pytest.importorskip("torch")
    lsmp = load_lsmp()
    assert torch.cuda.is_available()
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    cases = [(64, 60, 56, 3, 1.0, 0), (20, 16, 12, 3, 0.25, 1), (96, 80, 1, 2, 0.5, 0), (64, 60, 56, 3, 1.0, 0)]
    side = torch.cuda.Stream()
    for k, (nx, ny, nz, dim, r12, inv) in enumerate(cases):
        inner = sphere_interior(nx, ny, nz, dim)
        p, m_ref, v_ref = handle_mass_vol(lsmp, inner, nx, ny, nz, dim, r12, inv)
        phi = torch.from_numpy(inner.ravel()).to("cuda")
        torch.cuda.synchronize()
        stream = side if k == len(cases) - 1 else None
        handle = ctypes_stream(stream)
        ctx = torch.cuda.stream(stream) if stream is not None else _null()
        with ctx:
            lsmp.mass_vol_noalloc(phi.data_ptr(), nx, ny, nz, dim, p.dx, p.dy, p.dz, p.ls_wid, p.rho_12,
                                  p.inverse, out.data_ptr(), handle)
        torch.cuda.synchronize()
        m, v = out.cpu().numpy().tolist()
        say(f"call {k + 1}: {nx}x{ny}x{nz} dim {dim}: mass {m!r} / {m_ref!r}, vol {v!r} / {v_ref!r}"
            f"{' (side stream)' if stream is not None else ''}")
        assert m == m_ref and v == v_ref


def test_noalloc_rejects_bad_arguments():
    lsmp = load_lsmp()
    with pytest.raises(lsmp.LsmpError, match="-1"):
        lsmp.mass_vol_noalloc(None, 8, 8, 8, 3, 0.1, 0.1, 0.1, 0.15, 1.0, 0, None)
    with pytest.raises(lsmp.LsmpError, match="-1"):
        lsmp.mass_vol_noalloc(1, 8, 8, 2, 2, 0.1, 0.1, 1.0, 0.15, 1.0, 0, 1)


def ctypes_stream(stream):
    return None if stream is None else stream.cuda_stream


class _null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False
