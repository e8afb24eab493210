"""Operator files (tv3_geometry_save/load, tv3_sparse_save/load; host only).

A description written and read back is the same description: every array
equal, and the oracle's operator built from it gives the same products.
Damaged files (wrong magic, truncation, out-of-range indices, broken row
pointers) are refused with an argument error.
"""
import numpy as np
import pytest

import tv3_oracle as O
import tval3_amd
from tval3_amd import files, synth


def _geo_equal(a, b):
    assert (a.num_mps, a.num_paths, a.rv, a.scale) == (b.num_mps, b.num_paths, b.rv, b.scale)
    for k in ("use_path", "xe", "ye", "ze", "xr", "yr", "zr"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


@pytest.mark.parametrize("geo_fn", [
    lambda: synth.ring_geometry(32, 40, scale=0.5),
    lambda: synth.multi_mp_ring_geometry(24, 32, 3, skip_every=4, extra_rows=5, seed=2),
    lambda: synth.hemisphere_geometry(40, 24),
])
def test_geometry_round_trip(tmp_path, geo_fn):
    geo = geo_fn()
    path = tmp_path / "g.tv3geom"
    files.save_geometry(path, geo)
    back = files.load_geometry(path)
    _geo_equal(back, geo)
    x = np.random.default_rng(0).random(geo.n).astype(np.float32)
    np.testing.assert_array_equal(O.Op.geometry(back).gemv(x), O.Op.geometry(geo).gemv(x))


def test_sparse_round_trip(tmp_path):
    geo = synth.ring_geometry(24, 32)
    ptr, lin = O.Op.geometry(geo).walk()
    arrs = synth.csr_from_walk(ptr, lin, geo.n, scale=0.25)
    path = tmp_path / "s.tv3spar"
    files.save_sparse(path, geo.m, geo.n, *arrs)
    m, n, *back = files.load_sparse(path)
    assert (m, n) == (geo.m, geo.n)
    for a, b in zip(back, arrs):
        np.testing.assert_array_equal(a, b)
    x = np.random.default_rng(1).random(n).astype(np.float32)
    np.testing.assert_array_equal(O.Op.sparse(m, n, *back).gemv(x), O.Op.sparse(m, n, *arrs).gemv(x))


def _damage(src, dst, fn):
    data = bytearray(open(src, "rb").read())
    open(dst, "wb").write(bytes(fn(data)))


def _put_int(data, off, v):
    data[off:off + 4] = np.int32(v).tobytes()
    return data


def test_damaged_files_are_refused(tmp_path):
    geo = synth.ring_geometry(16, 16)
    g = tmp_path / "g.bin"
    files.save_geometry(g, geo)
    bad = tmp_path / "bad.bin"
    cases = {
        "not a geometry file": lambda d: b"XXXXXXXX" + d[8:],
        "length does not match": lambda d: d[:-4],
        "invalid counts": lambda d: _put_int(d, 8, 0),          # num_mps = 0
        "use_path entry out of range": lambda d: _put_int(d, 8 + 32, geo.num_paths + 1),
    }
    for msg, fn in cases.items():
        _damage(g, bad, fn)
        with pytest.raises(tval3_amd.Tval3Error, match=msg):
            files.load_geometry(bad)
    ptr, lin = O.Op.geometry(geo).walk()
    arrs = synth.csr_from_walk(ptr, lin, geo.n)
    s = tmp_path / "s.bin"
    files.save_sparse(s, geo.m, geo.n, *arrs)
    m, n, z = geo.m, geo.n, arrs[1].size
    ri_off = 8 + 12 + 4 * (m + 1)            # first CSR column index
    cp_off = ri_off + 8 * z                  # first CSC column pointer
    cases = {
        "not a sparse-matrix file": lambda d: b"TV3GEOM1" + d[8:],
        "length does not match": lambda d: d + b"\0\0\0\0",
        "CSR column index out of range": lambda d: _put_int(d, ri_off, n),
        "row or column pointers": lambda d: _put_int(d, cp_off, 3),
    }
    for msg, fn in cases.items():
        _damage(s, bad, fn)
        with pytest.raises(tval3_amd.Tval3Error, match=msg):
            files.load_sparse(bad)
    with pytest.raises(tval3_amd.Tval3Error, match="cannot open"):
        files.load_sparse(tmp_path / "missing.bin")


@pytest.mark.gpu
def test_plans_from_files(tmp_path):
    """Plans made from loaded files give the products of plans made from the
    original arrays, bit for bit (multi-position geometry and its CSR/CSC)."""
    ctx = tval3_amd.Context()
    geo = synth.multi_mp_ring_geometry(24, 32, 2, skip_every=3, extra_rows=4, seed=9)
    files.save_geometry(tmp_path / "g", geo)
    back = files.load_geometry(tmp_path / "g")
    x = synth.disk_phantom(32, seed=3)
    ya = ctx.gemv(ctx.plan_geometry(geo), x).cpu().numpy()
    yb = ctx.gemv(ctx.plan_geometry(back), x).cpu().numpy()
    np.testing.assert_array_equal(ya, yb)
    ptr, lin = O.Op.geometry(geo).walk()
    arrs = synth.csr_from_walk(ptr, lin, geo.n)
    files.save_sparse(tmp_path / "s", geo.m, geo.n, *arrs)
    m, n, *sarrs = files.load_sparse(tmp_path / "s")
    ys = ctx.gemv(ctx.plan_sparse(m, n, *sarrs), x).cpu().numpy()
    np.testing.assert_array_equal(ys, O.Op.sparse(m, n, *arrs).gemv(x))
