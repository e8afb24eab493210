"""The C++ overload layer (include/tval3_gpu_2d.hpp, tval3_gpu_3d.hpp): the
reference's tval3_cpu_2d / tval3_cpu_3d call surface over the C-ABI.

The drivers under tests/cpp are built by build().  The argument checks run
without a GPU and must give the reference's messages (check_params,
2d/tval3_cpu_2d.cpp:696-750, 3d/tval3_cpu_3d.cpp:1028-1165).  The solves run
float and double scalars, the double/double overload and both complex
overloads, and are compared with the oracle given the same options (the layer
selects the reference's round-by-round backtracking); double/double and
complex<double> results are held to the FP64 bar of f64_bar.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import tv3_oracle as O
import f64_bar as F
import tval3_amd
from tval3_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def _driver(name):
    path = os.path.join(CPP, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    return path


EXPECT_2D = {
    "dense format": "Argument error: Matrices must be in row major format!",
    "dense Ut size": "Argument error: U and Ut must have the same size!",
    "dense U length": "Argument error: the length of U must be equal to A.dim_x * layers!",
    "dense b length": "Argument error: b.len must be equal to A.dim_y * layers!",
    "sparse format": "Argument error: A must be available in CSR and CSC format!",
    "geometry format": "Argument error: Matrices must be in row major format!",
    "geometry b length": "Argument error: b.len must be equal to A.num_emitters * A.num_receivers!",
}
EXPECT_3D = {
    "dense format": "Argument error: Matrices must be in row major format!",
    "dense Ut size": "Argument error: U and Ut must have the same size!",
    "dense U length": "Argument error: the length of U must be equal to A.dim_x!",
    "dense b length": "Argument error: b.len must be equal to A.dim_y!",
    "sparse format": "Argument error: A must be available in csr and csc format!",
    "geometry b length": "Argument error: b.len must be equal to A.num_emitters * A.num_receivers!",
}


@pytest.mark.parametrize("name,expect", [("shim_2d", EXPECT_2D), ("shim_3d", EXPECT_3D)])
def test_argument_checks(name, expect):
    out = subprocess.run([_driver(name), "args"], capture_output=True, text=True,
                         timeout=60, check=True).stdout
    got = dict(line.split(": ", 1) for line in out.strip().splitlines())
    for case, msg in expect.items():
        assert got[case] == msg, case
    # the double/double and complex overloads run the reference's checks
    assert got["double b length"] == expect["dense b length"]
    assert got["complex b length"] == expect["dense b length"]
    if name == "shim_2d":
        assert got["dense layers b length"] == "Argument error: b.len must be equal to A.dim_y * layers!"


# ---- GPU solves through the overload layer ------------------------------------
def _write(d, meta, **arrays):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {v}\n")
    for k, a in arrays.items():
        a.tofile(os.path.join(d, k + ".bin"))


def _complex_op(op, kind, d):
    """The oracle operator of the complex runs: the same entries as complex
    values (the projector is the same operator)."""
    if kind == "dense":
        m, n = op.m, op.n
        A = np.fromfile(os.path.join(d, "A.bin"), np.float32).reshape(m, n)
        return O.Op.dense_c(A.astype(np.complex64))
    if kind == "sparse":
        ld = lambda k, t: np.fromfile(os.path.join(d, k + ".bin"), t)  # noqa: E731
        return O.Op.sparse_c(op.m, op.n, ld("csr_ptr", np.int32), ld("csr_ind", np.int32),
                             ld("csr_val", np.float32).astype(np.complex64), ld("csc_ptr", np.int32),
                             ld("csc_ind", np.int32), ld("csc_val", np.float32).astype(np.complex64))
    return op


def _run_and_compare(name, kind, d, op, b, dims, maxit, profile=False):
    print(f"shim case {name} {kind} dims={dims} m={b.size} profile={profile}", flush=True)
    env = dict(os.environ, TVAL3_GPU_PROFILE="1" if profile else "0")
    r = subprocess.run([_driver(name), kind, d], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 0, r.stderr
    # TVAL3_GPU_PROFILE=1: one slot table per solve in the reference's
    # PROFILING print format; it changes no result (checked below)
    lines = [l for l in r.stdout.splitlines() if "(runs: " in l]
    if profile:
        pat = re.compile(r"^.+:\t \d+\.\d\d ms \(runs: \d+, total: \d+\.\d ms\)$")
        assert lines and all(pat.match(l) for l in lines), r.stdout[-2000:]
        assert sum(l.startswith("gemv CblasNoTrans: A x:") for l in lines) >= 5
    else:
        assert not lines
    n = int(np.prod(dims))
    for tag, sd in (("f", 0), ("d", 1)):
        Ug = np.fromfile(os.path.join(d, f"U_{tag}.bin"), np.float32)
        assert Ug.size == n
        outer, total, _, _ = open(os.path.join(d, f"info_{tag}.txt")).read().split()
        opts = tval3_amd.default_options(maxit=maxit, ls_exact=1, scal_double=sd)
        Uo, info_o, _ = O.solve(op, b, dims, opts)
        rel = float(np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo))
        assert int(total) == info_o.total_iters and int(outer) == info_o.outer_iters
        assert rel <= 1e-6, (tag, rel)
    # the double/double overload against the oracle's double/double solve
    Ud = np.fromfile(os.path.join(d, "U_dd.bin"), np.float64)
    assert Ud.size == n
    outer, total, _, _ = open(os.path.join(d, "info_dd.txt")).read().split()
    opts = tval3_amd.default_options(maxit=maxit, ls_exact=1)
    Uo, info_o, _ = O.solve_f64(op, b, dims, opts)
    assert int(total) == info_o.total_iters and int(outer) == info_o.outer_iters
    assert F.rel(Ud, Uo) <= F.bound(op, b, dims, opts, Uo), ("dd", F.rel(Ud, Uo))
    # the complex overloads on b rotated by 0.6 + 0.8i, imaginary part kept:
    # complex<float> as the float path (1e-6 here), complex<double> to the
    # FP64 bar
    cop = _complex_op(op, kind, d)
    opts = tval3_amd.default_options(maxit=maxit, ls_exact=1, isreal=0)
    for tag, t in (("f", np.complex64), ("d", np.complex128)):
        bc = (b.astype(np.float64) * (0.6 + 0.8j)).astype(t)
        Uc = np.fromfile(os.path.join(d, f"U_c{tag}.bin"), t)
        assert Uc.size == n
        outer, total, _, _ = open(os.path.join(d, f"info_c{tag}.txt")).read().split()
        Uo, info_o, _ = O.solve_c(cop, bc, dims, opts)
        assert int(total) == info_o.total_iters and int(outer) == info_o.outer_iters, tag
        bar = 1e-6 if tag == "f" else F.bound(cop, bc, dims, opts, Uo, O.solve_c)
        assert F.rel(Uc, Uo) <= bar, ("c" + tag, F.rel(Uc, Uo))
        assert np.abs(Uc.imag).max() > 0


@pytest.mark.gpu
def test_shim_2d_dense(tmp_path):
    q = p = 24
    m = 230
    A = synth.gaussian_A(m, q * p, seed=3)
    x = synth.ellipse_phantom(p, q, 6, seed=3)
    b = synth.add_noise(A @ x, 0.01, seed=3)
    d = str(tmp_path / "dense2d")
    _write(d, {"dx": q, "dy": p, "m": m, "n": q * p, "maxit": 80}, A=A, b=b)
    _run_and_compare("shim_2d", "dense", d, O.Op.dense(A), b, (q, p, 1), 80)


@pytest.mark.gpu
def test_shim_2d_dense_layers(tmp_path):
    """Multi-layer call: b holds three measurement vectors of one A, U the
    three images stacked along y (layers = b.len / A.dim_y)."""
    q = p = 16
    L, m = 3, 100
    A = synth.gaussian_A(m, q * p, seed=4)
    x = np.concatenate([synth.ellipse_phantom(p, q, 4, seed=4 + l).ravel() for l in range(L)])
    b = np.concatenate([synth.add_noise(A @ x[l * q * p:(l + 1) * q * p], 0.01, seed=4 + l)
                        for l in range(L)]).astype(np.float32)
    d = str(tmp_path / "dense2d_layers")
    _write(d, {"dx": q, "dy": L * p, "m": m, "n": q * p, "maxit": 60}, A=A, b=b)
    _run_and_compare("shim_2d", "dense", d, O.Op.dense(A), b, (q, L * p, 1), 60)


@pytest.mark.gpu
def test_shim_2d_sparse(tmp_path):
    G = 32
    geo = synth.ring_geometry(G, G)
    ptr, lin = O.Op.geometry(geo).walk()
    m, n = geo.m, geo.n
    rp, ri, rv, cp, ci, cv = synth.csr_from_walk(ptr, lin, n)
    op = O.Op.sparse(m, n, rp, ri, rv, cp, ci, cv)
    x = synth.disk_phantom(G, seed=1)
    b = synth.add_noise(op.gemv(x), 0.005, seed=1)
    d = str(tmp_path / "sparse2d")
    _write(d, {"dx": G, "dy": G, "m": m, "n": n, "maxit": 60},
           b=b, csr_ptr=rp.astype(np.int32), csr_ind=ri.astype(np.int32),
           csr_val=rv.astype(np.float32), csc_ptr=cp.astype(np.int32),
           csc_ind=ci.astype(np.int32), csc_val=cv.astype(np.float32))
    _run_and_compare("shim_2d", "sparse", d, op, b, (G, G, 1), 60)


@pytest.mark.gpu
def test_shim_2d_geometry(tmp_path):
    # the 2-D geometry layout: ray e * R + r for every pair (no compaction)
    G, K = 32, 32
    xe, ye = synth.ring_transducers(K, G)
    E = R = K
    used = np.array([[1 if e < r else 0 for r in range(R)] for e in range(E)], np.int32).ravel()
    up = np.where(used != 0, np.arange(E * R, dtype=np.int32) + 1, 0).astype(np.int32)
    z = np.zeros(K, np.int32)
    geo = synth.GeometryArrays(1, up, xe, ye, z, xe, ye, z, (G, G, 1), E * R)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(G, seed=2)
    b = synth.add_noise(op.gemv(x), 0.005, seed=2) * (used != 0)
    b = b.astype(np.float32)
    d = str(tmp_path / "geo2d")
    _write(d, {"dx": G, "dy": G, "ld": G, "scale": 1.0, "maxit": 60},
           b=b, use_path=used, xe=xe, ye=ye, xr=xe, yr=ye)
    _run_and_compare("shim_2d", "geometry", d, op, b, (G, G, 1), 60)


@pytest.mark.gpu
def test_shim_3d_geometry(tmp_path):
    grid, K = 16, 40
    geo = synth.hemisphere_geometry(K, grid)
    op = O.Op.geometry(geo)
    x = synth.ellipsoid_phantom(grid, 4, seed=4)
    b = synth.add_noise(op.gemv(x), 0.005, seed=4)
    d = str(tmp_path / "geo3d")
    _write(d, {"dx": grid, "dy": grid, "dz": grid, "mps": geo.num_mps, "E": geo.E,
               "R": geo.R, "paths": geo.num_paths, "rv_x": geo.rv[0], "rv_y": geo.rv[1],
               "rv_z": geo.rv[2], "scale": geo.scale, "maxit": 40},
           b=b, use_path=geo.use_path, xe=geo.xe, ye=geo.ye, ze=geo.ze,
           xr=geo.xr, yr=geo.yr, zr=geo.zr)
    _run_and_compare("shim_3d", "geometry", d, op, b, (grid, grid, grid), 40, profile=True)


@pytest.mark.gpu
def test_shim_3d_dense(tmp_path):
    q, p, r = 8, 8, 4
    n = q * p * r
    m = 100
    A = synth.gaussian_A(m, n, seed=5)
    x = np.random.default_rng(5).random(n).astype(np.float32)
    b = synth.add_noise(A @ x, 0.01, seed=5)
    d = str(tmp_path / "dense3d")
    _write(d, {"dx": q, "dy": p, "dz": r, "m": m, "n": n, "maxit": 50}, A=A, b=b)
    _run_and_compare("shim_3d", "dense", d, O.Op.dense(A), b, (q, p, r), 50)
