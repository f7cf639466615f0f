"""The oracle against vectors written by the reference's own Fortran modules
(random.f90, tensor_element.f90 compiled with flang; tests/golden/make_golden.py).  These
pins hold bit for bit without the reference tree."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


def _load():
    return np.load(GOLDEN, allow_pickle=False)


def test_mt_uniform_stream(oracle_mod):
    g = _load()
    for i, seed in enumerate(g["mt_seeds"]):
        want = g[f"mt_uniform_{i}"]
        assert np.array_equal(oracle_mod.MT(int(seed)).uniform(len(want)), want), int(seed)


def test_mt_polar_normals(oracle_mod):
    g = _load()
    want = g["mt_normal3d"]
    assert np.array_equal(oracle_mod.MT(int(g["mt_normal3d_seed"][0])).normal3d(len(want)), want)


def test_volume_averaged_n(oracle_mod):
    g = _load()
    for a, n_ref, v_ref in zip(g["van_args"], g["van_n"], g["van_vol"]):
        n, vol = oracle_mod.volume_averaged_n(*[float(x) for x in a])
        assert vol == v_ref
        assert np.array_equal(n, n_ref.reshape(3, 3).T)


def test_plane_to_plane(oracle_mod):
    g = _load()
    for i, (nx, ny, dz, dzo, rz, img) in enumerate(g["p2p_cases"]):
        nx, ny, img = int(nx), int(ny), int(img)
        p = oracle_mod.plane_to_plane(nx, ny, 2e-7, 2e-7, float(dz), float(dzo), float(rz), img, img)
        assert np.array_equal(p, g[f"p2p_{i}"]), i
