"""Writes tests/golden/reference_vectors.npz from the reference's own Fortran modules
(random.f90, tensor_element.f90), compiled with flang into oracle/_ref/libmrmref.so by
`make -C oracle ref` in a tree where the reference sources are present.  The vectors let
tests/test_golden.py pin the oracle's MT19937 stream, polar normals and Newell tensor
elements bit for bit where the reference tree (and so oracle/_ref) is absent.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O   # noqa: E402

DP = C.POINTER(C.c_double)

MT_SEEDS = (34, 0, -7, 123456)
MT_N = 600
NORMAL_SEED, NORMAL_N = 35, 300
VAN_ARGS = (
    (0, 0, 0, 2e-7, 2e-7, 8e-7, 2e-7, 2e-7, 8e-7),
    (4e-7, -2e-7, 3e-7, 2e-7, 2e-7, 4e-7, 2e-7, 2e-7, 2e-7),
    (1e-5, 3e-6, -2e-7, 2e-7, 2e-7, 2e-7, 2e-7, 2e-7, 2e-7),
    (0.0, -1.5e-7, 0.0, 1.5e-7, 1.5e-7, 1.5e-7, 1.5e-7, 1.5e-7, 1.5e-7),
)
P2P_CASES = ((8, 8, 4e-7, 4e-7, 0.0, 2), (10, 9, 2e-7, 2e-7, -4e-7, 0), (6, 5, 4e-7, 2e-7, 3e-7, 1))


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` where the reference tree is present")
    R = O.ref()
    out = {}
    for i, seed in enumerate(MT_SEEDS):
        R.ref_random_create(i + 1, seed)
        u = np.zeros(MT_N)
        R.ref_random_uniform(i + 1, MT_N, u.ctypes.data_as(DP))
        out[f"mt_uniform_{i}"] = u
    out["mt_seeds"] = np.array(MT_SEEDS, np.int64)
    R.ref_random_create(9, NORMAL_SEED)
    nrm = np.zeros((NORMAL_N, 3))
    R.ref_random_normal3d(9, NORMAL_N, nrm.ctypes.data_as(DP))
    out["mt_normal3d"] = nrm
    out["mt_normal3d_seed"] = np.array([NORMAL_SEED], np.int64)
    van_n, van_v = [], []
    for a in VAN_ARGS:
        nr = np.zeros(9)
        van_v.append(R.ref_volume_averaged_n(*a, nr.ctypes.data_as(DP)))
        van_n.append(nr)
    out["van_args"] = np.array(VAN_ARGS, np.float64)
    out["van_n"] = np.array(van_n)          # Fortran n(3,3) in memory order
    out["van_vol"] = np.array(van_v)
    out["p2p_cases"] = np.array(P2P_CASES, np.float64)
    for i, (nx, ny, dz, dzo, rz, img) in enumerate(P2P_CASES):
        pr = np.zeros(9 * nx * ny)
        R.ref_plane_to_plane(nx, ny, 2e-7, 2e-7, dz, dzo, rz, img, img, pr.ctypes.data_as(DP))
        out[f"p2p_{i}"] = pr
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
