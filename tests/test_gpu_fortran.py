"""The Fortran binding end to end on the device: mrm_amd/fortran/mrmg_drive (built by
__graft_entry__.build() with flang against mrmg_m.f90 and libmrmg.so) runs one LLG
command through the iso_c_binding interfaces, with the reference's own array layouts,
then the same command under a moving write-head field table (mrmg_set_source, the
FLD_VELOCITY table of applied.f90:884-903) with a launch shape set by mrmg_tune; the same
runs made through the Python binding give the same m bit for bit."""
import math
import os
import subprocess

import numpy as np
import pytest

from mrm_amd import abi, media
from mrm_amd.llg import Llg
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mrm_amd", "fortran", "mrmg_drive")


def head_table():
    """The driver's table: nodes xp (5), yp (4) and the field (yp, xp, 3) = Fortran (3,nxp,nyp)."""
    xp = np.array([-2e-6 + 1.2e-6 * i for i in range(5)])
    yp = np.array([-1e-6 + 1.5e-6 * j for j in range(4)])
    tab = np.zeros((4, 5, 3))
    for j in range(4):
        for i in range(5):
            tab[j, i] = (300.0 * (i - 2), 50.0 * (j - 1) * i, -2e3 - 500.0 * i + 100.0 * j * j)
    return xp, yp, tab


def python_run(source=False):
    nx = ny = 16
    dx = dy = 2e-7
    dz = 4e-7
    jj, ii = np.mgrid[0:ny, 0:nx]
    g2g = ((jj // 2) * (nx // 2) + ii // 2 + 1).astype(np.int32)      # grid2grain (ny, nx)
    g2gi, cells, counts = media.grain_cells(g2g)
    ng = len(counts)
    n = np.arange(1, ng + 1)
    uni = np.stack([0.01 * (n % 5), 0.02 * (n % 3), np.ones(ng)], axis=1)
    uni = uni / np.sqrt(uni[:, 0] * uni[:, 0] + uni[:, 1] * uni[:, 1] + uni[:, 2] * uni[:, 2])[:, None]
    m = np.tile([0.1, 0.0, 0.995], (ng, 1))
    m = m / np.sqrt(m[:, 0] * m[:, 0] + m[:, 1] * m[:, 1] + m[:, 2] * m[:, 2])[:, None]
    la = abi.LayerArrays(ng=ng, grid2grain=g2g, grain2gridindex=g2gi, grain2grid=cells,
                         grain_vol=np.full(ng, 4.0 * dx * dy * dz), ms=np.full(ng, 700.0),
                         hk=np.stack([1e4 + 100.0 * n, np.zeros(ng)], axis=1), uni_axis=uni,
                         alpha=0.1, gamma=1.76e7, thm_alpha=1.0, floor_ms=0.01, thermal=False)
    x = Llg(nx, ny, 1, dx, dy, [dz])
    x.build_tensor(num_images=1)
    x.set_layer(0, la)
    d = np.array([0.1, 0.0, -1.0]) / math.sqrt(1.01)
    x.set_ufield(0, abi.make_ufield(d, t_start=0.0, t_end=1e-10, s_start=0.0, s_end=1e4))
    if source:
        xp, yp, tab = head_table()
        x.set_source(0, 0, abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos=xp, ypos=yp, vals=tab, vel=(0.0, 2000.0),
                                            t_start=0.0, t_end=2e-12, s_start=0.0, s_end=1.0))
        x.tune("lanes_per_grain", 4)
    x.set_state(0, m)
    x.set_llg(dt=5e-14, llg_type=abi.MRMG_TYPE_HEUN)
    x.step(50)
    out = x.get_state(0)[0]
    x.close()
    return out


def test_fortran_driver_matches_python_binding(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("the Fortran driver is not built (no flang where build() ran)")
    say("fortran driver")
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    say("  " + r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    for name, source in (("m_fortran.bin", False), ("m_fortran_src.bin", True)):
        mf = np.fromfile(tmp_path / name, dtype="<f8").reshape(-1, 3)   # Fortran m(3,ng)
        mp = python_run(source)
        assert mf.shape == mp.shape
        say(f"  {name}: max|dm| Fortran vs Python binding: {np.abs(mf - mp).max():.2e}")
        assert np.array_equal(mf, mp)
        assert np.abs(mp[:, 2] - mp[0, 2]).max() > 0      # the grains moved differently
    # the head table changed the run
    assert not np.array_equal(python_run(False), python_run(True))
