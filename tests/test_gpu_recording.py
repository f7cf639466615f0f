"""The recording workload on the device: the Karlqvist head tables of mrm_amd.recording
moved by FLD_VELOCITY (applied.f90:884-903, set/interpolate 1471-1522, 1757-1777), the
per-bit polarity flips of an NRZ script, and the readback of the written track.

* The grain field under the moving head equals the mean over the grain's cells of the
  closed-form Karlqvist field at the head's position at the evaluation time (bilinear
  interpolation of the table: within 5e-3 of the peak field), and a polarity flip negates
  it exactly.
* Parity: a moving head table with demag, exchange and thermal noise on a small grid
  matches the oracle to 1e-10 in m.
* Full track (512 x 4096 x 4 cells of 1.5 nm, FP32): a shortened NRZ write runs bit by bit
  through mrmg_set_source without rebuilding the context; directions stay unit vectors,
  the track ahead of the head keeps its initial magnetisation (the head field's tail there
  only tilts it), and the readback there has the sign of that magnetisation.
"""
import numpy as np
import pytest

from mrm_amd import abi, media, recording
from mrm_amd.llg import Llg
from tests.test_gpu_parity import pair, random_state, say

pytestmark = pytest.mark.gpu


def grain_centres_cells(L, dx, dy):
    g2gi = L.grain2gridindex
    cells = L.grain2grid.reshape(-1, 2)
    ci = np.repeat(np.arange(L.ng), np.diff(g2gi))
    return ci, (cells[:, 0] - 0.5) * dx, (cells[:, 1] - 0.5) * dy


def test_head_field_is_karlqvist_at_the_moving_gap():
    nx, ny, nl, d = 32, 256, 4, 1.5e-7
    med = media.config_media("cfg2", nx=nx, ny=ny, exchange=False, ku=0.0, layer_ku=[0.0] * 4, thermal=False,
                             temperature=0.0)
    x = Llg(nx, ny, nl, d, d, med.dz, demag=False)
    x.set_media(med)
    y0 = 0.5 * ny * d
    x.set_llg(dt=5e-14)
    hg = recording.deep_gap_field(15000.0, 20e-7, 5e-7)
    depths = recording.layer_depths(list(med.dz), np.zeros(nl), 5e-7)
    for pol in (1.0, -1.0):
        for k, s in enumerate(recording.head_sources(nx, ny, d, d, med.dz, start=y0, polarity=pol)):
            x.set_source(k, 0, s)
        x.step(1, t0=0.0)                      # the stored field is stage 2's, at t = dt
        yg = y0 + 2000.0 * 5e-14
        worst = 0.0
        for k, L in enumerate(med.layers):
            _, h = x.get_state(k)
            ci, cx, cy = grain_centres_cells(L, d, d)
            hl, hp = recording.karlqvist(cy - yg, depths[k], 20e-7, hg)
            cnt = np.bincount(ci)
            want_y = pol * np.bincount(ci, weights=hl) / cnt
            want_z = -pol * np.bincount(ci, weights=hp) / cnt
            worst = max(worst, np.abs(h[:, 1] - want_y).max(), np.abs(h[:, 2] - want_z).max(), np.abs(h[:, 0]).max())
        say(f"head field polarity {pol:+.0f}: max deviation from Karlqvist {worst:.2f} Oe")
        assert worst <= 5e-3 * 15000.0
    x.close()


def test_head_table_parity_with_oracle(oracle_mod):
    O = oracle_mod
    med = media.config_media("cfg2", nx=48, ny=96, seed=4)
    s, gpu = pair(O, med)
    srcs = recording.head_sources(med.nx, med.ny, med.dx, med.dy, med.dz, start=20e-7)
    for k, src in enumerate(srcs):
        s.set_source(k, 0, src)
        gpu.set_source(k, 0, src)
    for k, m in enumerate(random_state(med, 3, tilt=0.2)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n = 40
    noise = np.random.default_rng(9).standard_normal((n, med.G, 3))
    gpu.set_llg(dt=5e-14)
    s.step(n, 5e-14, noise=noise)
    gpu.step(n, noise=noise)
    dm = max(float(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max()) for k in range(med.n_lay))
    say(f"moving head table vs oracle: max|dm| {dm:.2e}")
    assert dm <= 1e-10
    gpu.close()


def test_nrz_write_on_the_full_track():
    med = media.config_media("cfg2", temperature=0.0, thermal=False)
    x = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=abi.MRMG_FP32)
    x.build_tensor(num_images=2)
    x.set_media(med)
    bits = recording.nrz_bits(4, seed=7)
    y0 = 0.25 * med.ny * med.dy
    say(f"nrz write on {med.nx}x{med.ny}x{med.n_lay}: bits {bits.tolist()}")
    # a shortened write: 1.5 nm bits at 20 m/s, dt 50 fs (15 steps per bit)
    t = recording.write_nrz(x, bits, 1.5e-7, 2000.0, 5e-14,
                            lambda pol: recording.head_sources(med.nx, med.ny, med.dx, med.dy, med.dz, start=y0,
                                                               polarity=pol))
    assert t > 0
    ahead = []
    for k, L in enumerate(med.layers):
        m, _ = x.get_state(k)
        assert np.abs(np.linalg.norm(m, axis=1) - 1.0).max() < 1e-12
        ci, cx, cy = grain_centres_cells(L, med.dx, med.dy)
        yc = np.bincount(ci, weights=cy) / np.bincount(ci)
        far = yc > y0 + 2000.0 * t + 1.5e-5
        ahead.append(m[far, 2].min())
    say(f"  min m_z 150 nm ahead of the head: {min(ahead):.4f}")
    assert min(ahead) > 0.9
    # readback far ahead of the head, where the media is still uniform +z
    ys, v = recording.readback_scan(x, y0 + 3e-5, y0 + 3.5e-5, 1e-7, 0.5 * med.nx * med.dx)
    assert np.all(np.isfinite(v)) and v.min() > 0
    x.close()
