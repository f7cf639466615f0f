"""Readback and switching record on the device: outputs the reference does not compute
(its nearest relative, off_field.f90, only plots a stray-field plane).  They are pinned
by closed forms: the point response of one magnetic cell, the integral of the Gaussian
sensitivity over uniform magnetisation, the Stoner-Wohlfarth switching field that the
recorded switching fields approach as the field ramp slows; and, for the record's
bookkeeping, by the oracle's trajectory stepped one step at a time."""
import numpy as np
import pytest

from mrm_amd import abi, media
from mrm_amd.llg import Llg
from tests.test_gpu_parity import random_state, say

pytestmark = pytest.mark.gpu


def cell_layer(g2g, dx, dy, dz, ms=650.0):
    """A layer whose grains are given by a grid2grain map (0 = non-magnetic cell)."""
    g2gi, cells, counts = media.grain_cells(g2g)
    ng = len(counts)
    return abi.LayerArrays(ng=ng, grid2grain=g2g, grain2gridindex=g2gi, grain2grid=cells,
                           grain_vol=counts * dx * dy * dz, ms=np.full(ng, ms), hk=np.zeros((ng, 2)),
                           alpha=0.1, thermal=False, temp=np.zeros(ng))


def periodic_gauss(c, x0, L, s):
    d = c - x0
    d = d - L * np.rint(d / L)
    return np.exp(-d * d / (2.0 * s * s))


def test_readback_point_response():
    """One magnetic cell: the readback is Ms (dir . m) dV times the two Gaussians at the
    periodic distance from the cell centre; at the centre it is Ms (dir . m) dV itself."""
    nx, ny, dx, dy, dz = 32, 24, 2e-7, 3e-7, 5e-7
    i0, j0 = 5, 17
    g2g = np.zeros((ny, nx), np.int32)
    g2g[j0, i0] = 1
    x = Llg(nx, ny, 1, dx, dy, [dz], demag=False)
    x.set_layer(0, cell_layer(g2g, dx, dy, dz))
    m = np.array([[0.6, 0.0, 0.8]])
    x.set_state(0, m)
    xc, yc = (i0 + 0.5) * dx, (j0 + 0.5) * dy
    sx, sy = 3.1e-7, 4.5e-7
    xs = np.concatenate([np.linspace(-2e-7, nx * dx + 2e-7, 41), [xc, xc - nx * dx / 2]])
    ys = np.concatenate([np.full(41, yc + 2e-7), [yc, yc - ny * dy + 1e-8]])
    for d, comp in (((0.0, 0.0, 1.0), 0.8), ((1.0, 0.0, 0.0), 0.6), ((0.0, 0.0, -2.0), -1.6)):
        got = x.readback(xs, ys, (sx, sy), direction=d, weights=[0.5])
        want = 650.0 * comp * 0.5 * dx * dy * dz * periodic_gauss(xc, xs, nx * dx, sx) * periodic_gauss(yc, ys, ny * dy,
                                                                                                       sy)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-30)
        assert abs(got[41] / (650.0 * comp * 0.5 * dx * dy * dz) - 1.0) < 1e-15      # at the centre
    # the peak of a down-track scan sits on the cell centre
    scan = x.readback(np.arange(nx) * dx + 0.5 * dx, yc, (sx, sy))
    assert int(np.argmax(scan)) == i0
    x.close()


@pytest.mark.parametrize("precision", [abi.MRMG_FP64, abi.MRMG_FP32])
def test_readback_uniform_magnetisation(precision):
    """Uniform M in every cell of two layers: the readback is Ms sum_k w_k dz_k 2 pi sx sy
    (Poisson summation: the discrete sum of the Gaussians equals the integral up to
    exp(-2 pi^2 s^2 / d^2), and the periodic tails are below exp(-(L/2)^2 / (2 s^2)))."""
    nx, ny, dx, dy = 32, 24, 2e-7, 2e-7
    dz = [4e-7, 2e-7]
    g2g = (np.arange(nx * ny, dtype=np.int32).reshape(ny, nx) // 3 + 1).astype(np.int32)
    x = Llg(nx, ny, 2, dx, dy, dz, precision=precision, demag=False)
    for k in range(2):
        x.set_layer(k, cell_layer(g2g, dx, dy, dz[k], ms=700.0))
        x.set_state(k, np.tile([0.0, 0.0, 1.0], (int(g2g.max()), 1)))
    sx, sy = 1.5 * dx, 1.25 * dy
    rng = np.random.default_rng(2)
    xs, ys = rng.uniform(0, nx * dx, 7), rng.uniform(0, ny * dy, 7)
    got = x.readback(xs, ys, (sx, sy), weights=[1.0, 0.5])
    want = 700.0 * (dz[0] + 0.5 * dz[1]) * 2 * np.pi * sx * sy
    tol = 1e-12 if precision == abi.MRMG_FP64 else 1e-6
    say(f"uniform readback precision {precision}: max rel err {np.abs(got / want - 1).max():.2e}")
    np.testing.assert_allclose(got, want, rtol=tol)
    x.close()


def test_readback_random_state_direct_sum():
    """A thermal-free Voronoi media in a random state: the readback equals the direct sum
    over cells of w dz dx dy (dir . Ms m) gx gy."""
    med = media.make_media(48, 40, dz=[4e-7, 4e-7], n_lay=2, pitch=8e-7, seed=9)
    x = Llg(med.nx, med.ny, 2, med.dx, med.dy, med.dz, demag=False)
    x.set_media(med)
    ms = random_state(med, 4)
    for k, m in enumerate(ms):
        x.set_state(k, m)
    sx, sy = 5e-7, 8e-7
    d = np.array([0.3, -0.2, 0.9])
    w = [1.0, 0.35]
    xs = np.linspace(0, med.nx * med.dx, 50)
    ys = np.full(50, 0.4 * med.ny * med.dy)
    got = x.readback(xs, ys, (sx, sy), direction=d, weights=w)
    ci = (np.arange(med.nx) + 0.5) * med.dx
    cj = (np.arange(med.ny) + 0.5) * med.dy
    want = np.zeros(50)
    for k, L in enumerate(med.layers):
        g = L.grid2grain.reshape(med.ny, med.nx)
        proj = np.where(g > 0, (ms[k] @ d)[np.maximum(g, 1) - 1] * L.ms[np.maximum(g, 1) - 1], 0.0)
        for p in range(50):
            gx = periodic_gauss(ci, xs[p], med.nx * med.dx, sx)
            gy = periodic_gauss(cj, ys[p], med.ny * med.dy, sy)
            want[p] += w[k] * med.dz[k] * med.dx * med.dy * (gy[:, None] * gx[None, :] * proj).sum()
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())
    x.close()


def test_switch_record_matches_oracle_steps(oracle_mod):
    """The record's step index, time and applied field against the first crossing found
    by stepping the oracle one step at a time on the same media and ramp; the device run
    is split over several mrmg_step calls, so the step counter carries across calls."""
    O = oracle_mod
    med = media.make_media(48, 40, dz=[4e-7, 4e-7], n_lay=2, pitch=8e-7, seed=21, layer_ku=[2e6, 4e6], alpha=0.5)
    fd = (np.sin(np.radians(5.0)), 0.0, -np.cos(np.radians(5.0)))
    T, H, dt = 1.05e-10, 20000.0, 1e-13
    uf = [abi.make_ufield(fd, t_start=0.0, t_end=T, s_start=0.0, s_end=H) for _ in range(2)]
    g = abi.make_geom(med.nx, med.ny, 2, med.dx, med.dy, med.dz, demag=False)
    s = O.Sim(g)
    x = Llg(med.nx, med.ny, 2, med.dx, med.dy, med.dz, demag=False)
    m0 = random_state(med, 8, tilt=0.1)
    for k, L in enumerate(med.layers):
        for sim in (s, x):
            sim.set_layer(k, L)
            sim.set_ufield(k, uf[k])
            sim.set_state(k, m0[k])
    n = 1000
    x.set_llg(dt=dt, llg_type=-3)
    x.switch_arm((0.0, 0.0, -1.0), 0.0)
    t = 0.0
    for chunk in (1, 820, 179):      # the grains switch between steps 800 and 910
        t = x.step(chunk, t0=t).t
    rec = [np.full(L.ng, -1) for L in med.layers]
    rect = [np.zeros(L.ng) for L in med.layers]
    to = 0.0
    for it in range(n):
        s.step(1, dt, t0=to)
        to = to + dt
        for k in range(2):
            new = (rec[k] < 0) & (s.get_state(k)[0][:, 2] < 0)
            rec[k][new] = it
            rect[k][new] = to
    assert to == t
    total = 0
    for k in range(2):
        step, tt, h = x.switch_record(k)
        total += int((step >= 0).sum())
        assert np.array_equal(step, rec[k]), k
        on = step >= 0
        assert np.array_equal(tt[on], rect[k][on])
        np.testing.assert_allclose(h[on], (H * tt[on] / T)[:, None] * np.array(fd)[None, :], rtol=1e-13, atol=1e-9)
    say(f"switch record: {total} of {sum(L.ng for L in med.layers)} grains marked, identical to the oracle's steps")
    assert total > 150
    # stopping leaves the record as it is
    x.switch_stop()
    before = x.switch_record(0)[0].copy()
    x.step(50, t0=t)
    assert np.array_equal(before, x.switch_record(0)[0])
    x.close()


def stoner_wohlfarth(hk, psi):
    return hk / (np.cos(psi) ** (2.0 / 3.0) + np.sin(psi) ** (2.0 / 3.0)) ** 1.5


def test_switching_fields_approach_stoner_wohlfarth():
    """Non-interacting grains (no exchange, no demag, T = 0, alpha = 1) with easy axes in a
    20-degree cone under a field 5 degrees off -z ramped linearly: every recorded switching
    field lies above the grain's Stoner-Wohlfarth field, the excess shrinks as ramp
    rate^(2/3) (critical slowing down at the saddle-node) from a 4x slower ramp, and the
    extrapolation to zero ramp rate recovers the astroid within 0.5 %."""
    med = media.make_media(32, 32, n_lay=1, pitch=8e-7, seed=3, exchange=False, ku=6e6, ku_sigma=0.1,
                           cone_deg=20.0, alpha=1.0, thermal=False)
    L = med.layers[0]
    psi = np.radians(5.0)
    fd = np.array([np.sin(psi), 0.0, -np.cos(psi)])
    ax = L.uni_axis.reshape(-1, 3)
    hsw = stoner_wohlfarth(L.hk.reshape(-1, 2)[:, 0], np.arccos(np.clip(np.abs(ax @ -fd), 0, 1)))
    ratios = []
    for T in (2e-9, 8e-9):
        x = Llg(med.nx, med.ny, 1, med.dx, med.dy, med.dz, demag=False)
        x.set_layer(0, L)
        x.set_ufield(0, abi.make_ufield(fd, t_start=0.0, t_end=T, s_start=0.0, s_end=20000.0))
        x.set_state(0, ax.copy())
        x.set_llg(dt=1e-13, llg_type=-3)
        x.switch_arm((0.0, 0.0, -1.0), 0.0)
        x.step(int(round(T / 1e-13)))
        step, t, h = x.switch_record(0)
        x.close()
        assert (step >= 0).all()
        r = np.linalg.norm(h, axis=1) / hsw
        say(f"ramp {T:.0e} s: recorded/SW field min {r.min():.4f} mean {r.mean():.4f} max {r.max():.4f}")
        assert (r > 1.0).all()
        ratios.append(r)
    lag_fast, lag_slow = ratios[0] - 1.0, ratios[1] - 1.0
    q = 4.0 ** (-2.0 / 3.0)
    assert abs(np.median(lag_slow / lag_fast) - q) < 0.05
    extrap = ratios[1] - (ratios[0] - ratios[1]) * q / (1.0 - q)
    say(f"extrapolated to zero ramp rate: {extrap.min():.4f} .. {extrap.max():.4f} of the SW field")
    assert np.abs(extrap - 1.0).max() < 5e-3


def test_readout_argument_checks():
    nx, ny, dx, dy, dz = 16, 16, 2e-7, 2e-7, 4e-7
    g2g = np.ones((ny, nx), np.int32)
    x = Llg(nx, ny, 1, dx, dy, [dz], demag=False)
    x.set_layer(0, cell_layer(g2g, dx, dy, dz))
    x.set_state(0, np.array([[0.0, 0.0, 1.0]]))
    with pytest.raises(abi.MrmgError):
        x.readback([0.0], [0.0], (0.0, 1e-7))          # widths must be > 0
    with pytest.raises(ValueError):
        x.readback([0.0], [0.0], (1e-7, 1e-7), weights=[1.0, 2.0])
    assert x.readback(np.zeros(0), np.zeros(0), (1e-7, 1e-7)).size == 0
    with pytest.raises(abi.MrmgError):
        x.switch_record(0)                               # never armed
    # a new layer (more grains) ends the record: stepping marks nothing and reading refuses
    x.switch_arm()
    x.step(2)
    g2 = (np.arange(nx * ny, dtype=np.int32).reshape(ny, nx) + 1).astype(np.int32)
    x.set_layer(0, cell_layer(g2, dx, dy, dz))
    x.set_state(0, np.tile([0.0, 0.0, -1.0], (nx * ny, 1)))
    x.step(2)
    with pytest.raises(abi.MrmgError):
        x.switch_record(0)
    x.close()
