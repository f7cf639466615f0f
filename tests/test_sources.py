"""Moving applied sources of the oracle (get_appl, applied.f90:827-1034; get_applied,
field.f90:545-556), pinned by what the mathematics fixes:

* bilinear table interpolation (set + interpolate) reproduces an affine field exactly
  inside the table and holds the edge values outside it, at the position the source
  has moved to (offset + vel (t - t_origin) - ref);
* the super-Gaussian temperature (set_sg + get_sg) is dtemp exp(-(u/sigma)^n) per axis;
* a grain sees the mean of its cells;
* a static spot drives every grain's LLB |m| to the mean-field fixed point at its own
  temperature, and the temperature scaling of Hk follows it.
"""
import numpy as np
import pytest

from mrm_amd import abi, media
from tests.test_llb import _fixed_point

NX, NY, DX, DY = 24, 20, 2e-7, 3e-7


def _sim(O, n_lay=1, demag=False, grains=None, **lay):
    g2g = np.arange(1, NX * NY + 1, dtype=np.int32).reshape(NY, NX) if grains is None else grains
    g2gi, cells, counts = media.grain_cells(g2g)
    ng = len(counts)
    L = abi.LayerArrays(ng=ng, grid2grain=g2g, grain2gridindex=g2gi, grain2grid=cells,
                        grain_vol=np.full(ng, 1e-19), ms=np.full(ng, 700.0), hk=np.zeros((ng, 2)), alpha=0.1,
                        gamma=1.76e7, **lay)
    s = O.Sim(abi.make_geom(NX, NY, n_lay, DX, DY, [2e-7] * n_lay, demag=demag))
    for k in range(n_lay):
        s.set_layer(k, L)
    return s, L, g2g


def _affine_table(nxp=7, nyp=6):
    xpos = np.cumsum(np.r_[0.0, np.linspace(0.6, 1.4, nxp - 1)]) * 1e-7 - 1e-7
    ypos = np.cumsum(np.r_[0.0, np.linspace(1.3, 0.7, nyp - 1)]) * 1.5e-7 - 2e-7
    X, Y = np.meshgrid(xpos, ypos)
    coef = np.array([[100.0, 3e8, -2e8], [-50.0, 1e8, 4e8], [800.0, -6e8, 1e8]])
    vals = np.stack([c[0] + c[1] * X + c[2] * Y for c in coef], axis=-1)   # (nyp, nxp, 3)
    return xpos, ypos, vals, coef


@pytest.mark.parametrize("vel,t", [((0.0, 0.0), 0.0), ((2000.0, 0.0), 1e-11), ((-500.0, 1500.0), 3e-11)])
def test_table_interpolation_affine_and_clamped(oracle_mod, vel, t):
    O = oracle_mod
    s, _, _ = _sim(O)
    xpos, ypos, vals, coef = _affine_table()
    off, ref, t0 = (1.1e-7, -0.4e-7), (0.3e-7, 0.2e-7), 2e-12
    src = abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos, ypos, vals, vel=vel, offset=off, ref=ref, t_origin=t0,
                           s_start=0.5, s_end=1.5, t_start=0.0, t_end=4e-11, scale=2.0)
    s.set_source(0, 0, src)
    got = s.source_cells(0, 0, t)
    sc = ((1.5 - 0.5) * min(1.0, max(0.0, t / 4e-11)) + 0.5) * 2.0
    # sampled position of cell centre i: (i - 1/2) dx - (offset + vel (t - t0) - ref)
    xs = (np.arange(NX) + 0.5) * DX - (off[0] + vel[0] * (t - t0) - ref[0])
    ys = (np.arange(NY) + 0.5) * DY - (off[1] + vel[1] * (t - t0) - ref[1])
    XS, YS = np.meshgrid(np.clip(xs, xpos[0], xpos[-1]), np.clip(ys, ypos[0], ypos[-1]))
    for c in range(3):
        want = sc * (coef[c][0] + coef[c][1] * XS + coef[c][2] * YS)
        np.testing.assert_allclose(got[c], want, rtol=1e-12, atol=1e-9 * np.abs(want).max())
    inside = (xs > xpos[0]).any() and (xs < xpos[-1]).any()
    assert inside


def test_super_gaussian_closed_form(oracle_mod):
    O = oracle_mod
    s, _, _ = _sim(O)
    sig, n, dT = (6e-7, 9e-7), 4, 350.0
    off, ref, vel, t0, t = (24e-7, 30e-7), (1e-7, -2e-7), (0.0, 1500.0), 0.0, 2e-11
    src = abi.SourceArrays(abi.MRMG_SRC_TEMP_SG, vel=vel, offset=off, ref=ref, t_origin=t0, sigma=sig, expn=n,
                           dtemp=dT)
    s.set_source(0, 1, src)
    got = s.source_cells(0, 1, t)[0]
    ux = (np.arange(NX) + 0.5) * DX + (ref[0] - vel[0] * (t - t0) - off[0])
    uy = (np.arange(NY) + 0.5) * DY + (ref[1] - vel[1] * (t - t0) - off[1])
    want = dT * np.outer(np.exp(-(uy / sig[1]) ** n), np.exp(-(ux / sig[0]) ** n))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * dT)


def test_grain_field_is_cell_mean(oracle_mod):
    """Grains of several cells: after one step the field of a grain (no demag, no
    anisotropy, no exchange) is the mean of the source over its cells at t + dt plus the
    uniform field."""
    O = oracle_mod
    g2g = media.voronoi_grid(NX, NY, DX, DY, 6e-7, 3, boundary=True)
    s, L, _ = _sim(O, grains=g2g)
    xpos, ypos, vals, _ = _affine_table()
    vals = vals + 40.0 * np.sin(np.arange(vals.size).reshape(vals.shape))    # not affine any more
    s.set_source(0, 0, abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos, ypos, vals, vel=(1000.0, 0.0),
                                        offset=(1e-7, 0.0)))
    s.set_ufield(0, abi.make_ufield((0.0, 0.6, 0.8), scale=500.0))
    dt = 1e-13
    s.step(1, dt)
    h = s.get_state(0)[1]
    cells = s.source_cells(0, 0, dt)
    want = np.zeros((L.ng, 3))
    for n in range(L.ng):
        sel = g2g == n + 1
        want[n] = cells[:, sel].mean(axis=1)
    want += np.array([0.0, 0.6, 0.8]) * 500.0
    np.testing.assert_allclose(h, want, rtol=1e-12, atol=1e-9)


def test_static_spot_llb_fixed_points(oracle_mod):
    """A super-Gaussian spot over a 300 K background: every grain relaxes to the LLB
    mean-field |m| of its own temperature (Tc 700 K, so the spot centre is above Tc)."""
    O = oracle_mod
    Tc, S, H = 700.0, 0.5, 1e4
    s, L, _ = _sim(O, temp=np.full(NX * NY, 300.0), tc=np.full(NX * NY, Tc), spin=S, thermal=False)
    s.set_ufield(0, abi.make_ufield((0, 0, 1), scale=H))
    spot = abi.SourceArrays(abi.MRMG_SRC_TEMP_SG, offset=(NX * DX / 2, NY * DY / 2), sigma=(8e-7, 8e-7), expn=2,
                            dtemp=600.0)
    s.set_source(0, 1, spot)
    s.set_state(0, np.tile([0.0, 0.0, 1.0], (NX * NY, 1)))
    s.step(6000, 1e-15, llb=True)
    T = 300.0 + s.source_cells(0, 1, 0.0)[0].ravel()
    want = np.array([_fixed_point(H, t, Tc, S) for t in T])
    far = np.abs(T - Tc) > 0.05 * Tc          # critical slowing down within 5 % of Tc
    np.testing.assert_allclose(s.get_mag(0)[far], want[far], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(s.get_mag(0)[~far], want[~far], rtol=2e-2)
    assert T.max() > 1.05 * Tc and T.min() < Tc


def test_temperature_source_scales_hk(oracle_mod):
    """hkexp with a moving temperature: the uniaxial field of each grain scales with
    (1 - T/Tc)^hkexp at the grain's current temperature."""
    O = oracle_mod
    Tc, hk, e = 700.0, 1e4, 1.5
    n = NX * NY
    s, L, _ = _sim(O, temp=np.full(n, 250.0), tc=np.full(n, Tc), tc_nom=Tc, hkexp=e, uni_axis=np.tile([0, 0, 1.0], (n, 1)))
    L2 = abi.LayerArrays(**{**{f: getattr(L, f) for f in abi.LayerArrays.FIELDS_D + abi.LayerArrays.FIELDS_I},
                            "ng": n, "hk": np.tile([hk, 0.0], (n, 1)), "alpha": 0.1, "tc_nom": Tc, "hkexp": e,
                            "gamma": 1.76e7})
    s.set_layer(0, L2)
    spot = abi.SourceArrays(abi.MRMG_SRC_TEMP_SG, offset=(10e-7, 30e-7), vel=(3000.0, 0.0), sigma=(6e-7, 9e-7),
                            expn=2, dtemp=300.0)
    s.set_source(0, 1, spot)
    m = np.tile([0.6, 0.0, 0.8], (n, 1))
    s.set_state(0, m)
    dt = 1e-16
    s.step(1, dt)
    h = s.get_state(0)[1]
    T = 250.0 + s.source_cells(0, 1, dt)[0].ravel()
    # the corrector's field is at the predicted m, a hair away from m: compare z parts
    want = hk * np.clip(1 - T / Tc, 0, 1) ** e * 0.8
    np.testing.assert_allclose(h[:, 2], want, rtol=1e-6)


def _tables():
    x = np.array([0.0, 0.3, 0.55, 0.8, 0.95, 1.1])
    return {"ms": (x, np.array([1.0, 0.92, 0.75, 0.45, 0.2, 0.0])),
            "alpha": (x, np.array([1.0, 1.1, 1.3, 1.8, 2.5, 3.0])),
            "hk": (x[:4], np.array([1.0, 0.7, 0.4, 0.1]))}


def test_scale_tables_static(oracle_mod):
    """READ_SCALE tables (temp_scaling.f90:397-416) at a static temperature: each scaling is
    the piecewise-linear table at T/Tc, held at the end values outside (numpy.interp)."""
    O = oracle_mod
    n = NX * NY
    rng = np.random.default_rng(4)
    T = rng.uniform(0.0, 900.0, n)
    tc = np.full(n, 700.0)
    tabs = _tables()
    s, L, _ = _sim(O, temp=T, tc=tc, tc_nom=700.0, tables=tabs)
    ms, al, hk = s.get_scales(0)
    x = T / tc
    np.testing.assert_allclose(ms, np.interp(x, *tabs["ms"]), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(al, np.interp(x, *tabs["alpha"]), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(hk, np.interp(x, *tabs["hk"]), rtol=1e-14, atol=1e-15)
    # exact node hits and both ends
    T2 = np.concatenate([tabs["ms"][0] * 700.0, [-50.0, 5000.0]])
    T2 = np.resize(T2, n)
    s2, _, _ = _sim(O, temp=T2, tc=tc, tc_nom=700.0, tables=tabs)
    np.testing.assert_allclose(s2.get_scales(0)[0], np.interp(T2 / 700.0, *tabs["ms"]), rtol=1e-15, atol=0)


def test_scale_tables_follow_moving_temperature(oracle_mod):
    """Under a moving spot every field evaluation re-evaluates the tables at the grain
    temperature; power laws take precedence for Ms and Hk when their exponent is >= 0."""
    O = oracle_mod
    n = NX * NY
    tc = np.full(n, 700.0)
    tabs = _tables()
    s, L, _ = _sim(O, temp=np.full(n, 300.0), tc=tc, tc_nom=700.0, tables=tabs, hkexp=1.5)
    spot = abi.SourceArrays(abi.MRMG_SRC_TEMP_SG, offset=(10e-7, 30e-7), vel=(3000.0, 0.0), sigma=(6e-7, 9e-7),
                            expn=2, dtemp=500.0)
    s.set_source(0, 1, spot)
    s.set_state(0, np.tile([0.0, 0.0, 1.0], (n, 1)))
    dt = 1e-15
    s.step(1, dt)
    T = 300.0 + s.source_cells(0, 1, dt)[0].ravel()     # temperature of the corrector's evaluation
    ms, al, hk = s.get_scales(0)
    np.testing.assert_allclose(ms, np.interp(T / 700.0, *tabs["ms"]), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(al, np.interp(T / 700.0, *tabs["alpha"]), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(hk, np.clip(1 - T / 700.0, 0, 1) ** 1.5, rtol=1e-13, atol=1e-15)
    assert T.max() > 700.0 and T.min() < 400.0
