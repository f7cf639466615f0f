"""Time functions of applied fields and sources (funct, applied.f90:328-396; set_fld's
normalisation, applied.f90:276-310) in the oracle, pinned by closed forms: the plain
SIN/COS/EXP values, the STE trapezoid's levels, slopes and half-period antisymmetry, the
NEW pulse train's continuity at the duty edge and at the period wrap, and the field a
bare grain sees."""
import math

import numpy as np
import pytest

from mrm_amd import abi, media
from tests.test_sources import NX, NY, _affine_table, _sim


def test_sin_cos_exp(oracle_mod):
    O = oracle_mod
    for t in (0.0, 1e-10, 3.7e-10):
        g, ph = 2.0e10, 0.3
        assert O.funct(abi.MRMG_FN_SIN, t, g, ph) == math.sin(t * g + ph)
        assert O.funct(abi.MRMG_FN_COS, t, g, ph) == math.cos(t * g + ph)
        assert O.funct(abi.MRMG_FN_EXP, t, -1e9, ph) == math.exp(-t * 1e9 + ph)


def test_trapezoid_wave(oracle_mod):
    """STE: period T, ramps of width p: 0 -> 1 over [0, p/2], 1 until (T-p)/2, down to -1
    over p, -1 until T - p/2, back to 0; f(x + T/2) = -f(x); periodic, also for t < 0."""
    O = oracle_mod
    T, p = 4e-9, 0.8e-9
    f = lambda t: O.funct(abi.MRMG_FN_STE, t, T, 0.0, p)
    assert f(0.0) == 0.0
    assert abs(f(p / 4) - 0.5) < 1e-12
    assert f(p / 2 + 1e-12) == 1.0 and f(T / 4) == 1.0
    assert abs(f(T / 2)) < 1e-12
    assert f(3 * T / 4) == -1.0
    xs = np.linspace(0.0, T / 2, 97)[1:-1]
    for x in xs:
        assert abs(f(x + T / 2) + f(x)) < 1e-9
        assert abs(f(x + 3 * T) - f(x)) < 1e-9
        assert abs(f(x - 2 * T) - f(x)) < 1e-9
    # continuity: no jump larger than the slope allows between close samples
    ts = np.linspace(0, T, 4001)
    v = np.array([f(t) for t in ts])
    assert np.abs(np.diff(v)).max() <= 2.0 / p * (ts[1] - ts[0]) * 1.0001


@pytest.mark.parametrize("duty,md", [(0.3, 0.5), (0.5, 1.0), (0.8, 0.2)])
def test_pulse_train_continuity(oracle_mod, duty, md):
    """NEW: an exponential charge/discharge train of period T and time constant tau; it is
    continuous at the duty edge and across the period, periodic, and stays within
    [1 - 2 md, 1 + md]."""
    O = oracle_mod
    T, tau = 2e-9, 0.3e-9
    f = lambda t: O.funct(abi.MRMG_FN_NEW, t, T, 0.0, tau, duty, md)
    e = 1e-18
    assert abs(f(duty * T - e) - f(duty * T + e)) < 1e-7
    assert abs(f(T - e) - f(e)) < 1e-7
    a, y, w = T / tau, math.exp(duty * T / tau), math.exp(T / tau)
    assert abs(f(0.0) - (1 + (y - w) * md / (w - 1))) < 1e-14
    ts = np.linspace(0, 3 * T, 901)
    v = np.array([f(t) for t in ts])
    assert np.all(v <= 1 + md + 1e-12) and np.all(v >= 1 - 2 * md - 1e-12)
    for t in ts[:100]:
        assert abs(f(t + T) - f(t)) < 1e-9


def _bare(O):
    g2g = np.arange(1, NX * NY + 1, dtype=np.int32).reshape(NY, NX)
    s, L, _ = _sim(O, grains=g2g)
    return s


def test_uniform_field_time_function(oracle_mod):
    """A bare grain (no demag, anisotropy or exchange) sees (amp f(t) + dc) scale dir at the
    corrector's time t + dt."""
    O = oracle_mod
    s = _bare(O)
    amp, dc, sc, g, ph = 300.0, 50.0, 2.0, 3e10, 0.2
    s.set_ufield(0, abi.make_ufield((0.0, 0.6, 0.8), scale=sc, func=abi.MRMG_FN_SIN, amp=amp, gamma=g, phase=ph, dc=dc))
    s.set_state(0, np.tile([1.0, 0.0, 0.0], (NX * NY, 1)))
    dt, t0 = 1e-13, 2e-11
    s.step(1, dt, t0=t0)
    h = s.get_state(0)[1]
    want = (amp * math.sin((t0 + dt) * g + ph) + dc) * sc
    np.testing.assert_allclose(h[:, 1], 0.6 * want, rtol=1e-12)
    np.testing.assert_allclose(h[:, 2], 0.8 * want, rtol=1e-12)


def test_constant_function_normalisation(oracle_mod):
    """set_fld: amp == 0 or gamma == 0 turns a function into the constant amp f(0) + dc."""
    O = oracle_mod
    s = _bare(O)
    amp, dc, ph = 300.0, 50.0, 0.7
    s.set_ufield(0, abi.make_ufield((0.0, 0.0, 1.0), scale=3.0, func=abi.MRMG_FN_COS, amp=amp, gamma=0.0, phase=ph, dc=dc))
    s.set_state(0, np.tile([1.0, 0.0, 0.0], (NX * NY, 1)))
    s.step(1, 1e-13, t0=5e-10)
    h = s.get_state(0)[1]
    np.testing.assert_allclose(h[:, 2], 3.0 * (amp * math.cos(ph) + dc), rtol=1e-12)


def test_source_time_function(oracle_mod):
    """A moving field table scaled by the STE wave: the cell values are the table's affine
    field times (amp f(t) + dc) scale."""
    O = oracle_mod
    s, _, _ = _sim(O)
    xpos, ypos, vals, coef = _affine_table()
    T, p, amp, dc = 4e-11, 1e-11, 2.0, 0.5
    src = abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos, ypos, vals, scale=1.5, func=abi.MRMG_FN_STE, amp=amp,
                           gamma=T, param=p, dc=dc)
    s.set_source(0, 0, src)
    ref = abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos, ypos, vals)
    s2, _, _ = _sim(O)
    s2.set_source(0, 0, ref)
    for t in (0.0, 0.3e-11, 1.2e-11, 2.9e-11):
        f = O.funct(abi.MRMG_FN_STE, t, T, 0.0, p)
        np.testing.assert_allclose(s.source_cells(0, 0, t), (amp * f + dc) * 1.5 * s2.source_cells(0, 0, t),
                                   rtol=1e-12, atol=1e-9)
