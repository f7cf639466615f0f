"""Pins of two parts of the oracle that its deterministic tests leave open.

push_m's step control (llg.f90:1195-1271): a macrospin precessing without damping
(alpha = 0) about a constant field H perpendicular to m.  The field does not depend on m,
so every quantity push_m sees follows from the geometry of one Heun step of pure
precession by the angle phi = gamma H dtt:

  d1 = phi e1,  m1 = (m + d1)/sqrt(1 + phi^2) (rotated by a = atan(phi)),  d2 = phi e1',
  d = (d1 + d2)/2:  |d|^2 = phi^2 (1 + cos a)/2,
  the step's rotation psi = acos(1 - |d|^2 / 2), its precession angle
  beta = atan2(phi (1 + cos a)/2, 1 - phi sin a / 2).

With these the rejection count, the accepted step and the time follow from the rules the
reference states: reject when psi/max_dm > 1 and retry with max_dmdt * 0.9/rot_ratio
(llg.f90:1229-1236); after an accepted step keep max_dmdt when rot_ratio >= 2/3, else take
min(2, max_dmdt/rot_ratio) (1244-1253); get_fields' dt = fact1/max(1,|h|) with
fact1 = (1+a^2)/((1+a) gamma) dt_frac (384-386), so phi = dt_frac * max_dmdt here.

The thermal LLB push (llg.f90:955-979 Heun, 1010-1023 rotational) is pinned in
test_llb_thermal_equilibrium.py.
"""
import math

import numpy as np

from mrm_amd import abi
from tests.test_oracle import _single_grain

GAMMA = 1.76e7


def psi(phi):
    a = math.atan(phi)
    return math.acos(1.0 - 0.25 * phi * phi * (1.0 + math.cos(a)))


def beta(phi):
    a = math.atan(phi)
    return math.atan2(0.5 * phi * (1.0 + math.cos(a)), 1.0 - 0.5 * phi * math.sin(a))


def _precessing(O, H):
    g, L = _single_grain(alpha=0.0)
    s = O.Sim(g)
    s.set_layer(0, L)
    s.set_ufield(0, abi.make_ufield((0, 0, 1), scale=H))
    s.set_state(0, np.array([[1.0, 0.0, 0.0]]))
    return s


def test_push_m_rejects_once_and_shrinks_by_point_nine(oracle_mod):
    """phi0 = 0.5 against max_dm = 10 deg: psi(0.5) = 28 deg, so the first attempt is
    rejected once and retried at max_dmdt = 0.9 max_dm/psi(phi0); that step turns by
    psi(phi1) in [2/3, 1] max_dm, so max_dmdt is kept and no later step of the command is
    rejected.  llg() resets max_dmdt to 1 when it starts (llg.f90:206), so the next command
    is rejected once again."""
    H, dt_frac, max_dm = 1.0e4, 0.5, math.radians(10.0)
    s = _precessing(oracle_mod, H)
    dt0 = dt_frac / (GAMMA * H)
    rot0 = psi(dt_frac) / max_dm
    assert rot0 > 1.0
    k1 = 0.9 / rot0
    phi1 = dt_frac * k1
    assert 2.0 / 3.0 <= psi(phi1) / max_dm <= 1.0
    d = s.step(6, 1e-14, t0=0.0, dt_frac=dt_frac, dt_max=1.0, max_dm=math.degrees(max_dm))
    assert d.rejected == 1
    assert abs(d.dt / (dt0 * k1) - 1.0) < 1e-13
    assert abs(d.max_dmdt / k1 - 1.0) < 1e-13
    assert abs(d.t / (6 * dt0 * k1) - 1.0) < 1e-13
    b = beta(phi1)
    np.testing.assert_allclose(s.get_state(0)[0][0], [math.cos(6 * b), math.sin(6 * b), 0.0], atol=1e-12)
    d2 = s.step(2, d.dt, t0=d.t, dt_frac=dt_frac, dt_max=1.0, max_dm=math.degrees(max_dm))
    assert d2.rejected == 1
    assert abs(d2.t / (8 * dt0 * k1) - 1.0) < 1e-13
    np.testing.assert_allclose(s.get_state(0)[0][0], [math.cos(8 * b), math.sin(8 * b), 0.0], atol=1e-12)


def test_push_m_doubles_small_steps_up_to_two(oracle_mod):
    """rot_ratio 0.26 < 2/3 after the first step: max_dmdt becomes min(2, 1/0.26) = 2, the
    next steps turn by psi(2 phi0) < 2/3 max_dm and max_dmdt stays at the cap of 2."""
    H, dt_frac, max_dm = 2.0e4, 0.04, math.radians(9.0)
    s = _precessing(oracle_mod, H)
    dt0 = dt_frac / (GAMMA * H)
    assert psi(dt_frac) / max_dm < 2.0 / 3.0 and psi(2 * dt_frac) / max_dm < 2.0 / 3.0
    d = s.step(4, 1e-14, t0=0.0, dt_frac=dt_frac, dt_max=1.0, max_dm=math.degrees(max_dm))
    assert d.rejected == 0
    assert abs(d.t / (7 * dt0) - 1.0) < 1e-13      # dt0 + 2 dt0 + 2 dt0 + 2 dt0
    assert d.max_dmdt == 2.0
    b = beta(dt_frac) + 3 * beta(2 * dt_frac)
    np.testing.assert_allclose(s.get_state(0)[0][0], [math.cos(b), math.sin(b), 0.0], atol=1e-12)


def test_push_m_keeps_steps_between_two_thirds_and_one(oracle_mod):
    """rot_ratio 0.82: accepted, and max_dmdt stays 1; the time is n dt0."""
    H, dt_frac, max_dm = 5.0e3, 0.1, math.radians(7.0)
    s = _precessing(oracle_mod, H)
    dt0 = dt_frac / (GAMMA * H)
    assert 2.0 / 3.0 <= psi(dt_frac) / max_dm <= 1.0
    d = s.step(6, 1e-14, t0=0.0, dt_frac=dt_frac, dt_max=1.0, max_dm=math.degrees(max_dm))
    assert d.rejected == 0 and d.max_dmdt == 1.0
    assert abs(d.t / (6 * dt0) - 1.0) < 1e-13
    b = 6 * beta(dt_frac)
    np.testing.assert_allclose(s.get_state(0)[0][0], [math.cos(b), math.sin(b), 0.0], atol=1e-12)
