"""The oracle's switching-field sample (tests/golden/sfd_cfg1_64.npz, written by
tests/golden/make_sfd.py) that the device's distribution test compares against.

* Its protocol is the one make_sfd.py states, and every switching field is a check-point
  value of the ramp (0 < H <= 20 kOe, a multiple of the 40 Oe chunk field).
* The distribution test has power at this sample size: the planted sample (thermal
  variance doubled) is rejected by the two-sample KS test on the column means at
  p < 1e-10, while disjoint halves of the oracle's own seeds are not (p > 1e-3).
* Hotter noise switches more grains and at lower fields (the direction thermal activation
  fixes).
"""
import os
import sys

import numpy as np
from scipy import stats

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_sfd  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "sfd_cfg1_64.npz"))


def columns(h):
    h = np.where(np.isfinite(h), h, 1e9)
    return h.reshape(h.shape[0], 4, -1).mean(axis=1).ravel()


def test_protocol_and_values():
    M = make_sfd
    assert list(FIX["protocol"]) == [M.NX, M.NY, M.MEDIA_SEED, M.T_RAMP, M.H_MAX, M.DT, M.STEPS, M.CHUNK]
    assert list(FIX["seeds"]) == M.SEEDS
    assert FIX["sfd"].shape == FIX["planted"].shape == (len(M.SEEDS), int(FIX["ngs"].sum()))
    step = M.chunk_field(0)
    for a in (FIX["sfd"], FIX["planted"]):
        v = a[np.isfinite(a)]
        assert v.min() > 0 and v.max() <= M.H_MAX
        assert np.allclose(v / step, np.rint(v / step), atol=1e-9)


def test_distribution_test_has_power():
    a, b = FIX["sfd"], FIX["planted"]
    assert stats.ks_2samp(columns(a), columns(b)).pvalue < 1e-10
    assert stats.ks_2samp(columns(a[:5]), columns(a[5:])).pvalue > 1e-3
    assert stats.ks_2samp(columns(a[::2]), columns(a[1::2])).pvalue > 1e-3


def test_hotter_noise_switches_earlier():
    a, b = FIX["sfd"], FIX["planted"]
    assert np.isfinite(b).mean() > np.isfinite(a).mean()
    assert np.nanmedian(b) < np.nanmedian(a)
