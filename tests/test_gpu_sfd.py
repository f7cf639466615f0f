"""Switching-field statistics of the thermal two-layer media (T = 300 K, Heun, dt = 50 fs)
on the device against the oracle, in distribution.

The oracle's samples (tests/golden/sfd_cfg1_64.npz, written by tests/golden/make_sfd.py)
come from the reference's own generator (MT19937 per layer, injected); the device draws
its thermal field from its Philox stream, so single trajectories differ and only the
distributions can agree.  A grain's switching field is the ramp field when its m_z first
turns negative, read from the device's switching record and rounded up to the oracle's
20-step check points; grains that never switch during the ramp are censored (+inf for the
test).  The two-sample Kolmogorov-Smirnov test compares one value per grain column (the
mean over its layers, the layers of a column switching together), pooled over the 10
seeds.  Acceptance: p >= 1e-3 against the oracle's sample.  Power: the same device sample
must be rejected (p < 1e-3) against the oracle's planted sample, whose thermal variance is
doubled.  Both the FP64 form and the FP32 form are tested.
"""
import os

import numpy as np
import pytest
from scipy import stats

from mrm_amd import abi, media
from mrm_amd.llg import Llg
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sfd_cfg1_64.npz")
CENSORED = 1e9


def column_values(h, ngs):
    """(seeds, G) switching fields -> per-column means over the layers (censored -> +inf)."""
    h = np.where(np.isfinite(h), h, CENSORED)
    n = ngs[0]
    assert all(v == n for v in ngs)
    return h.reshape(h.shape[0], len(ngs), n).mean(axis=1).ravel()


def device_sfd(prec, seeds):
    f = np.load(FIX)
    nx, ny, mseed, t_ramp, h_max, dt, steps, chunk = f["protocol"]
    nx, ny, mseed, steps, chunk = int(nx), int(ny), int(mseed), int(steps), int(chunk)
    med = media.config_media("cfg1", nx=nx, ny=ny, seed=mseed)
    x = Llg(nx, ny, med.n_lay, med.dx, med.dy, med.dz, precision=prec)
    x.build_tensor(num_images=2)
    x.set_media(med)
    for k, u in enumerate(media.ramp_field(med.n_lay, t_ramp=t_ramp, h_max=h_max)):
        x.set_ufield(k, u)
    x.set_llg(dt=dt)
    out = np.full((len(seeds), med.G), np.nan)
    off = np.concatenate([[0], np.cumsum(med.ngs)])
    for i, sd in enumerate(seeds):
        for k in range(med.n_lay):
            x.set_state(k, np.tile([0.0, 0.0, 1.0], (med.ngs[k], 1)))
        x.step_index = 0
        x.switch_arm((0.0, 0.0, -1.0), 0.0)
        x.step(steps, philox_seed=1000 + sd, diag=False)
        for k in range(med.n_lay):
            st, _, _ = x.switch_record(k)
            c = st // chunk
            v = h_max * np.minimum((c + 1) * chunk * dt / t_ramp, 1.0)
            out[i, off[k]:off[k + 1]] = np.where(st >= 0, v, np.nan)
    x.close()
    return out, med.ngs


@pytest.mark.parametrize("prec", [abi.MRMG_FP64, abi.MRMG_FP32], ids=["fp64", "fp32"])
def test_sfd_matches_oracle_in_distribution(prec):
    f = np.load(FIX)
    h, ngs = device_sfd(prec, list(f["seeds"]))
    assert list(ngs) == list(f["ngs"])
    dev = column_values(h, ngs)
    ora = column_values(f["sfd"], ngs)
    bad = column_values(f["planted"], ngs)
    p = stats.ks_2samp(dev, ora).pvalue
    p_bad = stats.ks_2samp(dev, bad).pvalue
    sw_d = np.isfinite(h).mean()
    sw_o = np.isfinite(f["sfd"]).mean()
    fin = np.isfinite(h)
    say(f"SFD {'fp64' if prec == abi.MRMG_FP64 else 'fp32'}: device switched {sw_d:.3f} "
        f"(oracle {sw_o:.3f}), mean {h[fin].mean():.0f} Oe (oracle {np.nanmean(f['sfd']):.0f}); "
        f"KS p {p:.3g}, against the doubled-variance sample p {p_bad:.3g}")
    assert p >= 1e-3
    assert p_bad < 1e-3
