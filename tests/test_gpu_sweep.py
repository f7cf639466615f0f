"""A seeded sweep of configurations against the oracle: grid sizes with factors 2, 3 and 5
(odd ones included), one to four layers of different thickness and spacing, 0-2 periodic
images, every push type, with and without exchange and thermal noise (injected MT
normals), on each column path and lanes-per-grain shape.  Each case runs 12 FP64 steps on
the device and in the oracle from the same start and must agree to 1e-10 in m."""
import numpy as np
import pytest

from mrm_amd import abi, media
from tests.test_gpu_parity import pair, random_state, say

pytestmark = pytest.mark.gpu

SIZES = [16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 90, 96]
N_CASES = 16


def case(i):
    rng = np.random.default_rng(4000 + i)
    nx, ny = (int(v) for v in rng.choice(SIZES, 2))
    nl = int(rng.integers(1, 5))
    dz = [float(v) for v in rng.choice([2e-7, 3e-7, 4e-7], nl)]
    il = [0.0] + [float(v) for v in rng.choice([0.0, 5e-8, 1e-7], nl - 1)]
    return dict(nx=nx, ny=ny, nl=nl, dz=dz, il=il, pitch=float(rng.choice([6e-7, 8e-7])),
                images=int(rng.integers(0, 3)), ltype=int(rng.choice([-3, -2, -1, 1, 2])),
                thermal=bool(rng.random() < 0.6), exchange=bool(rng.random() < 0.8),
                ku=[float(v) for v in rng.uniform(2e6, 8e6, nl)], cone=float(rng.uniform(0.0, 20.0)),
                alpha=float(rng.choice([0.02, 0.1, 0.5])), path=int(rng.integers(0, 3)),
                lpg=int(rng.choice([4, 8, 16])), tilt=float(rng.uniform(5.0, 30.0)), seed=int(rng.integers(1, 1000)))


@pytest.mark.parametrize("i", range(N_CASES))
def test_parity_sweep(oracle_mod, i):
    O = oracle_mod
    c = case(i)
    say(f"sweep case {i}: {c}")
    med = media.make_media(c["nx"], c["ny"], dz=c["dz"], n_lay=c["nl"], pitch=c["pitch"], seed=c["seed"],
                           layer_ku=c["ku"], cone_deg=c["cone"], alpha=c["alpha"], exchange=c["exchange"],
                           thermal=c["thermal"], temperature=300.0 if c["thermal"] else 0.0)
    tensor = O.calculate_tensor(c["nx"], c["ny"], c["nl"], med.dx, med.dy, med.dz, il_space=c["il"],
                                num_images=c["images"])
    uf = media.ramp_field(c["nl"], tilt_deg=c["tilt"], h_max=1.5e4, t_ramp=5e-12)
    s, gpu = pair(O, med, tensor=tensor, ufields=uf, il=c["il"])
    try:
        gpu.tune("col_path", c["path"])
    except abi.MrmgError:
        pass                      # the grid does not fit that path; the default one runs
    gpu.tune("lanes_per_grain", c["lpg"])
    for k, m in enumerate(random_state(med, c["seed"], tilt=0.3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n, dt = 12, 5e-14
    noise = None
    if c["thermal"]:
        noise = O.mt_noise(list(range(21, 21 + c["nl"])), [L.ng for L in med.layers], [True] * c["nl"], n)
    gpu.set_llg(dt=dt, llg_type=c["ltype"])
    s.step(n, dt, type=c["ltype"], noise=noise)
    gpu.step(n, noise=noise)
    d = max(float(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max()) for k in range(c["nl"]))
    say(f"  path {gpu.get_tune('col_path')}, {med.G} grains: max|dm| {d:.2e}")
    assert d <= 1e-10


def llb_case(i):
    rng = np.random.default_rng(5000 + i)
    nl = int(rng.integers(1, 4))
    return dict(nx=int(rng.choice(SIZES)), ny=int(rng.choice(SIZES)), nl=nl,
                dz=[float(v) for v in rng.choice([2e-7, 4e-7], nl)], ltype=int(rng.choice([-3, -2, -1, 1, 2])),
                tc=float(rng.uniform(550.0, 750.0)), alpha=float(rng.choice([0.05, 0.1, 0.3])),
                t_lo=float(rng.uniform(200.0, 400.0)), seed=int(rng.integers(1, 1000)), path=int(rng.integers(0, 3)))


@pytest.mark.parametrize("i", range(8))
def test_llb_parity_sweep(oracle_mod, i):
    """LLB (push_llb) with grain temperatures spread below and above Tc, thermal noise and
    demag, random grids and push types: 12 steps within 1e-9 in m and |m| (the device's
    and host's tanh/sinh/exp differ in the last bits)."""
    O = oracle_mod
    c = llb_case(i)
    say(f"llb sweep case {i}: {c}")
    med = media.make_media(c["nx"], c["ny"], dz=c["dz"], n_lay=c["nl"], pitch=6e-7, seed=c["seed"], thermal=True,
                           temperature=450.0, tc_nom=c["tc"], tc_sigma=0.05, alpha=c["alpha"])
    rng = np.random.default_rng(c["seed"])
    for L in med.layers:
        L.temp = rng.uniform(c["t_lo"], c["tc"] * 1.1, L.ng)
    s, gpu = pair(O, med, ufields=media.ramp_field(c["nl"], h_max=2e4, t_ramp=1e-11))
    try:
        gpu.tune("col_path", c["path"])
    except abi.MrmgError:
        pass
    for k, m in enumerate(random_state(med, c["seed"], tilt=0.4)):
        mag = rng.uniform(0.3, 0.95, len(m))
        s.set_state(k, m)
        gpu.set_state(k, m)
        s.set_mag(k, mag)
        gpu.set_mag(k, mag)
    n, dt = 12, 1e-15
    noise = O.mt_noise(list(range(41, 41 + c["nl"])), [L.ng for L in med.layers], [True] * c["nl"], n)
    gpu.set_llg(dt=dt, llg_type=c["ltype"], llb=True)
    s.step(n, dt, type=c["ltype"], noise=noise, llb=True)
    gpu.step(n, noise=noise)
    dm = max(float(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max()) for k in range(c["nl"]))
    dmag = max(float(np.abs(s.get_mag(k) - gpu.get_mag(k)).max()) for k in range(c["nl"]))
    say(f"  {med.G} grains: max|dm| {dm:.2e} max|d|m|| {dmag:.2e}")
    assert dm <= 1e-9 and dmag <= 1e-9


@pytest.mark.parametrize("i", range(6))
def test_adaptive_parity_sweep(oracle_mod, i):
    """Adaptive dt with push_m's rejection on random grids and stacks: the same accepted
    steps, rejection count, dt and trajectory as the oracle over two calls."""
    O = oracle_mod
    rng = np.random.default_rng(6000 + i)
    nl = int(rng.integers(1, 4))
    nx, ny = (int(v) for v in rng.choice(SIZES, 2))
    max_dm = float(rng.choice([1.0, 3.0, 20.0]))
    dt_frac = float(rng.choice([0.05, 0.2]))
    med = media.make_media(nx, ny, dz=[3e-7] * nl, n_lay=nl, pitch=8e-7, seed=int(rng.integers(1, 999)),
                           thermal=True, temperature=300.0, alpha=float(rng.choice([0.02, 0.2])))
    say(f"adaptive sweep case {i}: {nx}x{ny}x{nl}, max_dm {max_dm}, dt_frac {dt_frac}")
    s, gpu = pair(O, med, ufields=media.ramp_field(nl, h_max=2e4, t_ramp=2e-11))
    for k, m in enumerate(random_state(med, 3 + i, tilt=0.3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n = 10
    noise = O.mt_noise(list(range(61, 61 + nl)), [L.ng for L in med.layers], [True] * nl, 2 * n)
    gpu.set_llg(dt=1e-14, dt_frac=dt_frac, dt_max=1e-12, max_dm=max_dm)
    t, dt, rej = 0.0, 1e-14, 0
    for ch in range(2):
        do = s.step(n, dt, t0=t, noise=noise[ch * n:(ch + 1) * n], dt_frac=dt_frac, dt_max=1e-12, max_dm=max_dm)
        dg = gpu.step(n, t0=t, noise=noise[ch * n:(ch + 1) * n])
        assert do.rejected == dg.rejected
        assert abs(do.dt - dg.dt) <= 1e-12 * do.dt and abs(do.t - dg.t) <= 1e-12 * do.t
        t, dt, rej = do.t, do.dt, rej + do.rejected
    d = max(float(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max()) for k in range(nl))
    say(f"  {med.G} grains, {rej} rejections: max|dm| {d:.2e}")
    assert d <= 1e-10
