"""GPU parity of the moving applied sources (get_appl, applied.f90:827-1034; get_applied,
field.f90:545-556) and of everything a moving temperature drives: Hk and Ms scaling
(temp_scaling.f90:365-431), the thermal variance (field.f90:407-423) and the LLB push."""
import numpy as np
import pytest

from mrm_amd import abi, media
from tests.test_gpu_llb import llb_media, start, compare
from tests.test_gpu_parity import pair, random_state, say

pytestmark = pytest.mark.gpu


def head_table(med, peak=1.5e4):
    """A head-like field table: a gap field along x under a pole, on its own node mesh."""
    xpos = np.linspace(-40e-7, 40e-7, 33)
    ypos = np.linspace(-10e-7, med.ny * med.dy + 10e-7, 9)
    X, Y = np.meshgrid(xpos, ypos)
    g = 20e-7
    hx = peak / np.pi * (np.arctan((X + g / 2) / 5e-7) - np.arctan((X - g / 2) / 5e-7))
    hz = peak / (2 * np.pi) * np.log(((X + g / 2) ** 2 + 25e-14) / ((X - g / 2) ** 2 + 25e-14))
    vals = np.stack([hx, 0.05 * hx * np.sin(Y / 7e-7), hz], axis=-1)
    return abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos, ypos, vals, vel=(2000.0, 0.0), offset=(10e-7, 0.0),
                            ref=(0.0, 0.0), t_origin=0.0, t_start=0.0, t_end=2e-12, s_start=0.0, s_end=1.0)


def laser(med, dtemp=450.0, vel=(0.0, 1500.0)):
    return abi.SourceArrays(abi.MRMG_SRC_TEMP_SG, offset=(med.nx * med.dx / 2, med.ny * med.dy / 3), vel=vel,
                            sigma=(12e-7, 12e-7), expn=2, dtemp=dtemp)


def both(s, gpu, k, slot, src):
    s.set_source(k, slot, src)
    gpu.set_source(k, slot, src)


@pytest.mark.parametrize("ltype", [-3, -2, 1])
def test_moving_field_table_parity(oracle_mod, ltype):
    O = oracle_mod
    med = media.make_media(64, 48, dz=[4e-7, 4e-7], n_lay=2, pitch=8e-7, seed=31, thermal=True, temperature=300.0,
                           layer_ku=[2e6, 1e7])
    s, gpu = pair(O, med)
    for k in range(2):
        both(s, gpu, k, 0, head_table(med))
    for k, m in enumerate(random_state(med, 2, tilt=0.3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n = 30
    noise = O.mt_noise([3, 4], [L.ng for L in med.layers], [True, True], n)
    gpu.set_llg(dt=5e-14, llg_type=ltype)
    s.step(n, 5e-14, type=ltype, noise=noise)
    gpu.step(n, noise=noise)
    d = max(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max() for k in range(2))
    dh = max(np.abs(s.get_state(k)[1] - gpu.get_state(k)[1]).max() for k in range(2))
    say(f"moving field table type {ltype}: |dm| {d:.2e} |dh| {dh:.2e}")
    assert d <= 1e-10 and dh <= 1e-7


@pytest.mark.parametrize("ltype", [-3, -2])
def test_laser_llb_parity(oracle_mod, ltype):
    """LLB under a moving super-Gaussian spot over 300 K (through Tc), with Ms from |m|
    and Hk ~ (1-T/Tc)^2."""
    O = oracle_mod
    med = llb_media(nx=64, ny=48, seed=41, use_mag=True, hkexp=2.0)
    for L in med.layers:
        L.temp = np.full(L.ng, 300.0)
    s, gpu = pair(O, med, ufields=[abi.make_ufield((0, 0, -1), scale=8000.0)] * med.n_lay)
    for k in range(med.n_lay):
        both(s, gpu, k, 1, laser(med))
    start(s, gpu, med, seed=5)
    n = 30
    noise = O.mt_noise([7, 8], [L.ng for L in med.layers], [True] * med.n_lay, n)
    gpu.set_llg(dt=1e-15, llg_type=ltype, llb=True)
    s.step(n, 1e-15, type=ltype, noise=noise, llb=True)
    gpu.step(n, noise=noise)
    compare(s, gpu, med, 1e-9, f"laser llb type {ltype}")


def test_temperature_table_llg_ms_power_law(oracle_mod):
    """LLG with Ms ~ (1-T/Tc)^msexp under a moving temperature table: the cells' Ms, the
    exchange weights and the thermal variance follow the temperature."""
    O = oracle_mod
    med = llb_media(nx=64, ny=48, seed=43, msexp=1.0, hkexp=1.5, alpha=0.05)
    for L in med.layers:
        L.temp = np.full(L.ng, 280.0)
    xpos = np.linspace(-30e-7, 30e-7, 13)
    ypos = np.linspace(-30e-7, 30e-7, 11)
    X, Y = np.meshgrid(xpos, ypos)
    tab = abi.SourceArrays(abi.MRMG_SRC_TEMP_TABLE, xpos, ypos, 300.0 * np.exp(-(X ** 2 + Y ** 2) / 2e-12),
                           vel=(-1000.0, 500.0), offset=(60e-7, 40e-7))
    s, gpu = pair(O, med, ufields=media.ramp_field(med.n_lay, h_max=1.5e4, t_ramp=1e-11))
    for k in range(med.n_lay):
        both(s, gpu, k, 1, tab)
    start(s, gpu, med, seed=6)
    for k in range(med.n_lay):          # LLG: |m| = 1
        s.set_mag(k, np.ones(med.layers[k].ng))
        gpu.set_mag(k, np.ones(med.layers[k].ng))
    n = 30
    noise = O.mt_noise([9, 10], [L.ng for L in med.layers], [True] * med.n_lay, n)
    gpu.set_llg(dt=5e-14)
    s.step(n, 5e-14, noise=noise)
    gpu.step(n, noise=noise)
    d = max(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max() for k in range(med.n_lay))
    say(f"temperature table llg msexp: |dm| {d:.2e}")
    assert d <= 1e-10


def test_laser_llb_adaptive_parity(oracle_mod):
    O = oracle_mod
    med = llb_media(nx=48, ny=40, seed=47, use_mag=True)
    for L in med.layers:
        L.temp = np.full(L.ng, 300.0)
    s, gpu = pair(O, med, ufields=[abi.make_ufield((0, 0, -1), scale=8000.0)] * med.n_lay)
    for k in range(med.n_lay):
        both(s, gpu, k, 1, laser(med, vel=(1500.0, 0.0)))
    start(s, gpu, med, seed=8)
    n = 20
    noise = O.mt_noise([11, 12], [L.ng for L in med.layers], [True] * med.n_lay, n)
    gpu.set_llg(dt=1e-15, llb=True, dt_frac=0.05, dt_max=4e-15, max_dm=2.0)
    d0 = s.step(n, 1e-15, noise=noise, llb=True, dt_frac=0.05, dt_max=4e-15, max_dm=2.0)
    d1 = gpu.step(n, noise=noise)
    say(f"adaptive laser llb: rejected {d0.rejected}/{d1.rejected}")
    assert d0.rejected == d1.rejected
    assert abs(d0.t - d1.t) <= 1e-12 * d0.t
    compare(s, gpu, med, 1e-9, "adaptive laser llb")


def test_source_argument_checks():
    med = media.make_media(16, 16, dz=[4e-7], n_lay=1, pitch=6e-7, seed=1)
    from mrm_amd.llg import Llg
    x = Llg(16, 16, 1, med.dx, med.dy, med.dz, demag=False)
    x.set_media(med)
    bad = abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, [0.0, 1e-7, 0.5e-7], [0.0, 1e-7], np.zeros((2, 3, 3)))
    with pytest.raises(abi.MrmgError):
        x.set_source(0, 0, bad)                 # nodes not increasing
    with pytest.raises(abi.MrmgError):
        x.set_source(0, 1, head_table(med))     # a field table in the temperature slot


@pytest.mark.parametrize("ltype", [-3, 2])
def test_time_functions_parity(oracle_mod, ltype):
    """Time functions (funct, applied.f90:328-396) on the device: a uniform field under the
    NEW pulse train, a moving head table under SIN and a laser under the STE wave."""
    O = oracle_mod
    med = llb_media(nx=48, ny=40, seed=51, hkexp=1.5, tc_nom=700.0)
    for L in med.layers:
        L.temp = np.full(L.ng, 300.0)
    period = 1.5e-12
    uf = abi.make_ufield((0.1, 0.0, -1.0), scale=6000.0, func=abi.MRMG_FN_NEW, amp=1.0, gamma=period, param=0.4e-12,
                         duty=0.4, moddepth=0.7)
    s, gpu = pair(O, med, ufields=[uf] * med.n_lay)
    head = head_table(med)
    head.kw.update(func=abi.MRMG_FN_SIN, amp=0.8, gamma=4e12, phase=0.3, dc=0.2)
    spot = laser(med, dtemp=300.0)
    spot.kw.update(func=abi.MRMG_FN_STE, amp=1.0, gamma=period, param=0.5e-12, dc=1.0)
    for k in range(med.n_lay):
        both(s, gpu, k, 0, head)
        both(s, gpu, k, 1, spot)
    for k, m in enumerate(random_state(med, 9, tilt=0.3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n = 40
    noise = O.mt_noise([61, 62], [L.ng for L in med.layers], [True] * med.n_lay, n)
    gpu.set_llg(dt=5e-14, llg_type=ltype)
    s.step(n, 5e-14, type=ltype, noise=noise)
    gpu.step(n, noise=noise)
    d = max(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max() for k in range(med.n_lay))
    dh = max(np.abs(s.get_state(k)[1] - gpu.get_state(k)[1]).max() for k in range(med.n_lay))
    say(f"time functions type {ltype}: |dm| {d:.2e} |dh| {dh:.2e}")
    assert d <= 1e-10 and dh <= 1e-7


def test_time_function_argument_checks():
    from mrm_amd.llg import Llg
    med = media.make_media(16, 16, dz=[4e-7], n_lay=1, pitch=6e-7, seed=3)
    x = Llg(med.nx, med.ny, 1, med.dx, med.dy, med.dz)
    try:
        with pytest.raises(abi.MrmgError):
            x.set_ufield(0, abi.make_ufield((0, 0, 1), func=9, amp=1.0, gamma=1.0))
        with pytest.raises(abi.MrmgError):
            x.set_ufield(0, abi.make_ufield((0, 0, 1), func=abi.MRMG_FN_STE, amp=1.0, gamma=1e-9, param=0.0))
        x.set_ufield(0, abi.make_ufield((0, 0, 1), t_start=0.0, t_end=0.0, func=abi.MRMG_FN_SIN, amp=1.0, gamma=1e9))
    finally:
        x.close()


def _scale_tables():
    x = np.array([0.0, 0.3, 0.55, 0.8, 0.95, 1.1])
    return {"ms": (x, np.array([1.0, 0.92, 0.75, 0.45, 0.2, 0.0])),
            "alpha": (x, np.array([1.0, 1.1, 1.3, 1.8, 2.5, 3.0])),
            "hk": (x[:4], np.array([1.0, 0.7, 0.4, 0.1]))}


@pytest.mark.parametrize("ltype,moving", [(-3, True), (2, True), (-3, False)])
def test_scale_tables_parity(oracle_mod, ltype, moving):
    """READ_SCALE tables for Ms, alpha and Hk (temp_scaling.f90:397-416): at a static
    temperature (set once) and under a moving spot (every evaluation: Hk and alpha at the
    evaluation's temperature, the cells' Ms one stage ahead), thermal LLG."""
    O = oracle_mod
    med = llb_media(nx=48, ny=40, seed=71, tc_nom=700.0, temperature=300.0)
    for L in med.layers:
        L.tables = {k: (np.ascontiguousarray(v[0]), np.ascontiguousarray(v[1])) for k, v in _scale_tables().items()}
        if not moving:
            L.temp = np.random.default_rng(5).uniform(250.0, 800.0, L.ng)
        else:
            L.temp = np.full(L.ng, 300.0)
    s, gpu = pair(O, med, ufields=[abi.make_ufield((0.1, 0.0, -1.0), scale=4000.0)] * med.n_lay)
    if moving:
        for k in range(med.n_lay):
            both(s, gpu, k, 1, laser(med, dtemp=450.0))
    for k, m in enumerate(random_state(med, 12, tilt=0.3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n = 40
    noise = O.mt_noise([71, 72], [L.ng for L in med.layers], [True] * med.n_lay, n)
    gpu.set_llg(dt=5e-14, llg_type=ltype)
    s.step(n, 5e-14, type=ltype, noise=noise)
    gpu.step(n, noise=noise)
    d = max(np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max() for k in range(med.n_lay))
    dh = max(np.abs(s.get_state(k)[1] - gpu.get_state(k)[1]).max() for k in range(med.n_lay))
    say(f"scale tables type {ltype} moving {moving}: |dm| {d:.2e} |dh| {dh:.2e}")
    assert d <= 1e-10 and dh <= 1e-7
