"""GPU parity: the HIP path (through the C ABI of libmrmg.so) against the oracle on the
same inputs.  FP64 tolerances are written per test; the precision form is held to
max |dm| <= 1e-4 against the FP64 GPU path."""
import sys

import numpy as np
import pytest

from mrm_amd import abi, media
from mrm_amd.llg import Llg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    import ctypes as C
    lib = abi.load()
    n = C.c_int(0)
    if lib.mrmg_device_count(C.byref(n)) != 0 or n.value < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")


def pair(O, med, tensor=None, demag=True, precision=abi.MRMG_FP64, ufields=None, il=None):
    """An oracle Sim and a GPU Llg over the same media."""
    g = abi.make_geom(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, il, abi.MRMG_FP64, demag)
    s = O.Sim(g)
    gpu = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, il, precision=precision, demag=demag)
    if demag:
        if tensor is None:
            tensor = O.calculate_tensor(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, il_space=il, num_images=2)
        s.set_tensor(tensor)
        gpu.set_tensor(tensor)
    for k, L in enumerate(med.layers):
        s.set_layer(k, L)
        gpu.set_layer(k, L)
    for k, u in enumerate(ufields or []):
        s.set_ufield(k, u)
        gpu.set_ufield(k, u)
    return s, gpu


def check_loop_stats(do, dg, n_lay, tol=1e-10):
    """llg()'s loop-top statistics of the last step (llg.f90:217-237): oracle vs device."""
    def close(a, b, scale):
        return np.abs(np.array(a) - np.array(b)).max() <= tol * scale
    hs = max(1.0, float(np.abs(np.array(do.have_all)).max()), float(np.abs(np.array(do.have)[:n_lay]).max()))
    assert close(np.array(do.mave)[:n_lay], np.array(dg.mave)[:n_lay], 1.0)
    assert close(np.array(do.mvar)[:n_lay], np.array(dg.mvar)[:n_lay], max(1e-6, float(np.abs(np.array(do.mvar)).max())))
    assert close(np.array(do.have)[:n_lay], np.array(dg.have)[:n_lay], hs)
    assert close(do.mave_all, dg.mave_all, 1.0)
    assert close(do.have_all, dg.have_all, hs)
    assert close(do.mvar_all, dg.mvar_all, max(1e-6, float(np.abs(np.array(do.mvar_all)).max())))


def random_state(med, seed, tilt=None):
    rng = np.random.default_rng(seed)
    out = []
    for L in med.layers:
        if tilt is None:
            m = rng.standard_normal((L.ng, 3))
        else:
            m = np.tile([0.0, 0.0, 1.0], (L.ng, 1)) + tilt * rng.standard_normal((L.ng, 3))
        out.append(m / np.linalg.norm(m, axis=1, keepdims=True))
    return out


def max_dm(s, gpu, nl):
    d = 0.0
    for k in range(nl):
        d = max(d, np.abs(s.get_state(k)[0] - gpu.get_state(k)[0]).max())
    return d


def test_device_and_footprint(oracle_mod):
    med = media.config_media("cfg0")
    s, gpu = pair(oracle_mod, med)
    assert gpu.footprint() > 0


@pytest.mark.parametrize("nx,ny,nl", [(64, 64, 1), (48, 45, 3), (40, 30, 2), (20, 27, 1)])
def test_demag_field_parity(oracle_mod, nx, ny, nl):
    """Scatter + r2c + tensor product + c2r + grain average: FP64, |dH| <= 1e-12 max|H|."""
    say(f"demag parity {nx}x{ny}x{nl}")
    med = media.make_media(nx, ny, dz=[2e-7 + 1e-7 * k for k in range(nl)], n_lay=nl, pitch=6e-7, seed=nx + ny)
    il = [0.0] + [5e-8] * (nl - 1)
    s, gpu = pair(oracle_mod, med, il=il)
    for k, m in enumerate(random_state(med, 3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    hc_o, hg_o = s.demag_field()
    hc_g, hg_g = gpu.demag_field()
    scale = np.abs(hc_o).max()
    assert np.abs(hc_g - hc_o).max() <= 1e-12 * scale
    assert np.abs(hg_g - hg_o).max() <= 1e-12 * scale


def test_tensor_build_parity(oracle_mod):
    """FP64 Newell tensor built on the GPU against the oracle.  The 64-term g1/g2 sums
    cancel catastrophically: the reference's own tensor is not even symmetric under
    r -> -r beyond ~1e-8 relative (its noise floor).  The device tensor (different log/atan
    ulps) must agree with the oracle to within 10x that floor, and to 1e-8 of the max."""
    say("tensor build parity 32x24x2")
    nx, ny, nl = 32, 24, 2
    dz = [2e-7, 3e-7]
    gpu = Llg(nx, ny, nl, 2e-7, 2e-7, dz, il_space=[0.0, 1e-7])
    got = gpu.build_tensor(num_images=2, return_real=True)
    ref = oracle_mod.calculate_tensor(nx, ny, nl, 2e-7, 2e-7, dz, il_space=[0.0, 1e-7], num_images=2)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-8 * scale
    # noise floor: asymmetry of the even diagonal components under (i,j) -> (-i,-j)
    R = ref.reshape(nl * nl, 9, ny, nx)
    flip = R[:, :, (-np.arange(ny)) % ny][:, :, :, (-np.arange(nx)) % nx]
    floor = max(np.abs(R[:, c] - flip[:, c]).max() for c in (0, 4, 8))
    assert floor > 0
    assert np.abs(got - ref).max() <= 10 * floor
    # the symmetric components are bit-identical on the device as well
    T = got.reshape(nl * nl, 9, ny * nx)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert np.array_equal(T[:, 3 * a + b], T[:, 3 * b + a])


@pytest.mark.parametrize("sul_mu,head_mu", [(9.0, 1.0), (1.0, 50.0)])
def test_tensor_images_parity(oracle_mod, sul_mu, head_mu):
    """The device tensor with the SUL or head image (tensor.f90:157-181) against the oracle,
    to the same bound as the direct tensor; the image makes xz != zx, so the library keeps
    9 slots and the demag field still matches."""
    say(f"tensor images sul_mu {sul_mu} head_mu {head_mu}")
    nx, ny, nl = 32, 24, 2
    dz = [2e-7, 3e-7]
    il = [0.0, 1e-7]
    gpu = Llg(nx, ny, nl, 2e-7, 2e-7, dz, il_space=il)
    got = gpu.build_tensor(num_images=1, return_real=True, sul_mu=sul_mu, head_mu=head_mu, fly_height=6e-7,
                           space_to_sul=30e-7)
    ref = oracle_mod.calculate_tensor(nx, ny, nl, 2e-7, 2e-7, dz, il_space=il, num_images=1, sul_mu=sul_mu,
                                      head_mu=head_mu, fly_height=6e-7, space_to_sul=30e-7)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-8 * scale
    T = ref.reshape(nl * nl, 9, ny * nx)
    assert np.abs(T[:, 2] - T[:, 6]).max() > 1e-6 * scale     # xz != zx with an image
    med = media.make_media(nx, ny, dz=dz, n_lay=nl, pitch=6e-7, seed=77)
    s, g2 = pair(oracle_mod, med, tensor=ref, il=il)
    for k, m in enumerate(random_state(med, 3)):
        s.set_state(k, m)
        g2.set_state(k, m)
    hc0, hg0 = s.demag_field()
    hc1, hg1 = g2.demag_field()
    assert np.abs(hc0 - hc1).max() <= 1e-12 * np.abs(hc0).max()


def test_heun_step_parity_cfg0(oracle_mod):
    """Configuration 0 media, demag on, ramped tilted field, T = 0: Heun FP64 matches the
    oracle to 1e-10 after one step and after 100."""
    say("heun parity cfg0")
    med = media.config_media("cfg0")
    s, gpu = pair(oracle_mod, med, ufields=media.ramp_field(1))
    for k, m in enumerate(random_state(med, 5, tilt=0.05)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    dt = 5e-14
    s.step(1, dt, t0=0.0)
    gpu.set_llg(dt=dt)
    gpu.step(1, t0=0.0)
    assert max_dm(s, gpu, 1) <= 1e-10
    do = s.step(99, dt, t0=dt)
    dg = gpu.step(99, t0=dt)
    assert max_dm(s, gpu, 1) <= 1e-10
    assert do.t == dg.t
    assert abs(do.errmx - dg.errmx) <= 1e-8 * do.errmx
    assert abs(do.mave[0][2] - dg.mave[0][2]) < 1e-10
    check_loop_stats(do, dg, 1)
    # the stored field is the stage-2 field of the last step
    ho = s.get_state(0)[1]
    hg = gpu.get_state(0)[1]
    assert np.abs(ho - hg).max() <= 1e-9 * np.abs(ho).max()


@pytest.mark.parametrize("ltype", [2, -2, 1, -1])
def test_rotational_push_parity(oracle_mod, ltype):
    say(f"rotational push parity type {ltype}")
    med = media.config_media("cfg0", nx=32, ny=32)
    s, gpu = pair(oracle_mod, med, ufields=media.ramp_field(1))
    for k, m in enumerate(random_state(med, 6, tilt=0.3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    s.step(30, 5e-14, type=ltype)
    gpu.set_llg(dt=5e-14, llg_type=ltype)
    gpu.step(30)
    assert max_dm(s, gpu, 1) <= 1e-10


def test_thermal_injected_parity_multilayer(oracle_mod):
    """Four sub-layers (two physical layers, interlayer CSR), T = 300 K, the reference's MT
    streams (one per layer, seed 4357 + thermal_seed) injected into both paths."""
    say("thermal multilayer parity 64x64x4")
    med = media.config_media("cfg1", nx=64, ny=64, thm_alpha=1.0)
    s, gpu = pair(oracle_mod, med, ufields=media.ramp_field(4))
    n = 20
    noise = oracle_mod.mt_noise([34, 35, 36, 37], med.ngs, [True] * 4, n)
    s.step(n, 5e-14, noise=noise)
    gpu.set_llg(dt=5e-14)
    gpu.step(n, noise=noise)
    assert max_dm(s, gpu, 4) <= 1e-10


def test_anisotropy_variants_parity(oracle_mod):
    """Cubic anisotropy, a static per-grain applied field, ms_scale = 0 and hk_scale <= 0
    grains, no intralayer exchange in one layer."""
    say("anisotropy variants parity")
    med = media.make_media(32, 32, n_lay=2, dz=4e-7, pitch=6e-7, seed=8)
    rng = np.random.default_rng(1)
    for k, L in enumerate(med.layers):
        ng = L.ng
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        cub = np.tile(q.T.reshape(1, 3, 3), (ng, 1, 1))   # [n][c][r] = axis c component r
        mss = np.ones(ng)
        mss[::17] = 0.0
        hks = np.ones(ng)
        hks[::13] = 0.0
        kw = dict(ng=ng, grid2grain=L.grid2grain, grain2gridindex=L.grain2gridindex, grain2grid=L.grain2grid,
                  grain_vol=L.grain_vol, ms=L.ms, hk=np.tile([4000.0, 1500.0], (ng, 1)), cubic_axis=cub,
                  uni_axis=L.uni_axis if k == 0 else None, ms_scale=mss, hk_scale=hks,
                  h_app=rng.standard_normal((ng, 3)) * 300.0, alpha=0.05)
        if k == 0:
            kw.update(hxgrainindex=L.hxgrainindex, hxgrain=L.hxgrain, hx=L.hx,
                      lhxgrainindex=L.lhxgrainindex, lhxgrain=L.lhxgrain, lhx=L.lhx)
        else:
            kw.update(uhxgrainindex=L.uhxgrainindex, uhxgrain=L.uhxgrain, uhx=L.uhx)
        med.layers[k] = abi.LayerArrays(**kw)
    s, gpu = pair(oracle_mod, med, ufields=media.ramp_field(2))
    for k, m in enumerate(random_state(med, 2)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    s.step(25, 5e-14)
    gpu.set_llg(dt=5e-14)
    gpu.step(25)
    assert max_dm(s, gpu, 2) <= 1e-10


def test_philox_noise_is_what_the_step_uses(oracle_mod):
    """Running with on-device Philox equals injecting the normals mrmg_philox_normals
    dumps (bit for bit), and those normals match the oracle fed with them."""
    say("philox equivalence")
    med = media.config_media("cfg1", nx=32, ny=32)
    a = Llg(32, 32, 4, med.dx, med.dy, med.dz, demag=False)
    b = Llg(32, 32, 4, med.dx, med.dy, med.dz, demag=False)
    for x in (a, b):
        x.set_media(med)
        x.set_llg(dt=5e-14)
    a.step(40, philox_seed=1234)
    nz = b.philox_normals(1234, 0, 40)
    b.step(40, noise=nz)
    for k in range(4):
        assert np.array_equal(a.get_state(k)[0], b.get_state(k)[0])
    flat = nz.ravel()
    # 40 x G x 3 = 30720 normals: 4 sigma bounds on the mean and variance
    assert abs(flat.mean()) < 4 / np.sqrt(flat.size) and abs(flat.var() - 1) < 4 * np.sqrt(2 / flat.size)


def test_philox_in_kernel_draw_equals_the_chunked_draw():
    """A run split into one-step calls draws its normals inside the grain kernel
    (philox_normals3); one multi-step call draws them ahead with k_philox_dump.  Both must
    give the same trajectory bit for bit, so results do not depend on how a run is split
    into calls; with demag on, since the cells are refreshed at each call's start."""
    say("philox split: 40 x step(1) vs step(40)")
    med = media.config_media("cfg1", nx=32, ny=32)
    N = None
    out = []
    for split in (False, True):
        x = Llg(32, 32, 4, med.dx, med.dy, med.dz)
        if N is None:
            N = x.build_tensor(num_images=1, return_real=True)
        else:
            x.set_tensor(N)
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(4)):
            x.set_ufield(k, u)
        x.set_llg(dt=5e-14)
        if split:
            t = 0.0
            for i in range(40):
                t = x.step(1, t0=t, philox_seed=77).t
        else:
            x.step(40, philox_seed=77)
        out.append([x.get_state(k)[0] for k in range(4)])
        x.close()
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_adaptive_philox_equals_injected_dump():
    """The adaptive path always draws Philox normals inside the grain kernel: it must equal
    the same adaptive run fed the normals mrmg_philox_normals dumps, rejections included."""
    say("adaptive philox vs injected dump")
    med = media.config_media("cfg1", nx=32, ny=32)
    a = Llg(32, 32, 4, med.dx, med.dy, med.dz, demag=False)
    b = Llg(32, 32, 4, med.dx, med.dy, med.dz, demag=False)
    for x in (a, b):
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(4)):
            x.set_ufield(k, u)
        x.set_llg(dt=1e-14, dt_frac=0.2, dt_max=1e-12, max_dm=2.0)
    nz = b.philox_normals(4321, 0, 30)
    da = a.step(30, philox_seed=4321)
    db = b.step(30, noise=nz)
    assert da.rejected == db.rejected and da.t == db.t and da.dt == db.dt
    for k in range(4):
        assert np.array_equal(a.get_state(k)[0], b.get_state(k)[0]), k


def test_philox_normals_match_numpy_restatement():
    from tests.philox_np import philox_normals
    say("philox vs numpy")
    med = media.config_media("cfg0", nx=16, ny=16)
    a = Llg(16, 16, 1, med.dx, med.dy, med.dz, demag=False)
    a.set_media(med)
    got = a.philox_normals(99, 7, 3)
    ref = philox_normals(99, 7, 3, sum(med.ngs))
    assert np.abs(got - ref).max() < 1e-12


def test_precision_form_vs_fp64():
    """FP32 precision form against the FP64 GPU path after 2000 steps: max |dm| <= 1e-4."""
    say("fp32 vs fp64 cfg0 2000 steps")
    med = media.config_media("cfg0")
    out = []
    for prec in (abi.MRMG_FP64, abi.MRMG_FP32):
        x = Llg(64, 64, 1, med.dx, med.dy, med.dz, precision=prec)
        x.build_tensor(num_images=2)
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(1)):
            x.set_ufield(k, u)
        x.set_llg(dt=5e-14)
        x.step(2000)
        out.append(x.get_state(0)[0])
    d = np.abs(out[0] - out[1])
    say("fp32 max|dm|", d.max(), "rms", np.sqrt((d ** 2).mean()))
    assert d.max() <= 1e-4
    assert np.abs(np.linalg.norm(out[1], axis=1) - 1).max() < 1e-5


def test_cfg1_full_size_parity(oracle_mod):
    """Configuration 1 at full size (256x256x4), GPU-built tensor handed to both paths,
    reference MT noise injected, 3 Heun steps at T = 300 K."""
    say("cfg1 full size parity")
    med = media.config_media("cfg1", thm_alpha=1.0)
    gpu = Llg(256, 256, 4, med.dx, med.dy, med.dz)
    N = gpu.build_tensor(num_images=2, return_real=True)
    s, g2 = pair(oracle_mod, med, tensor=N, ufields=media.ramp_field(4))
    noise = oracle_mod.mt_noise([34, 35, 36, 37], med.ngs, [True] * 4, 3)
    s.step(3, 5e-14, noise=noise)
    g2.set_llg(dt=5e-14)
    g2.step(3, noise=noise)
    assert max_dm(s, g2, 4) <= 1e-10


def test_cfg1_full_size_long_run(oracle_mod):
    """Configuration 1 at full size for 100 Heun steps (5 ps) at T = 300 K with the
    reference's MT noise, split over four calls: the trajectories stay within 1e-10."""
    say("cfg1 full size, 100 steps")
    med = media.config_media("cfg1", thm_alpha=1.0)
    gpu = Llg(256, 256, 4, med.dx, med.dy, med.dz)
    N = gpu.build_tensor(num_images=2, return_real=True)
    s, g2 = pair(oracle_mod, med, tensor=N, ufields=media.ramp_field(4))
    for k, m in enumerate(random_state(med, 12, tilt=0.2)):
        s.set_state(k, m)
        g2.set_state(k, m)
    noise = oracle_mod.mt_noise([34, 35, 36, 37], med.ngs, [True] * 4, 100)
    g2.set_llg(dt=5e-14)
    t = 0.0
    for c in range(4):
        s.step(25, 5e-14, t0=t, noise=noise[25 * c:25 * (c + 1)])
        t = g2.step(25, t0=t, noise=noise[25 * c:25 * (c + 1)]).t
    d = max_dm(s, g2, 4)
    say(f"  max|dm| after 100 steps: {d:.2e}")
    assert d <= 1e-10


def test_state_survives_layer_rebuild(oracle_mod):
    """set_layer after steps keeps the other layers' device state (the host copy is synced)."""
    med = media.config_media("cfg1", nx=32, ny=32)
    x = Llg(32, 32, 4, med.dx, med.dy, med.dz, demag=False)
    x.set_media(med)
    x.set_llg(dt=5e-14)
    x.step(5, philox_seed=3)
    before = [x.get_state(k)[0] for k in range(4)]
    x.set_layer(2, med.layers[2])
    after = [x.get_state(k)[0] for k in range(4)]
    for k in (0, 1, 3):
        assert np.array_equal(before[k], after[k])
    assert np.allclose(after[2], [0, 0, 1])


@pytest.mark.parametrize("max_dm", [180.0, 1.5])
def test_adaptive_dt_parity(oracle_mod, max_dm):
    """Adaptive dt (get_fields) and push_m's rejection/retry on the GPU against the oracle:
    same accepted steps, rejections, dt and trajectory (thermal, injected MT noise)."""
    say(f"adaptive dt parity max_dm={max_dm}")
    med = media.config_media("cfg1", nx=64, ny=64, thm_alpha=1.0)
    s, gpu = pair(oracle_mod, med, ufields=media.ramp_field(4))
    for k, m in enumerate(random_state(med, 4, tilt=0.2)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    n = 12
    noise = oracle_mod.mt_noise([34, 35, 36, 37], med.ngs, [True] * 4, 2 * n)
    gpu.set_llg(dt=1e-14, dt_frac=0.2, dt_max=1e-12, max_dm=max_dm)
    t, dt, rejected = 0.0, 1e-14, 0
    for c in range(2):
        do = s.step(n, dt, t0=t, noise=noise[c * n:(c + 1) * n], dt_frac=0.2, dt_max=1e-12, max_dm=max_dm)
        dg = gpu.step(n, t0=t, noise=noise[c * n:(c + 1) * n])
        assert do.rejected == dg.rejected
        assert abs(do.dt - dg.dt) <= 1e-12 * do.dt and abs(do.t - dg.t) <= 1e-12 * do.t
        check_loop_stats(do, dg, med.n_lay, tol=1e-9)
        t, dt, rejected = do.t, do.dt, rejected + do.rejected
    if max_dm < 10:
        assert rejected > 0
    assert max_dm_diff(s, gpu) <= 1e-10


def max_dm_diff(s, gpu):
    return max_dm(s, gpu, gpu.n_lay)


def test_launch_shapes_do_not_change_results():
    """Row/column launch shapes give bit-identical trajectories; lanes per grain only
    reorder the gather sums (last bits).  The autotuner restores the state it steps."""
    from mrm_amd.tune import autotune
    say("launch shapes")
    med = media.config_media("cfg1", nx=64, ny=64)
    N = None
    out = {}
    for cfg in ({}, {"row_threads": 64}, {"row_threads": 256, "col_threads": 256}, {"col_groups": 2},
                {"col_groups": 4, "col_threads": 256}, {"col_threads": 1024}, {"col_path": 3},
                {"col_path": 3, "col_groups": 2}, {"lanes_per_grain": 1}, {"lanes_per_grain": 2},
                {"lanes_per_grain": 4}, {"lanes_per_grain": 16}):
        x = Llg(64, 64, 4, med.dx, med.dy, med.dz)
        if N is None:
            N = x.build_tensor(num_images=1, return_real=True)
        else:
            x.set_tensor(N)
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(4)):
            x.set_ufield(k, u)
        for key, v in cfg.items():
            x.tune(key, v)
        x.set_llg(dt=5e-14)
        x.step(15, philox_seed=9)
        out[str(cfg)] = [x.get_state(k)[0] for k in range(4)]
    base = out["{}"]
    for key, ms in out.items():
        d = max(np.abs(p - q).max() for p, q in zip(base, ms))
        if "lanes" in key:
            assert d <= 1e-12, key
        else:
            assert d == 0.0, key
    before = [x.get_state(k)[0] for k in range(4)]
    autotune(x, steps=3)
    after = [x.get_state(k)[0] for k in range(4)]
    assert all(np.array_equal(p, q) for p, q in zip(before, after))
    # the one-key sweep of the open-boundary bench leg: it keeps a listed value and the state
    from mrm_amd.tune import tune_keys
    got = tune_keys(x, {"row_block": [2, 4, 8]}, steps=2)
    assert got["row_block"] in (2, 4, 8) and x.get_tune("row_block") == got["row_block"]
    assert all(np.array_equal(p, x.get_state(k)[0]) for k, p in enumerate(before))


def test_launch_shapes_agree_at_power_of_two_sizes():
    """At 256 x 256 the FFTs take the compile-time power-of-two passes: the fused, split and
    column-contiguous paths still agree bit for bit (row radix 16), and the generic passes
    (row and column radix 8) agree with them to rounding."""
    say("launch shapes 256x256x2")
    med = media.make_media(256, 256, n_lay=2, seed=4)
    N = None
    out = {}
    for cfg in ({}, {"col_path": 1}, {"col_path": 2}, {"col_path": 3}, {"col_path": 1, "row_block": 2},
                {"row_radix": 8, "col_radix": 8}):
        x = Llg(256, 256, 2, med.dx, med.dy, med.dz)
        if N is None:
            N = x.build_tensor(num_images=1, return_real=True)
        else:
            x.set_tensor(N)
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(2)):
            x.set_ufield(k, u)
        for key, v in cfg.items():
            x.tune(key, v)
        x.set_llg(dt=5e-14)
        x.step(10, philox_seed=3)
        out[str(cfg)] = [x.get_state(k)[0] for k in range(2)]
        x.close()
    base = out["{}"]
    for key, ms in out.items():
        d = max(np.abs(p - q).max() for p, q in zip(base, ms))
        say(f"  {key}: max|dm| {d:.3g}")
        if "radix" in key:
            assert d <= 1e-12, key
        else:
            assert d == 0.0, key


def test_step_rejects_bad_arguments():
    med = media.config_media("cfg0", nx=16, ny=16)
    x = Llg(16, 16, 1, med.dx, med.dy, med.dz)
    x.set_media(med)
    with pytest.raises(abi.MrmgError):
        x.step(1)          # demag on, no tensor
    y = Llg(16, 16, 1, med.dx, med.dy, med.dz, demag=False)
    y.set_media(med)
    y.set_llg(dt=5e-14, llg_type=7)
    with pytest.raises(abi.MrmgError):
        y.step(1)


def test_time_kernel_leaves_state_alone():
    """mrmg_time_kernel times each demag class of the active path by back-to-back launches
    and does not change the trajectory; classes of the other path are refused."""
    med = media.config_media("cfg1", nx=64, ny=64)
    runs = []
    for timed in (False, True):
        x = Llg(64, 64, 4, med.dx, med.dy, med.dz)
        x.build_tensor(num_images=1)
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(4)):
            x.set_ufield(k, u)
        x.set_llg(dt=5e-14)
        x.tune("col_path", 2)
        x.step(5, philox_seed=3)
        if timed:
            for name in ("row_fwd", "col", "row_inv"):
                ms = x.time_kernel(name, reps=5)
                say(f"  {name}: {ms * 1e3:.2f} us")
                assert 0.0 < ms < 10.0
            with pytest.raises(abi.MrmgError):
                x.time_kernel("prod")
            with pytest.raises(abi.MrmgError):
                x.time_kernel("grain1")
            x.tune("col_path", 1)
            for name in ("col_fwd", "prod", "col_inv"):
                assert x.time_kernel(name, reps=3) > 0.0
            x.tune("col_path", 2)
        x.step(5, t0=5 * 5e-14, philox_seed=3)
        runs.append([x.get_state(k)[0] for k in range(4)])
        x.close()
    assert all(np.array_equal(p, q) for p, q in zip(*runs))
