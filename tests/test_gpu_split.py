"""The split column path (column-contiguous spectra; row, column, product, column, row
kernels) against the oracle.  At small grids it is forced with the col_path tunable; at
the 512 x 4096 recording-track grid it is the only path that fits."""
import numpy as np
import pytest

from mrm_amd import abi, media
from mrm_amd.llg import Llg
from tests.test_gpu_parity import pair, random_state, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,nl", [(64, 64, 1), (48, 45, 3), (40, 30, 2), (20, 27, 1), (30, 250, 2)])
def test_split_demag_parity(oracle_mod, nx, ny, nl):
    say(f"split demag parity {nx}x{ny}x{nl}")
    med = media.make_media(nx, ny, dz=[2e-7 + 1e-7 * k for k in range(nl)], n_lay=nl, pitch=6e-7, seed=nx + ny)
    il = [0.0] + [5e-8] * (nl - 1)
    s, gpu = pair(oracle_mod, med, il=il)
    gpu.tune("col_path", 1)
    assert gpu.get_tune("col_path") == 1
    for k, m in enumerate(random_state(med, 3)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    hc0, hg0 = s.demag_field()
    hc1, hg1 = gpu.demag_field()
    scale = np.abs(hc0).max()
    err = max(np.abs(hc0 - hc1).max(), np.abs(hg0 - hg1).max()) / scale
    say(f"  rel err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("precision", [abi.MRMG_FP64, abi.MRMG_FP32])
def test_split_equals_fused_steps(oracle_mod, precision):
    """Thermal multilayer Heun steps: the split path against the oracle (FP64) and against
    the fused path (both forms)."""
    O = oracle_mod
    med = media.make_media(64, 64, dz=[4e-7, 4e-7], n_lay=2, pitch=8e-7, seed=4, thermal=True, temperature=300.0,
                           layer_ku=[2e6, 1e7])
    tensor = O.calculate_tensor(64, 64, 2, med.dx, med.dy, med.dz)
    s, a = pair(O, med, tensor=tensor, precision=precision, ufields=media.ramp_field(2))
    _, b = pair(O, med, tensor=tensor, precision=precision, ufields=media.ramp_field(2))
    b.tune("col_path", 1)
    noise = O.mt_noise([5, 6], [L.ng for L in med.layers], [True, True], 30)
    for x in (a, b):
        x.set_llg(dt=5e-14)
        x.step(30, noise=noise)
    dab = max(np.abs(a.get_state(k)[0] - b.get_state(k)[0]).max() for k in range(2))
    say(f"split vs fused ({'fp64' if precision == abi.MRMG_FP64 else 'fp32'}): {dab:.2e}")
    assert dab <= (1e-13 if precision == abi.MRMG_FP64 else 1e-5)
    if precision == abi.MRMG_FP64:
        s.step(30, 5e-14, noise=noise)
        d = max(np.abs(s.get_state(k)[0] - b.get_state(k)[0]).max() for k in range(2))
        say(f"split vs oracle: {d:.2e}")
        assert d <= 1e-10


def test_track_grid_demag_parity(oracle_mod):
    """512 x 4096 x 2 cells of 1.5 nm (the recording-track grid): the fused path does not
    fit, the split path runs; demag field against the oracle with the device-built tensor."""
    O = oracle_mod
    nx, ny, nl = 512, 4096, 2
    med = media.make_media(nx, ny, dx=1.5e-7, dy=1.5e-7, dz=[1.5e-7] * nl, n_lay=nl, pitch=6e-7, seed=8,
                           exchange=False)
    gpu = Llg(nx, ny, nl, med.dx, med.dy, med.dz)
    assert gpu.get_tune("col_path") == 1
    N = gpu.build_tensor(num_images=2, return_real=True)
    g = abi.make_geom(nx, ny, nl, med.dx, med.dy, med.dz)
    s = O.Sim(g)
    s.set_tensor(N)
    del N
    for k, L in enumerate(med.layers):
        s.set_layer(k, L)
        gpu.set_layer(k, L)
    for k, m in enumerate(random_state(med, 9, tilt=0.5)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    hc0, hg0 = s.demag_field()
    hc1, hg1 = gpu.demag_field()
    err = max(np.abs(hc0 - hc1).max(), np.abs(hg0 - hg1).max()) / np.abs(hc0).max()
    say(f"track grid demag rel err {err:.2e}")
    assert err <= 1e-12


def test_track_grid_precision_form():
    """512 x 4096 x 4 (the FP32 recording-track configuration's grid and media): the FP32
    form stays within 1e-4 of the FP64 form after 200 thermal Heun steps."""
    med = media.config_media("cfg2")
    forms = []
    N = None
    for prec in (abi.MRMG_FP64, abi.MRMG_FP32):
        x = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=prec)
        if N is None:
            N = x.build_tensor(num_images=2, return_real=True)
        else:
            x.set_tensor(N)
        x.set_media(med)
        for k, u in enumerate(media.ramp_field(med.n_lay)):
            x.set_ufield(k, u)
        x.set_llg(dt=5e-14)
        forms.append(x)
    del N
    for x in forms:
        x.step(200, philox_seed=3)
    dmax, ss, n = 0.0, 0.0, 0
    for k in range(med.n_lay):
        m64, m32 = forms[0].get_state(k)[0], forms[1].get_state(k)[0]
        assert np.abs(np.linalg.norm(m32, axis=1) - 1).max() < 1e-5
        d = np.abs(m64 - m32)
        dmax = max(dmax, float(d.max()))
        ss += float((d ** 2).sum())
        n += d.size
    say(f"cfg2 grid fp32 vs fp64 after 200 steps: max {dmax:.2e} rms {(ss / n) ** 0.5:.2e}")
    assert dmax <= 1e-4



@pytest.mark.parametrize("row_block", [2, 4, 8])
def test_column_contiguous_fused_path(oracle_mod, row_block):
    """col_path 2 (transposing row kernels + the fused column kernel on column-contiguous
    spectra) gives the oracle's demag field, for every rows-per-workgroup block."""
    med = media.make_media(48, 45, dz=[2e-7, 3e-7, 2e-7], n_lay=3, pitch=6e-7, seed=93)
    il = [0.0, 5e-8, 5e-8]
    s, gpu = pair(oracle_mod, med, il=il)
    gpu.tune("col_path", 2)
    gpu.tune("row_block", row_block)
    assert gpu.get_tune("col_path") == 2 and gpu.get_tune("row_block") == row_block
    for k, m in enumerate(random_state(med, 4)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    hc0, hg0 = s.demag_field()
    hc1, hg1 = gpu.demag_field()
    err = max(np.abs(hc0 - hc1).max(), np.abs(hg0 - hg1).max()) / np.abs(hc0).max()
    say(f"col_path 2, row_block {row_block}: rel err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("path,cr", [(0, 4), (2, 8), (2, 16)])
def test_column_kernel_1024_threads(oracle_mod, path, cr):
    """The 1024-thread column kernel (two product items per thread) gives the oracle's
    demag field and the 512-thread kernel's, bit for bit, on both fused paths."""
    med = media.make_media(64, 48, dz=[2e-7, 3e-7, 2e-7], n_lay=3, pitch=6e-7, seed=71)
    s, gpu = pair(oracle_mod, med)
    gpu.tune("col_path", path)
    gpu.tune("col_radix", cr)
    for k, m in enumerate(random_state(med, 2)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    hc5, hg5 = gpu.demag_field()
    gpu.tune("col_threads", 1024)
    assert gpu.get_tune("col_threads") == 1024
    hc1, hg1 = gpu.demag_field()
    hc0, hg0 = s.demag_field()
    err = max(np.abs(hc0 - hc1).max(), np.abs(hg0 - hg1).max()) / np.abs(hc0).max()
    say(f"col_threads 1024, path {path}, col radix {cr}: rel err {err:.2e}")
    assert err <= 1e-12
    assert np.array_equal(hc1, hc5) and np.array_equal(hg1, hg5)
    with pytest.raises(abi.MrmgError):
        gpu.tune("wave_fft", 1)          # the 1024-thread kernel has no wave-local form


@pytest.mark.parametrize("path,rr,cr", [(0, 4, 8), (0, 8, 4), (1, 4, 4), (2, 8, 8), (2, 4, 16)])
def test_fft_radix_choices(oracle_mod, path, rr, cr):
    """Row and column FFTs with a smaller largest radix (more, lighter Stockham passes)
    give the oracle's demag field on every column path."""
    med = media.make_media(64, 48, dz=[2e-7, 3e-7], n_lay=2, pitch=6e-7, seed=97)
    s, gpu = pair(oracle_mod, med)
    gpu.tune("col_path", path)
    gpu.tune("row_radix", rr)
    gpu.tune("col_radix", cr)
    for k, m in enumerate(random_state(med, 6)):
        s.set_state(k, m)
        gpu.set_state(k, m)
    hc0, hg0 = s.demag_field()
    hc1, hg1 = gpu.demag_field()
    err = max(np.abs(hc0 - hc1).max(), np.abs(hg0 - hg1).max()) / np.abs(hc0).max()
    say(f"path {path} row radix {rr} col radix {cr}: rel err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("path,row_block,row_threads,col_threads,cr,precision", [
    (2, 8, 256, 512, 16, abi.MRMG_FP64), (2, 4, 128, 512, 8, abi.MRMG_FP64), (2, 2, 64, 256, 4, abi.MRMG_FP64),
    (1, 8, 256, 512, 16, abi.MRMG_FP64), (2, 8, 128, 256, 8, abi.MRMG_FP32), (1, 4, 256, 512, 4, abi.MRMG_FP32)])
def test_wave_local_ffts(oracle_mod, path, row_block, row_threads, col_threads, cr, precision):
    """wave_fft 1: each wave transforms its own run of rows or columns with no workgroup
    barrier inside the FFT.  The butterflies are those of the workgroup form, so the demag
    field is bit-for-bit the same, and (FP64) within 1e-12 of the oracle.  Sizes with
    radix-3 and radix-5 passes and runs that leave waves idle (9 columns over 4 or 8 waves)."""
    med = media.make_media(48, 45, dz=[2e-7, 3e-7, 2e-7], n_lay=3, pitch=6e-7, seed=95)
    s, gpu = pair(oracle_mod, med, precision=precision)
    fields = []
    for wv in (0, 1):
        for key, v in (("col_path", path), ("row_block", row_block), ("row_threads", row_threads),
                       ("col_threads", col_threads), ("col_radix", cr), ("wave_fft", wv)):
            gpu.tune(key, v)
        assert gpu.get_tune("wave_fft") == wv and gpu.get_tune("col_path") == path
        for k, m in enumerate(random_state(med, 7)):
            s.set_state(k, m)
            gpu.set_state(k, m)
        fields.append(gpu.demag_field())
    assert np.array_equal(fields[0][0], fields[1][0]) and np.array_equal(fields[0][1], fields[1][1])
    if precision == abi.MRMG_FP64:
        hc0, hg0 = s.demag_field()
        hc1, hg1 = fields[1]
        err = max(np.abs(hc0 - hc1).max(), np.abs(hg0 - hg1).max()) / np.abs(hc0).max()
        say(f"wave ffts path {path} rb {row_block} rt {row_threads} ct {col_threads} cr {cr}: rel err {err:.2e}")
        assert err <= 1e-12
