"""mrmg_readback_noalloc: the readback with its scratch in the library's fixed device array
instead of buffers it allocates.  It must give mrmg_readback's values bit for bit (the same
kernels), on consecutive calls whose second input is smaller than the first (the scratch
keeps the last call's contents), and refuse what does not fit or needs an allocation."""
import ctypes as C

import numpy as np
import pytest

from mrm_amd import abi, media
from mrm_amd.llg import Llg

pytestmark = pytest.mark.gpu

MRMG_ERR_ARG, MRMG_ERR_STATE = 1, 3
PLANE_MAX, POS_MAX = 4 * 1024 * 1024, 65536


def say(msg):
    print(msg, flush=True)


def sensor(sx, sy, direction, weights):
    s = abi.Sensor(sigma_x=sx, sigma_y=sy, w=abi.dptr(weights))
    for i in range(3):
        s.dir[i] = direction[i]
    return s


def noalloc(x, s, xs, ys):
    out = np.full(xs.size, np.nan)
    rc = x.lib.mrmg_readback_noalloc(x.ctx, C.byref(s), xs.size, abi.dptr(xs), abi.dptr(ys), abi.dptr(out))
    abi.check(x.lib, rc)
    return out


def built_context(nx, ny, n_lay, precision, seed):
    """A Voronoi media in a random state, with a few steps run (which builds the context)."""
    med = media.make_media(nx, ny, dx=2e-7, dy=2e-7, dz=2e-7, n_lay=n_lay, seed=seed)
    x = Llg(nx, ny, n_lay, med.dx, med.dy, med.dz, precision=precision)
    x.build_tensor(num_images=1)
    x.set_media(med)
    rng = np.random.default_rng(seed)
    for k in range(n_lay):
        m = rng.standard_normal((med.ngs[k], 3))
        x.set_state(k, m / np.linalg.norm(m, axis=1, keepdims=True))
    x.set_llg(dt=5e-14)
    x.step(2)
    return x


def test_noalloc_readback_matches_readback_on_consecutive_calls():
    # the first call the larger: 128x96x3 FP64 and 300 head positions; then 64x64x1 FP32 and 37
    cases = [(128, 96, 3, abi.MRMG_FP64, 300, 11), (64, 64, 1, abi.MRMG_FP32, 37, 12)]
    for nx, ny, nl, prec, npos, seed in cases:
        say(f"no-alloc readback: {nx}x{ny}x{nl} precision {prec}, {npos} positions")
        x = built_context(nx, ny, nl, prec, seed)
        rng = np.random.default_rng(seed + 100)
        xs = np.ascontiguousarray(rng.random(npos) * nx * 2e-7)
        ys = np.ascontiguousarray(rng.random(npos) * ny * 2e-7)
        w = np.ascontiguousarray(np.linspace(1.0, 0.5, nl))
        d = (0.1, -0.2, 0.97)
        want = x.readback(xs, ys, (3e-7, 4e-7), direction=d, weights=w)
        got = noalloc(x, sensor(3e-7, 4e-7, d, w), xs, ys)
        assert np.array_equal(got, want), np.abs(got - want).max()
        # again on the same context: the scratch now holds this call's own data
        got2 = noalloc(x, sensor(3e-7, 4e-7, d, w), xs[: npos // 2], ys[: npos // 2])
        assert np.array_equal(got2, want[: npos // 2])
        x.close()


def test_noalloc_readback_refuses_what_it_cannot_do():
    nx, ny = 32, 32
    med = media.make_media(nx, ny, n_lay=1, seed=3)
    x = Llg(nx, ny, 1, med.dx, med.dy, med.dz, demag=False)
    x.set_media(med)
    s = sensor(2e-7, 2e-7, (0.0, 0.0, 1.0), None)
    xs = np.zeros(4)
    out = np.zeros(4)
    # not built yet: building allocates, so the no-allocation entry point declines
    assert x.lib.mrmg_readback_noalloc(x.ctx, C.byref(s), 4, abi.dptr(xs), abi.dptr(xs), abi.dptr(out)) == MRMG_ERR_STATE
    x.set_llg(dt=5e-14)
    x.step(1)
    big = np.zeros(POS_MAX + 1)
    assert x.lib.mrmg_readback_noalloc(x.ctx, C.byref(s), big.size, abi.dptr(big), abi.dptr(big),
                                       abi.dptr(big)) == MRMG_ERR_ARG
    assert b"exceeds" in x.lib.mrmg_last_error()
    # at the limit it runs
    lim = np.ascontiguousarray(np.linspace(0.0, nx * 2e-7, POS_MAX))
    got = noalloc(x, s, lim, lim)
    want = x.readback(lim, lim, (2e-7, 2e-7))
    assert np.array_equal(got, want)
    x.close()
