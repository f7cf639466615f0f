"""Batched steps of several contexts (mrmg_step_batch, the serving path): each demag
kernel once for all members over a table of their operands, one step captured as a
hipGraph and replayed.

* Each member ends bit for bit where its own mrmg_step with the same parameters ends:
  thermal Heun with Philox noise and deterministic runs, FP64 and FP32, the fused and the
  split column paths, graph and direct launches, a run split into chunks (the end time
  returned by a call is the next call's start) against one solo call, and the
  rotational push.
* Members whose grids or launch shapes differ are refused, and so is one that appears twice.
* A graph replays state-dependent work only: the same batch stepped twice from the same
  start gives the same result.
"""
import numpy as np
import pytest

from mrm_amd import abi, media
from mrm_amd.llg import Llg, step_batch
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu


def member(nx, ny, seed, prec, path=None, type_=abi.MRMG_TYPE_HEUN):
    med = media.config_media("cfg1", nx=nx, ny=ny, seed=seed)
    x = Llg(nx, ny, med.n_lay, med.dx, med.dy, med.dz, precision=prec)
    x.build_tensor(num_images=2)
    x.set_media(med)
    for k, u in enumerate(media.ramp_field(med.n_lay, h_max=1.5e4, t_ramp=2e-11)):
        x.set_ufield(k, u)
    x.set_llg(dt=5e-14, llg_type=type_)
    if path is not None:
        x.tune("col_path", path)
    return x, med


def states(x, nl):
    return [x.get_state(k)[0] for k in range(nl)] + [x.get_mag(k) for k in range(nl)]


def same(a, b):
    return all(np.array_equal(p, q) for p, q in zip(a, b))


@pytest.mark.parametrize("prec,path,graph", [
    (abi.MRMG_FP64, None, True), (abi.MRMG_FP32, None, True), (abi.MRMG_FP64, None, False),
    (abi.MRMG_FP32, 1, True), (abi.MRMG_FP64, 2, True)])
def test_batch_matches_solo(prec, path, graph):
    nx, ny, n = 48, 40, 24
    seeds = [3, 5, 8]
    solo, batch, meds = [], [], []
    for i, sd in enumerate(seeds):
        x, med = member(nx, ny, sd, prec, path)
        x.step(n, philox_seed=100 + i)
        solo.append(states(x, med.n_lay))
        x.close()
        y, med = member(nx, ny, sd, prec, path)
        batch.append(y)
        meds.append(med)
    t = step_batch(batch, n, philox_seeds=[100 + i for i in range(len(seeds))], graph=graph)
    ok = [same(solo[i], states(batch[i], meds[i].n_lay)) for i in range(len(seeds))]
    say(f"batch of {len(seeds)} {nx}x{ny} prec {prec} path {path} graph {graph}: bit-exact {ok}, t_end {t[0]:.6e}")
    assert all(ok)
    assert all(v == t[0] for v in t) and t[0] > 0.0
    for y in batch:
        y.close()


def test_batch_in_chunks_matches_one_solo_call():
    nx, ny = 40, 36
    x, med = member(nx, ny, 4, abi.MRMG_FP32)
    d = x.step(30, philox_seed=7)
    want = states(x, med.n_lay)
    y, _ = member(nx, ny, 4, abi.MRMG_FP32)
    z, _ = member(nx, ny, 9, abi.MRMG_FP32)
    t = [0.0, 0.0]
    for chunk in (7, 13, 10):
        t = step_batch([y, z], chunk, philox_seeds=[7, 8], t0s=t)
    say(f"chunked batch: bit-exact {same(want, states(y, med.n_lay))}, t {t[0]!r} vs solo {d.t!r}")
    assert same(want, states(y, med.n_lay))
    assert t[0] == d.t
    for q in (x, y, z):
        q.close()


def test_batch_rotational_push_and_no_noise():
    nx, ny = 32, 32
    for type_ in (-2, 2):
        x, med = member(nx, ny, 6, abi.MRMG_FP64, type_=type_)
        x.step(12)
        want = states(x, med.n_lay)
        y, _ = member(nx, ny, 6, abi.MRMG_FP64, type_=type_)
        z, _ = member(nx, ny, 2, abi.MRMG_FP64, type_=type_)
        step_batch([y, z], 12, graph=type_ < 0)
        assert same(want, states(y, med.n_lay)), type_
        for q in (x, y, z):
            q.close()


def test_batch_replay_is_repeatable():
    a, med = member(32, 24, 1, abi.MRMG_FP32)
    b, _ = member(32, 24, 2, abi.MRMG_FP32)
    start = [states(q, med.n_lay)[:med.n_lay] for q in (a, b)]
    step_batch([a, b], 15, philox_seeds=[1, 2])
    first = [states(q, med.n_lay) for q in (a, b)]
    for q, st in zip((a, b), start):
        for k in range(med.n_lay):
            q.set_state(k, st[k])
        q.step_index = 0
    step_batch([a, b], 15, philox_seeds=[1, 2])
    assert all(same(f, states(q, med.n_lay)) for f, q in zip(first, (a, b)))
    a.close()
    b.close()


def test_batch_refuses_mismatched_members():
    a, _ = member(32, 24, 1, abi.MRMG_FP32)
    b, _ = member(32, 32, 2, abi.MRMG_FP32)
    with pytest.raises(abi.MrmgError):
        step_batch([a, b], 2, philox_seeds=[1, 2])
    c, _ = member(32, 24, 3, abi.MRMG_FP32)
    c.tune("row_threads", 64 if c.get_tune("row_threads") != 64 else 128)
    with pytest.raises(abi.MrmgError):
        step_batch([a, c], 2, philox_seeds=[1, 2])
    with pytest.raises(abi.MrmgError):
        step_batch([a, a], 2, philox_seeds=[1, 2])
    for q in (a, b, c):
        q.close()
