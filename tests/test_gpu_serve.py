"""The request loop with continuous batching: requests of one padded grid are stepped
together (mrmg_step_batch), groups overlap on the device, and each result is the same
alone and in the mix, bit for bit."""
import numpy as np
import pytest

from mrm_amd import abi, media, serve
from mrm_amd.llg import Llg
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _copy(reqs):
    return [serve.Request(rid=r.rid, med=r.med, n_steps=r.n_steps, seed=r.seed, precision=r.precision,
                          arrival=r.arrival) for r in reqs]


def test_results_same_alone_and_in_the_mix():
    # one grid size, so the layer counts make three groups of a few requests each
    reqs = serve.build_load(10, rate=5000.0, sizes=(64, 64), steps=(40, 160), seed=5)
    alone = _copy(reqs)
    srv = serve.Server(workers=1, max_batch=1, chunk=37)
    srv.run(alone)
    srv.close()
    mixed = _copy(reqs)
    srv = serve.Server(workers=3, max_batch=8, chunk=50, batch_wait=2.0)   # a window that collects each group
    wall = srv.run(mixed)
    srv.close()
    st = serve.latency_stats(mixed, wall)
    say(f"serve: {st}, batch sizes {srv.batch_sizes}")
    assert max(srv.batch_sizes) > 1, "no two requests were stepped in one batch"
    overlap = any(b.start < a.end and a.start < b.end for a in mixed for b in mixed if a.rid < b.rid)
    assert overlap, "no two requests were in flight together"
    for a, b in zip(sorted(alone, key=lambda r: r.rid), sorted(mixed, key=lambda r: r.rid)):
        for ma, mb in zip(a.m, b.m):
            assert np.array_equal(ma, mb), a.rid
    # and the same as driving one context directly with one call
    r = reqs[3]
    x = Llg(r.med.nx, r.med.ny, r.med.n_lay, r.med.dx, r.med.dy, r.med.dz, precision=r.precision)
    x.build_tensor(num_images=2)
    x.set_media(r.med)
    for k, u in enumerate(media.ramp_field(r.med.n_lay)):
        x.set_ufield(k, u)
    x.set_llg(dt=r.dt)
    x.step(r.n_steps, philox_seed=r.seed, diag=False)
    ref = next(q for q in mixed if q.rid == 3)
    for k in range(r.med.n_lay):
        assert np.array_equal(x.get_state(k)[0], ref.m[k])
    x.close()


def test_copied_tensor_gives_the_built_ones_results():
    med = media.config_media("cfg1", nx=64, ny=48, seed=2)
    a = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=abi.MRMG_FP32)
    a.build_tensor(num_images=2)
    b = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=abi.MRMG_FP32)
    b.copy_tensor(a)
    out = []
    for x in (a, b):
        x.set_media(med)
        x.set_llg(dt=5e-14)
        x.step(10, philox_seed=4)
        out.append([x.get_state(k)[0] for k in range(med.n_lay)])
    assert all(np.array_equal(p, q) for p, q in zip(*out))
    c = Llg(med.nx, med.ny + 16, med.n_lay, med.dx, med.dy, med.dz, precision=abi.MRMG_FP32)
    with pytest.raises(abi.MrmgError):
        c.copy_tensor(a)
    for x in (a, b, c):
        x.close()


def test_admission_respects_the_memory_budget():
    reqs = serve.build_load(8, rate=1000.0, sizes=(64, 128), steps=(20, 60), seed=9)
    need = max(serve.footprint_estimate(r.med, r.precision) for r in reqs)
    budget = int(need * 2.5)
    srv = serve.Server(workers=3, mem_budget=budget, pool=2)
    srv.run(reqs)
    srv.close()
    assert srv.peak <= budget
    # every live context (admitted or idle in the pool) is counted at each admission, and
    # the estimate bounds what a context really holds
    assert max(srv.admissions) <= budget
    for r in reqs:
        assert 0 < r.footprint <= serve.footprint_estimate(r.med, r.precision), r.rid
    # at every admission the requests then in flight fit the budget together
    for r in reqs:
        live = [q for q in reqs if q.start <= r.start < q.end]
        assert sum(serve.footprint_estimate(q.med, q.precision) for q in live) <= budget
