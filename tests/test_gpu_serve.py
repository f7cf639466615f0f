"""The request loop: requests overlap on the device (one context and stream each), and
each result is the same alone and in the mix, bit for bit."""
import numpy as np
import pytest

from mrm_amd import abi, media, serve
from mrm_amd.llg import Llg
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _copy(reqs):
    return [serve.Request(rid=r.rid, med=r.med, tensor=r.tensor, n_steps=r.n_steps, seed=r.seed,
                          precision=r.precision, arrival=r.arrival) for r in reqs]


def test_results_same_alone_and_in_the_mix():
    reqs = serve.build_load(8, rate=400.0, sizes=(48, 128), steps=(40, 120), seed=5)
    alone = _copy(reqs)
    srv = serve.Server(slots=1)
    srv.run(alone)
    srv.close()
    mixed = _copy(reqs)
    srv = serve.Server(slots=4)
    wall = srv.run(mixed)
    srv.close()
    st = serve.latency_stats(mixed, wall)
    say(f"serve: {st}")
    overlap = any(b.start < a.end and a.start < b.end for a in mixed for b in mixed if a.rid < b.rid)
    assert overlap, "no two requests were in flight together"
    for a, b in zip(sorted(alone, key=lambda r: r.rid), sorted(mixed, key=lambda r: r.rid)):
        for ma, mb in zip(a.m, b.m):
            assert np.array_equal(ma, mb), a.rid
    # and the same as driving one context directly
    r = reqs[3]
    x = Llg(r.med.nx, r.med.ny, r.med.n_lay, r.med.dx, r.med.dy, r.med.dz, precision=r.precision)
    x.set_tensor(r.tensor)
    x.set_media(r.med)
    for k, u in enumerate(media.ramp_field(r.med.n_lay)):
        x.set_ufield(k, u)
    x.set_llg(dt=r.dt)
    x.step(r.n_steps, philox_seed=r.seed, diag=False)
    ref = next(q for q in mixed if q.rid == 3)
    for k in range(r.med.n_lay):
        assert np.array_equal(x.get_state(k)[0], ref.m[k])


def test_admission_respects_the_memory_budget():
    reqs = serve.build_load(6, rate=1000.0, sizes=(64, 96), steps=(20, 40), seed=9)
    need = max(serve.footprint_estimate(r.med, r.precision) for r in reqs)
    srv = serve.Server(slots=6, mem_budget=int(need * 1.5))
    srv.run(reqs)
    srv.close()
    budget = int(need * 1.5)
    assert srv.peak <= budget
    # every live context (in flight or idle in the pool) is counted at each admission,
    # and the estimate bounds what a context really holds
    assert max(srv.admissions) <= budget
    for r in reqs:
        assert 0 < r.footprint <= serve.footprint_estimate(r.med, r.precision), r.rid
    # at every request's start, the requests then in flight fit the budget together
    for r in reqs:
        live = [q for q in reqs if q.start <= r.start < q.end]
        assert sum(serve.footprint_estimate(q.med, q.precision) for q in live) <= budget
