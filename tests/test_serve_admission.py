"""Admission of the request loop on the CPU: a stand-in context records the device bytes
every live context holds (in flight or idle in the pool), and the loop must keep their sum
within the budget at every allocation, reuse contexts of the same key and close idle ones,
least recently used first, when a new request does not fit."""
import threading

import numpy as np

from mrm_amd import media, serve


class FakeDevice:
    def __init__(self):
        self.lock = threading.Lock()
        self.live = {}          # id -> bytes
        self.high = 0
        self.created = 0


class FakeCtx:
    """Holds `bytes` of a pretend device from creation until close()."""
    dev = None
    sizes = {}

    def __init__(self, nx, ny, n_lay, dx, dy, dz, precision=1):
        self.key = (nx, ny, n_lay)
        self.bytes = FakeCtx.sizes[self.key]
        with self.dev.lock:
            self.dev.live[id(self)] = self.bytes
            self.dev.high = max(self.dev.high, sum(self.dev.live.values()))
            self.dev.created += 1
        self.ngs = [0] * n_lay
        self.step_index = 0

    def set_tensor(self, n):
        pass

    def set_media(self, med):
        self.ngs = list(med.ngs)

    def set_ufield(self, k, u):
        pass

    def set_state(self, k, m):
        pass

    def set_llg(self, dt=None):
        pass

    def step(self, n, philox_seed=None, diag=False):
        pass

    def footprint(self):
        return self.bytes

    def get_state(self, k):
        return np.zeros((self.ngs[k], 3)), None

    def close(self):
        with self.dev.lock:
            self.dev.live.pop(id(self), None)


def _requests(n, seed):
    rng = np.random.default_rng(seed)
    meds = {}
    reqs = []
    for i in range(n):
        nx = int(rng.choice([16, 24, 32]))
        if nx not in meds:
            meds[nx] = (media.make_media(nx, nx, n_lay=1, seed=nx), np.zeros(1))
        med, tensor = meds[nx]
        reqs.append(serve.Request(rid=i, med=med, tensor=tensor, n_steps=1, seed=i, arrival=0.0))
    return reqs


def test_live_contexts_stay_within_the_budget(monkeypatch):
    dev = FakeDevice()
    FakeCtx.dev = dev
    reqs = _requests(40, 3)
    FakeCtx.sizes = {}
    for r in reqs:
        FakeCtx.sizes[(r.med.nx, r.med.ny, r.med.n_lay)] = serve.footprint_estimate(r.med, r.precision)
    monkeypatch.setattr(serve, "Llg", FakeCtx)
    biggest = max(FakeCtx.sizes.values())
    budget = int(biggest * 2.2)           # room for two contexts, not for the pool of idle ones
    srv = serve.Server(slots=4, mem_budget=budget)
    srv.run(reqs)
    assert dev.high <= budget, (dev.high, budget)
    assert max(srv.admissions) <= budget
    # contexts were reused (fewer created than requests) and idle ones closed when needed
    assert dev.created < len(reqs)
    srv.close()
    assert not dev.live and srv.used == 0


def test_pooled_contexts_are_counted(monkeypatch):
    """With a budget of one context, a request of another grid must close the idle one
    first: the old accounting let idle pooled contexts pile up beyond the budget."""
    dev = FakeDevice()
    FakeCtx.dev = dev
    reqs = _requests(12, 5)
    FakeCtx.sizes = {(r.med.nx, r.med.ny, r.med.n_lay): serve.footprint_estimate(r.med, r.precision) for r in reqs}
    monkeypatch.setattr(serve, "Llg", FakeCtx)
    budget = max(FakeCtx.sizes.values())
    srv = serve.Server(slots=3, mem_budget=budget)
    srv.run(reqs)
    assert dev.high <= budget
    assert len(dev.live) <= 1
    srv.close()
    assert not dev.live
