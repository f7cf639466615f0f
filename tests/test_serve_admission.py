"""Admission of the request loop on the CPU: a stand-in context records the device bytes
every live context holds (admitted or idle in the pool), and the loop must keep their sum
within the budget at every allocation, reuse contexts of the same key and close idle ones,
least recently used first, when a new request does not fit.  A stand-in batched step
records which contexts it was given: a context is never in two batches at once, and a
batch holds contexts of one key only."""
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

    def build_tensor(self, num_images=2):
        pass

    def copy_tensor(self, src):
        assert src.key == self.key

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


class FakeBatch:
    """Stands in for mrm_amd.llg.step_batch: checks its members and advances their time."""

    def __init__(self):
        self.lock = threading.Lock()
        self.running = set()
        self.calls = []

    def __call__(self, xs, n, philox_seeds=None, t0s=None, graph=True):
        with self.lock:
            ids = {id(x) for x in xs}
            assert not (ids & self.running), "a context in two batches at once"
            assert len({x.key for x in xs}) == 1, "a batch of two grids"
            self.running |= ids
            self.calls.append(len(xs))
        with self.lock:
            self.running -= ids
        return [t + n * 5e-14 for t in (t0s or [0.0] * len(xs))]


def _requests(n, seed):
    rng = np.random.default_rng(seed)
    meds = {}
    reqs = []
    for i in range(n):
        nx = int(rng.choice([16, 24, 32]))
        if nx not in meds:
            meds[nx] = media.make_media(nx, nx, n_lay=1, seed=nx)
        med = meds[nx]
        reqs.append(serve.Request(rid=i, med=med, n_steps=int(rng.integers(1, 4)), seed=i, arrival=0.0))
    return reqs


def test_live_contexts_stay_within_the_budget(monkeypatch):
    dev = FakeDevice()
    FakeCtx.dev = dev
    reqs = _requests(40, 3)
    FakeCtx.sizes = {}
    for r in reqs:
        FakeCtx.sizes[(r.med.nx, r.med.ny, r.med.n_lay)] = serve.footprint_estimate(r.med, r.precision)
    monkeypatch.setattr(serve, "Llg", FakeCtx)
    fb = FakeBatch()
    monkeypatch.setattr(serve, "step_batch", fb)
    biggest = max(FakeCtx.sizes.values())
    budget = int(biggest * 2.2)           # room for two contexts, not for the pool of idle ones
    srv = serve.Server(workers=4, mem_budget=budget, chunk=2)
    srv.run(reqs)
    assert fb.calls and all(r.m for r in reqs)
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
    monkeypatch.setattr(serve, "step_batch", FakeBatch())
    budget = max(FakeCtx.sizes.values())
    srv = serve.Server(workers=3, mem_budget=budget)
    srv.run(reqs)
    assert dev.high <= budget
    assert len(dev.live) <= 1
    srv.close()
    assert not dev.live
