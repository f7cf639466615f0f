"""A request loop over many independent simulations on one GPU.

Requests (a media, a grid, a temperature, a number of steps) arrive over time.  Worker
threads take them in arrival order as slots and device memory free up: a request is
admitted only when its context's device footprint fits the memory budget left.  Each
worker runs its request on its own context, and therefore on its own HIP stream, so the
requests in flight overlap on the device.  A context is kept after a request finishes
and reused by the next request of the same grid and precision, so its tensor spectra
and buffers stay resident.  The library releases the GIL during its calls (ctypes), so
the workers' device work proceeds in parallel.

Each request's result depends only on its own inputs (the thermal field is drawn from
its own seed, step and grain), so it is the same whether it runs alone or in the mix.
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from mrm_amd import abi, media
from mrm_amd.llg import Llg


@dataclass
class Request:
    rid: int
    med: object                 # mrm_amd.media.Media
    tensor: np.ndarray          # its real-space tensor (shared between requests of one grid)
    n_steps: int
    seed: int
    precision: int = abi.MRMG_FP32
    dt: float = 5e-14
    arrival: float = 0.0        # seconds after the start of the run
    ufields: list = None
    # filled in by the server
    start: float = 0.0
    end: float = 0.0
    footprint: int = 0          # device bytes its context held
    m: list = field(default_factory=list)


def footprint_estimate(med, precision):
    """Device bytes of one context (tensor spectra dominate), for admission."""
    P = 8 if precision == abi.MRMG_FP64 else 4
    nl, kc = med.n_lay, (med.nx // 2 + 1) * med.ny
    tensor = 6 * nl * nl * kc * 2 * P
    spectra = 2 * 3 * nl * kc * 2 * P
    cells = 2 * 3 * med.cells * P
    grains = 40 * 8 * med.G          # grain state and constants are FP64 in both forms
    noise = 16 * 3 * 8 * med.G       # Philox normals drawn 16 steps ahead
    return tensor + spectra + cells + grains + noise + (1 << 20)


class Server:
    """Admission by device footprint.  `used` counts every live context: the requests in
    flight and the idle contexts kept in the pool for reuse.  A request that does not fit
    first closes idle contexts, least recently used first, then waits for a running one
    to finish."""

    def __init__(self, slots=8, mem_budget=64 << 30):
        self.slots = slots
        self.budget = mem_budget
        self.used = 0
        self.peak = 0
        self.cv = threading.Condition()
        self.pool = {}            # context key -> [(last use, estimate, context)] idle
        self.clock = 0
        self.admissions = []      # (bytes of every live context) at each admission

    @staticmethod
    def key(r):
        """The grid, cell size, layer stack and precision decide a context's tensor and
        buffers; the tensor's identity covers anything else it was built from."""
        m = r.med
        il = tuple(getattr(m, "il_space", None) or ())
        return (m.nx, m.ny, m.n_lay, r.precision, float(m.dx), float(m.dy), tuple(m.dz), il, id(r.tensor))

    def _idle(self):
        return sum(len(v) for v in self.pool.values())

    def _evict_lru(self, keep=None):
        """Close the least recently used idle context other than `keep` (under the lock);
        False if there is none."""
        best = None
        for k, lst in self.pool.items():
            for i, (tl, need, x) in enumerate(lst):
                if x is keep:
                    continue
                if best is None or tl < best[0]:
                    best = (tl, k, i)
        if best is None:
            return False
        _, k, i = best
        _, need, x = self.pool[k].pop(i)
        if not self.pool[k]:
            del self.pool[k]
        x.close()
        self.used -= need
        return True

    def _admit(self, r, need):
        """Take a pooled context of r's key (its bytes are already counted) or reserve
        `need` bytes for a new one.  Returns the pooled context or None."""
        key = self.key(r)
        with self.cv:
            while True:
                lst = self.pool.get(key)
                if lst:
                    # a pooled context counts its own estimate; the request's may be larger
                    # (more grains), so the difference must fit as well
                    _, old, x = lst[-1]
                    if self.used + max(0, need - old) <= self.budget:
                        lst.pop()
                        if not lst:
                            del self.pool[key]
                        self.used += need - old
                        self.peak = max(self.peak, self.used)
                        self.admissions.append(self.used)
                        return x
                    if not self._evict_lru(keep=x):
                        self.cv.wait()
                    continue
                if self.used + need <= self.budget:
                    self.used += need
                    self.peak = max(self.peak, self.used)
                    self.admissions.append(self.used)
                    return None
                if not self._evict_lru():
                    self.cv.wait()

    def _context(self, r, x):
        if x is not None:
            return x
        med = r.med
        x = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=r.precision)
        x.set_tensor(r.tensor)
        return x

    def _release(self, r, x, need):
        """Keep the context idle in the pool (still counted) or close it."""
        with self.cv:
            if x is not None and self._idle() < self.slots:
                self.clock += 1
                self.pool.setdefault(self.key(r), []).append((self.clock, need, x))
                x = None
            else:
                self.used -= need
            self.cv.notify_all()
        if x is not None:
            x.close()

    def _run(self, r, x, t0):
        x.set_media(r.med)
        for k, u in enumerate(r.ufields or media.ramp_field(r.med.n_lay)):
            x.set_ufield(k, u)
        for k in range(r.med.n_lay):
            m = np.zeros((r.med.layers[k].ng, 3))
            m[:, 2] = 1.0
            x.set_state(k, m)
        x.step_index = 0
        x.set_llg(dt=r.dt)
        x.step(r.n_steps, philox_seed=r.seed, diag=False)
        r.footprint = x.footprint()
        r.m = [x.get_state(k)[0] for k in range(r.med.n_lay)]
        r.end = time.perf_counter() - t0

    def run(self, requests):
        """Serve the requests (sorted by arrival); returns wall seconds of the run."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        t0 = time.perf_counter()
        nxt = [0]
        errors = []

        def worker():
            while True:
                with self.cv:
                    if nxt[0] >= len(reqs):
                        return
                    r = reqs[nxt[0]]
                    nxt[0] += 1
                need = footprint_estimate(r.med, r.precision)
                if need > self.budget:
                    errors.append((r.rid, "larger than the memory budget"))
                    continue
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                pooled = self._admit(r, need)
                r.start = time.perf_counter() - t0
                x = None
                try:
                    x = self._context(r, pooled)
                    self._run(r, x, t0)
                except Exception as e:  # the request fails, the server goes on
                    errors.append((r.rid, str(e)))
                    if x is not None:
                        x.close()
                        x = None
                finally:
                    self._release(r, x, need)

        th = [threading.Thread(target=worker) for _ in range(self.slots)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errors:
            raise RuntimeError(f"requests failed: {errors[:3]}")
        return time.perf_counter() - t0

    def close(self):
        with self.cv:
            for lst in self.pool.values():
                for _, need, x in lst:
                    x.close()
                    self.used -= need
            self.pool = {}


FFT_SIZES = sorted({2 ** a * 3 ** b * 5 ** c for a in range(11) for b in range(4) for c in range(3)
                    if 64 <= 2 ** a * 3 ** b * 5 ** c <= 1024})


def fft_round(n):
    return min(FFT_SIZES, key=lambda s: abs(np.log(s / n)))


def build_load(n_requests, rate, sizes=(64, 512), steps=(500, 5000), seed=11, precision=abi.MRMG_FP32,
               tensor_of=None):
    """A Poisson stream (rate per second) of config-1-style requests: Nx, Ny log-uniform
    in sizes (rounded to 2^a 3^b 5^c), n_lay in {1, 2, 4} over 4 nm, an independent media
    seed, T in {0, 300, 400} K and a uniform number of steps.  tensor_of(med) gives the
    real-space tensor of a grid (built once per grid)."""
    tensor_of = tensor_of or device_tensor
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs, tensors = [], {}
    lo, hi = np.log(sizes[0]), np.log(sizes[1])
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        nx = fft_round(np.exp(rng.uniform(lo, hi)))
        ny = fft_round(np.exp(rng.uniform(lo, hi)))
        nl = int(rng.choice([1, 2, 4]))
        T = float(rng.choice([0.0, 300.0, 400.0]))
        physical = [0, 0, 1, 1] if nl == 4 else ([0, 1] if nl == 2 else [0])
        layer_ku = [2e6, 2e6, 1e7, 1e7] if nl == 4 else ([2e6, 1e7] if nl == 2 else [6e6])
        med = media.make_media(nx, ny, dz=[2e-7 * 4 / nl] * nl, n_lay=nl, layer_ku=layer_ku, physical=physical,
                               temperature=T, thermal=T > 0, seed=int(rng.integers(1, 1 << 30)))
        key = (nx, ny, nl)
        if key not in tensors:
            tensors[key] = tensor_of(med)
        reqs.append(Request(rid=i, med=med, tensor=tensors[key], n_steps=int(rng.integers(steps[0], steps[1] + 1)),
                            seed=int(rng.integers(1, 1 << 30)), precision=precision, arrival=t))
    return reqs


def device_tensor(med):
    """The FP64 tensor of a media's grid built on the device (num_images 2)."""
    x = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz)
    try:
        return x.build_tensor(num_images=2, return_real=True)
    finally:
        x.close()


def latency_stats(reqs, wall):
    lat = np.array([r.end - r.arrival for r in reqs])
    work = sum(r.med.cells * r.n_steps for r in reqs)
    return {"requests": len(reqs), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "wall_s": wall, "cell_updates_per_s": work / wall}
