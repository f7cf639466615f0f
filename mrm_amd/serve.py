"""A request loop over many independent simulations on one GPU, with continuous batching.

Requests (a media, a grid, a temperature, a number of steps) arrive over time.  A request
is admitted when its context's device footprint fits the memory budget; the idle contexts
kept for reuse count too, and are closed least recently used first when room is needed.
Admitted requests whose contexts share a grid, layer stack and precision form a group,
and a group's members are stepped together in chunks by mrmg_step_batch: each demag
kernel runs once for all of them and one step, captured as a hipGraph, is replayed.
Requests join their group between chunks and leave it when their steps are done, so a
group is never drained to let newcomers in (continuous batching); a group that is not
full may wait a batching window for more members before its first chunk.  Several worker
threads step different groups (or disjoint parts of one) at once, each batch on its lead
member's stream, so the work of different groups overlaps on the device.  A new context
of a grid that is already resident copies its tensor spectra on the device
(mrmg_copy_tensor) instead of building them.

Each request's result depends only on its own inputs (the thermal field is drawn from its
own seed and step counter, and a batched step gives each member what its own step would
give), so it is the same alone and in the mix, bit for bit.
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from mrm_amd import abi, media
from mrm_amd.llg import Llg, step_batch


@dataclass
class Request:
    rid: int
    med: object                 # mrm_amd.media.Media
    n_steps: int
    seed: int
    precision: int = abi.MRMG_FP32
    dt: float = 5e-14
    arrival: float = 0.0        # seconds after the start of the run
    tensor: np.ndarray = None   # a real-space tensor to use; None: built (or copied) on the device
    ufields: list = None
    # filled in by the server
    start: float = 0.0          # admission (its context set up and joining its group)
    end: float = 0.0
    footprint: int = 0          # device bytes its context held
    batches: int = 0            # batched calls it took part in
    m: list = field(default_factory=list)


def group_key(med, precision):
    """What decides a context's tensor and its batch: grid, cell size, layer stack, form."""
    il = tuple(getattr(med, "il_space", None) or ())
    return (med.nx, med.ny, med.n_lay, precision, float(med.dx), float(med.dy), tuple(med.dz), il)


def footprint_estimate(med, precision):
    """An upper bound of one context's device bytes, for admission: the tensor spectra with
    all n_lay^2 layer pairs and 9 slots, the spectra, per grain state and constants, per
    cell the grain map and at most one run (its run sums and lists), the exchange lists, the
    Philox normals drawn 16 steps ahead, and 2 MB of small tables."""
    P = 8 if precision == abi.MRMG_FP64 else 4
    nl, kc = med.n_lay, (med.nx // 2 + 1) * med.ny
    G, cells = med.G, med.cells
    E = 0
    for L in med.layers:
        for f in ("hxgrainindex", "uhxgrainindex", "lhxgrainindex"):
            v = getattr(L, f)
            if v is not None:
                E += int(v[-1])
    tensor = 9 * nl * nl * kc * 2 * P
    spectra = 2 * 3 * nl * kc * 2 * P
    grains = 640 * G + 3 * P * G
    per_cell = 4 * cells + 44 * cells
    exchange = 12 * E + 4 * (G + 1)
    noise = 16 * 3 * 8 * G
    return tensor + spectra + grains + per_cell + exchange + noise + (2 << 20)


class _Live:
    """An admitted request: its context, its remaining steps and its time."""

    def __init__(self, r, x, key, need):
        self.r, self.x, self.key, self.need = r, x, key, need
        self.left = r.n_steps
        self.t = 0.0
        self.busy = False       # being stepped, or lending its tensor
        self.ready = time.perf_counter()   # when it joined its group


class Server:
    """Continuous batching with admission by device footprint."""

    def __init__(self, workers=4, mem_budget=64 << 30, max_batch=16, chunk=200, graph=True, pool=8, admitters=4,
                 batch_wait=0.0):
        """batch_wait: seconds a group that is not full waits, from its oldest member's
        admission, for more members before it is stepped (the batching window)."""
        self.workers = workers
        self.batch_wait = batch_wait
        self.admitters = admitters
        self.budget = mem_budget
        self.max_batch = max_batch
        self.chunk = chunk
        self.graph = graph
        self.pool_max = pool
        self.used = 0
        self.peak = 0
        self.cv = threading.Condition()
        self.pool = {}            # key -> [(last use, bytes, context)] idle, still counted in used
        self.clock = 0
        self.admissions = []      # bytes of every live context at each admission
        self.live = []            # admitted, unfinished (_Live)
        self.batch_sizes = []

    # -- memory --------------------------------------------------------------------
    def _evict_lru(self):
        best = None
        for k, lst in self.pool.items():
            for i, (tl, need, x) in enumerate(lst):
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

    def _reserve(self, key, need):
        """Under the lock: a pooled context of key (its bytes re-counted as need) or None
        with need bytes reserved for a new one; waits for memory."""
        while True:
            lst = self.pool.get(key)
            if lst:
                _, old, x = lst[-1]
                if self.used + max(0, need - old) <= self.budget:
                    lst.pop()
                    if not lst:
                        del self.pool[key]
                    self.used += need - old
                    self._admitted()
                    return x
                if not self._evict_lru_except(x):
                    self.cv.wait()
                continue
            if self.used + need <= self.budget:
                self.used += need
                self._admitted()
                return None
            if not self._evict_lru():
                self.cv.wait()

    def _evict_lru_except(self, keep):
        best = None
        for k, lst in self.pool.items():
            for i, (tl, need, x) in enumerate(lst):
                if x is not keep and (best is None or tl < best[0]):
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

    def _admitted(self):
        self.peak = max(self.peak, self.used)
        self.admissions.append(self.used)

    def _release(self, lv):
        """Under the lock: keep the context idle in the pool (still counted) or close it."""
        if sum(len(v) for v in self.pool.values()) < self.pool_max:
            self.clock += 1
            self.pool.setdefault(lv.key, []).append((self.clock, lv.need, lv.x))
        else:
            lv.x.close()
            self.used -= lv.need
        self.cv.notify_all()

    # -- setup ---------------------------------------------------------------------
    def _context(self, r, key, pooled):
        if pooled is not None:
            return pooled
        med = r.med
        x = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=r.precision)
        try:
            if r.tensor is not None:
                x.set_tensor(r.tensor)
                return x
            donor = None
            with self.cv:
                for lv in self.live:
                    if lv.key == key and not lv.busy:
                        donor, lv.busy = lv, True
                        break
            if donor is not None:
                try:
                    x.copy_tensor(donor.x)
                finally:
                    with self.cv:
                        donor.busy = False
                        self.cv.notify_all()
            else:
                x.build_tensor(num_images=2)
            return x
        except Exception:
            x.close()
            raise

    def _setup(self, r, x):
        x.set_media(r.med)
        for k, u in enumerate(r.ufields or media.ramp_field(r.med.n_lay)):
            x.set_ufield(k, u)
        for k in range(r.med.n_lay):
            m = np.zeros((r.med.layers[k].ng, 3))
            m[:, 2] = 1.0
            x.set_state(k, m)
        x.step_index = 0
        x.set_llg(dt=r.dt)
        x.step(0, diag=False)      # the device build now, not in the first batched call

    # -- the loop ------------------------------------------------------------------
    def run(self, requests, log=None, every=20.0):
        """Serve the requests (sorted by arrival); returns wall seconds of the run.  log:
        called with a progress line every `every` seconds."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        t0 = time.perf_counter()
        errors = []
        state = {"admitting": True, "done": 0}
        finished = threading.Event()

        def monitor():
            while not finished.wait(every):
                with self.cv:
                    log(f"serving: {time.perf_counter() - t0:.0f} s, {len(self.live)} in flight, "
                        f"{state['done']} of {len(reqs)} done, {self.used / 1e9:.1f} GB admitted")

        nxt = [0]
        admitting = [self.admitters]

        def admit():
            while True:
                with self.cv:
                    if nxt[0] >= len(reqs):
                        break
                    r = reqs[nxt[0]]
                    nxt[0] += 1
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                key = group_key(r.med, r.precision)
                need = footprint_estimate(r.med, r.precision)
                if need > self.budget:
                    errors.append((r.rid, "larger than the memory budget"))
                    with self.cv:
                        state["done"] += 1
                        self.cv.notify_all()
                    continue
                with self.cv:
                    pooled = self._reserve(key, need)
                try:
                    x = self._context(r, key, pooled)
                    self._setup(r, x)
                except Exception as e:  # the request fails, the server goes on
                    errors.append((r.rid, str(e)))
                    with self.cv:
                        self.used -= need
                        state["done"] += 1
                        self.cv.notify_all()
                    continue
                r.start = time.perf_counter() - t0
                with self.cv:
                    self.live.append(_Live(r, x, key, need))
                    self.cv.notify_all()
            with self.cv:
                admitting[0] -= 1
                if admitting[0] == 0:
                    state["admitting"] = False
                self.cv.notify_all()

        def pick():
            """Under the lock: the members of the group whose oldest waiting member came
            first, among the groups that are full or whose oldest member has waited the
            batching window (or when admission is over); else ([], the next deadline)."""
            now = time.perf_counter()
            groups = {}
            for lv in self.live:
                if not lv.busy and lv.left > 0:
                    groups.setdefault(lv.key, []).append(lv)
            best, wake = None, None
            for g in groups.values():
                g.sort(key=lambda q: q.r.arrival)
                due = g[0].ready + self.batch_wait
                if len(g) >= self.max_batch or now >= due or not state["admitting"]:
                    if best is None or g[0].r.arrival < best[0].r.arrival:
                        best = g
                elif wake is None or due < wake:
                    wake = due
            if best is None:
                return [], wake
            group = best[:self.max_batch]
            for lv in group:
                lv.busy = True
            return group, None

        def work():
            while True:
                with self.cv:
                    while True:
                        group, wake = pick()
                        if group or (not state["admitting"] and not self.live):
                            break
                        self.cv.wait(None if wake is None else max(wake - time.perf_counter(), 1e-4))
                if not group:
                    return
                n = min(self.chunk, min(lv.left for lv in group))
                try:
                    ts = step_batch([lv.x for lv in group], n, philox_seeds=[lv.r.seed for lv in group],
                                    t0s=[lv.t for lv in group], graph=self.graph)
                except Exception as e:
                    with self.cv:
                        for lv in group:
                            errors.append((lv.r.rid, str(e)))
                            lv.left = 0
                            self.live.remove(lv)
                            lv.x.close()
                            self.used -= lv.need
                            state["done"] += 1
                        self.cv.notify_all()
                    continue
                finished = []
                for lv, t in zip(group, ts):
                    lv.t = t
                    lv.left -= n
                    lv.r.batches += 1
                    if lv.left == 0:
                        lv.r.m = [lv.x.get_state(k)[0] for k in range(lv.r.med.n_lay)]
                        lv.r.footprint = lv.x.footprint()
                        lv.r.end = time.perf_counter() - t0
                        finished.append(lv)
                with self.cv:
                    self.batch_sizes.append(len(group))
                    for lv in group:
                        lv.busy = False
                    for lv in finished:
                        self.live.remove(lv)
                        self._release(lv)
                        state["done"] += 1
                    self.cv.notify_all()

        th = [threading.Thread(target=admit) for _ in range(self.admitters)]
        th += [threading.Thread(target=work) for _ in range(self.workers)]
        mon = threading.Thread(target=monitor) if log else None
        for t in th:
            t.start()
        if mon:
            mon.start()
        for t in th:
            t.join()
        finished.set()
        if mon:
            mon.join()
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


FFT_SIZES = (64, 128, 256, 512, 1024)


def fft_round(n):
    """The nearest power of two in 64..1024 (log scale): requests of one rounded grid share
    their padded size and so can be stepped in one batch."""
    return min(FFT_SIZES, key=lambda s: abs(np.log(s / n)))


def _media(args):
    """One request's media (a process-pool task)."""
    nx, ny, nl, T, seed = args
    physical = [0, 0, 1, 1] if nl == 4 else ([0, 1] if nl == 2 else [0])
    layer_ku = [2e6, 2e6, 1e7, 1e7] if nl == 4 else ([2e6, 1e7] if nl == 2 else [6e6])
    return media.make_media(nx, ny, dz=[2e-7 * 4 / nl] * nl, n_lay=nl, layer_ku=layer_ku, physical=physical,
                            temperature=T, thermal=T > 0, seed=seed)


def build_load(n_requests, rate, sizes=(64, 1024), steps=(500, 5000), seed=11, precision=abi.MRMG_FP32, workers=1):
    """A Poisson stream (rate per second) of config-1-style requests: Nx, Ny log-uniform
    in sizes rounded to powers of two, n_lay in {1, 2, 4} over 4 nm, an independent media
    seed, T in {0, 300, 400} K and a uniform number of steps.  The tensors are built on the
    device by the server (one per grid, then copied).  workers > 1 generates the media in
    that many forked processes: call it before this process first uses the GPU."""
    rng = np.random.default_rng(seed)
    t = 0.0
    lo, hi = np.log(sizes[0]), np.log(sizes[1])
    draws = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        nx = fft_round(np.exp(rng.uniform(lo, hi)))
        ny = fft_round(np.exp(rng.uniform(lo, hi)))
        nl = int(rng.choice([1, 2, 4]))
        T = float(rng.choice([0.0, 300.0, 400.0]))
        mseed = int(rng.integers(1, 1 << 30))
        n = int(rng.integers(steps[0], steps[1] + 1))
        draws.append(((nx, ny, nl, T, mseed), n, int(rng.integers(1, 1 << 30)), t))
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            meds = pool.map(_media, [d[0] for d in draws], chunksize=4)
    else:
        meds = [_media(d[0]) for d in draws]
    return [Request(rid=i, med=m, n_steps=d[1], seed=d[2], precision=precision, arrival=d[3])
            for i, (d, m) in enumerate(zip(draws, meds))]


def latency_stats(reqs, wall):
    lat = np.array([r.end - r.arrival for r in reqs])
    work = sum(r.med.cells * r.n_steps for r in reqs)
    return {"requests": len(reqs), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "wall_s": wall, "cell_updates_per_s": work / wall}
