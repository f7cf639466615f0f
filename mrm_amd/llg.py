"""Host-side mirror of the reference's llg_m interface over the C ABI of libmrmg.so.

The reference's driver calls set_llg(...) to configure the integrator and llg(t) to run
the time loop of an `LLG` command (llg.f90:111-318, 1452-1487; cli.f90:889-896).  Here
the loop body runs on the GPU through mrmg_step; this module keeps the loop's host-side
bookkeeping: the step count of a fixed-dt run from del_t / t_max / min_it (the
termination test of llg.f90:277-281 with the convergence tests disabled, LLG_ERR 0), the
simulation time t, and the per-layer thermal streams.
"""
import ctypes as C

import numpy as np

from . import abi, formats


class Llg:
    """One simulation context (a grid, its layers, tensor and state) on the GPU."""

    def __init__(self, nx, ny, n_lay, dx, dy, dz, il_space=None, precision=abi.MRMG_FP64, demag=True, open=False):
        """open: open in-plane boundaries (the demag convolution over the media alone, on
        the grid zero-padded to 2nx x 2ny) instead of the reference's periodic ones."""
        self.lib = abi.load()
        self.geom = abi.make_geom(nx, ny, n_lay, dx, dy, dz, il_space, precision, demag)
        h = C.c_void_p()
        create = self.lib.mrmg_create_open if open else self.lib.mrmg_create
        abi.check(self.lib, create(C.byref(self.geom), C.byref(h)))
        self.ctx = h
        self.open = bool(open)
        self.nx, self.ny, self.n_lay = nx, ny, n_lay
        self.fx, self.fy = (2 * nx, 2 * ny) if open else (nx, ny)   # the FFT grid of the demag
        self.ngs = [0] * n_lay
        self._keep = {}
        # set_llg defaults (llg.f90:8-13, input.f90:386-412)
        self.dt = 1e-14
        self.type = abi.MRMG_TYPE_HEUN
        self.del_t = 1e22
        self.t_max = 1e22
        self.t_min = -1e22
        self.min_it = 10
        self.step_index = 0
        self.dt_frac = 0.0      # 0: fixed dt
        self.dt_max = 1e22
        self.max_dm = 180.0     # degrees
        self.llb = False        # llgp%llb

    # -- setup ---------------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mrmg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def set_tensor(self, n_real):
        n_real = np.ascontiguousarray(n_real, dtype=np.float64)
        abi.check(self.lib, self.lib.mrmg_set_tensor(self.ctx, abi.dptr(n_real)))

    def build_tensor(self, num_images=2, return_real=False, sul_mu=1.0, head_mu=1.0, fly_height=8e-7,
                     space_to_sul=40e-7):
        """calculate_tensor on the device (tensor.f90:82-216), with the SUL / head images
        when sul_mu / head_mu > 1.  The real tensor is on the FFT grid (fx, fy)."""
        out = None
        if return_real:
            out = np.zeros(self.fx * self.fy * 9 * self.n_lay * self.n_lay)
        if sul_mu > 1.0 or head_mu > 1.0:
            abi.check(self.lib, self.lib.mrmg_build_tensor_images(self.ctx, num_images, sul_mu, head_mu, fly_height,
                                                                  space_to_sul, abi.dptr(out)))
        else:
            abi.check(self.lib, self.lib.mrmg_build_tensor(self.ctx, num_images, abi.dptr(out)))
        return out

    def copy_tensor(self, src):
        """The tensor spectra of src (an Llg of the same precision, grid, boundaries and
        layer stack), copied on the device."""
        abi.check(self.lib, self.lib.mrmg_copy_tensor(self.ctx, src.ctx))

    def set_layer(self, k, la):
        s = la.struct()
        self._keep[("layer", k)] = la
        abi.check(self.lib, self.lib.mrmg_set_layer(self.ctx, k, C.byref(s)))
        self.ngs[k] = la.ng

    def set_media(self, media):
        for k, la in enumerate(media.layers):
            self.set_layer(k, la)

    def set_ufield(self, k, uf):
        abi.check(self.lib, self.lib.mrmg_set_ufield(self.ctx, k, C.byref(uf)))

    def set_source(self, k, slot, src):
        """A moving applied source of layer k (slot 0: field table, slot 1: temperature
        table or super-Gaussian; None clears): get_appl, applied.f90:827-1034."""
        st = src.struct() if src is not None else abi.Source(kind=abi.MRMG_SRC_NONE, t_end=1.0)
        abi.check(self.lib, self.lib.mrmg_set_source(self.ctx, k, slot, C.byref(st)))

    def set_state(self, k, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        abi.check(self.lib, self.lib.mrmg_set_state(self.ctx, k, abi.dptr(m)))

    def get_state(self, k):
        m = np.zeros((self.ngs[k], 3))
        h = np.zeros((self.ngs[k], 3))
        abi.check(self.lib, self.lib.mrmg_get_state(self.ctx, k, abi.dptr(m), abi.dptr(h)))
        return m, h

    def set_mag(self, k, mag):
        """|m| of layer k (data_s%mag, the LLB magnitude; 1 for LLG)."""
        mag = np.ascontiguousarray(mag, dtype=np.float64)
        assert mag.size == self.ngs[k]
        abi.check(self.lib, self.lib.mrmg_set_mag(self.ctx, k, abi.dptr(mag)))

    def get_mag(self, k):
        mag = np.zeros(self.ngs[k])
        abi.check(self.lib, self.lib.mrmg_get_mag(self.ctx, k, abi.dptr(mag)))
        return mag

    def profile(self, enable=True):
        abi.check(self.lib, self.lib.mrmg_profile(self.ctx, int(bool(enable))))

    def kernel_stats(self):
        arr = (abi.KStat * 16)()
        n = C.c_int(0)
        abi.check(self.lib, self.lib.mrmg_kernel_stats(self.ctx, arr, 16, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}

    def copy_probe(self, nbytes, reps=20):
        ms = C.c_double()
        abi.check(self.lib, self.lib.mrmg_copy_probe(self.ctx, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def memcpy_probe(self, nbytes, reps=20):
        """The same copy by hipMemcpyAsync device-to-device: average ms."""
        ms = C.c_double()
        abi.check(self.lib, self.lib.mrmg_memcpy_probe(self.ctx, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def time_kernel(self, name, reps=50):
        """Average duration (ms) of `reps` back-to-back launches of one demag kernel class."""
        ms = C.c_double()
        abi.check(self.lib, self.lib.mrmg_time_kernel(self.ctx, name.encode(), int(reps), C.byref(ms)))
        return ms.value

    def tune(self, key, value):
        abi.check(self.lib, self.lib.mrmg_tune(self.ctx, key.encode(), int(value)))

    def get_tune(self, key):
        v = C.c_int(0)
        abi.check(self.lib, self.lib.mrmg_get_tune(self.ctx, key.encode(), C.byref(v)))
        return v.value

    def footprint(self):
        b = C.c_size_t()
        abi.check(self.lib, self.lib.mrmg_footprint(self.ctx, C.byref(b)))
        return b.value

    # -- integrator ----------------------------------------------------------
    def set_llg(self, dt=None, llg_type=None, del_t=None, t_max=None, t_min=None, min_it=None, dt_frac=None,
                dt_max=None, max_dm=None, fixed_dt_p=None, llb=None):
        """set_llg (llg.f90:1452-1487): fixed dt unless dt_frac != 0 (then get_fields picks
        dt and push_m rejects steps rotating more than max_dm degrees); llb selects the
        Landau-Lifshitz-Bloch push (push_llb, llg.f90:916-1065)."""
        if llb is not None:
            self.llb = bool(llb)
        if dt_frac is not None:
            self.dt_frac = float(dt_frac)
        if fixed_dt_p:
            self.dt_frac = 0.0
        if dt_max is not None:
            self.dt_max = float(dt_max)
        if max_dm is not None:
            self.max_dm = float(max_dm)
        if dt is not None:
            self.dt = float(dt)
        if llg_type is not None:
            self.type = int(llg_type)
        if del_t is not None:
            self.del_t = float(del_t)
        if t_max is not None:
            self.t_max = float(t_max)
        if t_min is not None:
            self.t_min = float(t_min)
        if min_it is not None:
            self.min_it = max(1, int(min_it))

    def fixed_dt(self):
        """The fixed step get_fields takes: min(dt, dt_max) (fact0 = 1, llg.f90:383-386)."""
        return min(self.dt, self.dt_max) if self.dt_max > 0.0 else self.dt

    def steps_for(self, t):
        """Steps an `LLG` command takes from time t with fixed dt (llg.f90:268-281).  The
        convergence clause of cont_p is always true in the reference (get_applied returns
        .true., applied.f90:1033), so the time/step rule alone ends the loop."""
        it, del_t = 0, 0.0
        dt = self.fixed_dt()
        while True:
            t = t + dt
            del_t = del_t + dt
            it += 1
            if not ((del_t < self.del_t and t < max(self.t_max, self.t_min)) or it < self.min_it):
                return it

    def step(self, n_steps, t0=0.0, noise=None, philox_seed=None, diag=True):
        """Run n_steps fixed-dt steps from time t0.  noise: (n_steps, G, 3) normals to inject,
        or philox_seed to draw them on the device."""
        p = abi.StepParams(type=self.type, n_steps=int(n_steps), t0=float(t0), dt=self.dt,
                           noise=abi.MRMG_NOISE_NONE, seed=0, step0=self.step_index,
                           dt_frac=self.dt_frac, dt_max=self.dt_max, max_dm=self.max_dm, llb=int(self.llb))
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            assert nz.size == 3 * sum(self.ngs) * n_steps, "noise must be (n_steps, G, 3)"
            p.noise = abi.MRMG_NOISE_INJECT
        elif philox_seed is not None:
            p.noise = abi.MRMG_NOISE_PHILOX
            p.seed = int(philox_seed)
        d = abi.Diag() if diag else None
        abi.check(self.lib, self.lib.mrmg_step(self.ctx, C.byref(p), abi.dptr(nz), C.byref(d) if d else None))
        self.step_index += int(n_steps)
        if d is not None and self.dt_frac != 0.0:
            self.dt = d.dt          # llgp%dt carries over to the next call
        elif self.dt_frac == 0.0 and n_steps > 0:
            self.dt = self.fixed_dt()   # get_fields overwrote llgp%dt with the bounded step
        return d

    def llg(self, t, noise=None, philox_seed=None):
        """One `LLG` command: integrate from t until the termination test; returns (t, diag)."""
        n = self.steps_for(t)
        d = self.step(n, t0=t, noise=noise, philox_seed=philox_seed)
        return d.t, d

    # -- readback and switching record ----------------------------------------
    def readback(self, xpos, ypos, sigma, direction=(0.0, 0.0, 1.0), weights=None):
        """Reciprocity readback of the current state at head positions (xpos[i], ypos[i])
        with a Gaussian sensitivity of widths sigma = (sigma_x, sigma_y) (cm) along
        direction; weights: per-layer factors (spacing loss), default 1.  Returns emu."""
        xpos = np.ascontiguousarray(np.atleast_1d(xpos), dtype=np.float64)
        ypos = np.ascontiguousarray(np.broadcast_to(np.asarray(ypos, np.float64), xpos.shape))
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None and w.size != self.n_lay:
            raise ValueError("weights needs one value per layer")
        s = abi.Sensor(sigma_x=float(sigma[0]), sigma_y=float(sigma[1]), w=abi.dptr(w))
        for i in range(3):
            s.dir[i] = float(direction[i])
        out = np.zeros(xpos.size)
        abi.check(self.lib, self.lib.mrmg_readback(self.ctx, C.byref(s), xpos.size, abi.dptr(xpos), abi.dptr(ypos),
                                                   abi.dptr(out)))
        return out

    def switch_arm(self, direction=(0.0, 0.0, -1.0), threshold=0.0):
        """Start (and clear) the switching record: each later step marks the grains whose
        direction . m exceeds threshold for the first time."""
        d = np.ascontiguousarray(direction, dtype=np.float64)
        abi.check(self.lib, self.lib.mrmg_switch_arm(self.ctx, 1, abi.dptr(d), float(threshold)))

    def switch_stop(self):
        abi.check(self.lib, self.lib.mrmg_switch_arm(self.ctx, 0, None, 0.0))

    def switch_record(self, k):
        """(step, t, h_app) of layer k's grains: the step counter, end time and applied
        field (ng,3) of each grain's first crossing; step -1 where none happened."""
        step = np.zeros(self.ngs[k], np.int64)
        t = np.zeros(self.ngs[k])
        h = np.zeros((self.ngs[k], 3))
        abi.check(self.lib, self.lib.mrmg_switch_read(self.ctx, k, step.ctypes.data_as(C.POINTER(C.c_int64)),
                                                      abi.dptr(t), abi.dptr(h)))
        return step, t, h

    # -- checkpoints -----------------------------------------------------------
    def save_checkpoint(self, directory, t, iline, philox_seed=0, grain=True, ms_nom=None, generators=()):
        """A checkpoint save (checkpoint.f90:106-145) into directory: mrm_grain.chk when
        grain is true (the reference writes it once per run), the next of mrm_0.chk /
        mrm_1.chk, then mrm.chk.  The Philox stream (philox_seed and the step counter)
        goes in as one more random generator after generators (other save_random entries
        to carry, e.g. MT19937 states); the state, |m| and dt are also kept bit for bit.
        ms_nom: per-layer nominal Ms of the magnetisation header (default: mean grain Ms)."""
        layers = [self._keep[("layer", k)] for k in range(self.n_lay)]
        mag_layers, exact = [], []
        for k, la in enumerate(layers):
            m, _ = self.get_state(k)
            nom = float(ms_nom[k]) if ms_nom is not None else float(np.mean(la.ms))
            mag_layers.append((self.nx, self.ny, nom, m, la.ms))
            exact.append((m, self.get_mag(k)))
        gens = list(generators) + [formats.philox_generator(int(philox_seed), self.step_index)]
        dz = self.geom._keep[0]
        gl = formats.grain_dicts(layers, self.nx, self.ny, self.geom.dx, self.geom.dy, dz) if grain else None
        return formats.write_checkpoint(directory, t, iline, mag_layers, gens, gl,
                                        dict(layers=exact, dt=self.dt))

    def load_checkpoint(self, directory, physics=None):
        """A checkpoint read (checkpoint.f90:58-99): the state, the Philox counter and dt
        from the newest readable slot.  physics (one dict of LayerArrays keywords per layer,
        alpha, gamma, thermal, ...) also sets the layers from mrm_grain.chk, as the
        reference re-reads its media.  Returns read_checkpoint's dict with philox_seed
        added (None when the file holds no Philox stream); raises FileNotFoundError when no
        checkpoint is readable."""
        c = formats.read_checkpoint(directory, self.n_lay)
        if c is None:
            raise FileNotFoundError(f"no readable checkpoint in {directory}")
        if physics is not None:
            dz = self.geom._keep[0]
            for k, L in enumerate(c["grain"]):
                self.set_layer(k, formats.layer_arrays(L, dz[k], **physics[k]))
        for k in range(self.n_lay):
            if c["exact"] is not None:
                m, mag = c["exact"]["layers"][k]
                self.set_state(k, m)
                self.set_mag(k, mag)
            else:
                self.set_state(k, c["mag"][k][3])
        if c["exact"] is not None:
            self.dt = c["exact"]["dt"]
        c["philox_seed"] = None
        for g in c["generators"]:
            if g["name"] == formats.PHILOX_NAME:
                c["philox_seed"], self.step_index = formats.philox_state(g)
        return c

    def philox_normals(self, seed, step0, n_steps):
        out = np.zeros((n_steps, sum(self.ngs), 3))
        abi.check(self.lib, self.lib.mrmg_philox_normals(self.ctx, int(seed), int(step0), int(n_steps),
                                                         abi.dptr(out)))
        return out

    def demag_field(self):
        hc = np.zeros((3, self.n_lay, self.ny, self.nx))
        hg = np.zeros((sum(self.ngs), 3))
        abi.check(self.lib, self.lib.mrmg_demag_field(self.ctx, abi.dptr(hc), abi.dptr(hg)))
        return hc, hg


def step_batch(xs, n_steps, philox_seeds=None, t0s=None, graph=True):
    """n_steps fixed-dt steps of several contexts together (mrmg_step_batch): one launch of
    each demag kernel for all of them, one step captured as a hipGraph and replayed when
    graph is set.  xs: Llg contexts with equal grids, launch shapes and set_llg settings;
    philox_seeds: one per context (None: no thermal noise); t0s: start times (default 0).
    Each context's step counter advances as its own step() would advance it.  Returns the
    end times."""
    n = len(xs)
    if n == 0:
        return []
    lib = xs[0].lib
    ps = (abi.StepParams * n)()
    for b, x in enumerate(xs):
        if x.dt_frac != 0.0:
            raise ValueError("step_batch runs fixed-dt steps only")
        p = ps[b]
        p.type, p.n_steps, p.t0, p.dt = x.type, int(n_steps), float(t0s[b] if t0s is not None else 0.0), x.dt
        p.noise = abi.MRMG_NOISE_NONE if philox_seeds is None else abi.MRMG_NOISE_PHILOX
        p.seed = 0 if philox_seeds is None else int(philox_seeds[b])
        p.step0, p.dt_frac, p.dt_max, p.max_dm, p.llb = x.step_index, 0.0, x.dt_max, x.max_dm, int(x.llb)
    ctxs = (C.c_void_p * n)(*[x.ctx.value for x in xs])
    t_end = np.zeros(n)
    abi.check(lib, lib.mrmg_step_batch(ctxs, ps, n, int(bool(graph)), abi.dptr(t_end)))
    for x in xs:
        x.step_index += int(n_steps)
        if n_steps > 0:
            x.dt = x.fixed_dt()
    return list(t_end)
