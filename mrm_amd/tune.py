"""Launch-shape autotuner for a live simulation context.

The reference asks FFTW to plan its transforms exhaustively before the first step
(fftw.f90:25, FFTW_EXHAUSTIVE); this is the same idea for the HIP kernels: time the
candidate launch shapes of each kernel class on the real workload with HIP events
(mrmg_profile) and keep the fastest.  The simulation state, time and thermal counter
are restored afterwards, so tuning does not change the trajectory.
"""
import numpy as np

DEMAG = ("row_fwd", "col", "row_inv", "col_fwd", "prod", "col_inv")
SPACE = {
    "col_path": ([0, 1, 2, 3], DEMAG),        # fused column kernel, split path, fused on column-contiguous spectra,
                                              # split forward column FFTs then fused product + inverse
    "row_block": ([2, 4, 8], DEMAG),          # rows per transposing row workgroup (paths 1, 2)
    "wave_fft": ([0, 1], DEMAG),              # wave-local FFTs in the transposing row and path-2 column kernels
    "row_threads": ([64, 128, 256], ("row_fwd", "row_inv")),
    "row_radix": ([4, 8, 16], ("row_fwd", "row_inv")),   # largest Stockham radix of the row FFTs
    "col_radix": ([4, 8, 16], ("col", "col_fwd", "col_inv")),
    "col_threads": ([256, 512, 1024], ("col",)),
    "col_groups": (None, ("col",)),          # 1..min(n_lay, 4)
    "lanes_per_grain": ([1, 2, 4, 8, 16], ("grain1", "grain2")),
}


def _cost(x, steps, classes, philox_seed):
    x.profile(True)
    x.step(steps, philox_seed=philox_seed, diag=False)
    st = x.kernel_stats()
    x.profile(False)
    return sum(st[k][1] / max(st[k][0], 1) for k in classes if k in st)


ROW = ("row_fwd", "row_inv")


def _try(x, settings, classes, steps, philox_seed):
    try:
        for k, v in settings.items():
            x.tune(k, v)
    except Exception:
        return None
    x.step(2, philox_seed=philox_seed, diag=False)   # warm the shape
    return _cost(x, steps, classes, philox_seed)


def autotune(x, steps=10, philox_seed=12345, verbose=False):
    """Choose every tunable of context x (an mrm_amd.llg.Llg).  Returns {key: value}.
    The column path, the rows per transposing row workgroup, the FFT form, the row
    threads and the row radix are searched together (the paths use different row
    kernels); the others one at a time."""
    saved = [x.get_state(k)[0] for k in range(x.n_lay)]
    saved_mag = [x.get_mag(k) for k in range(x.n_lay)]
    step_index = x.step_index
    chosen = {}
    best, best_s = None, None
    joint = ("col_path", "row_block", "wave_fft", "row_threads", "row_radix")
    for path in SPACE["col_path"][0]:
        for rb in (SPACE["row_block"][0] if path != 0 else [x.get_tune("row_block")]):
            for wv in (SPACE["wave_fft"][0] if path != 0 else [0]):
                for rt in SPACE["row_threads"][0]:
                    for rr in SPACE["row_radix"][0]:
                        st = {"col_path": path, "row_block": rb, "wave_fft": wv, "row_threads": rt, "row_radix": rr}
                        c = _try(x, st, DEMAG, steps, philox_seed)
                        if verbose:
                            print(f"  tune {st}: {c if c is None else round(c * 1e3, 2)} us", flush=True)
                        if c is not None and (best is None or c < best):
                            best, best_s = c, st
    if best_s is not None:
        for k, v in best_s.items():
            x.tune(k, v)
        chosen.update(best_s)
    # coordinate sweeps over every tunable (the joint ones too): catches the interactions
    # the joint search leaves out, e.g. the column radix under wave-local FFTs
    for sweep in range(2):
        for key, (values, classes) in SPACE.items():
            if values is None:
                values = list(range(1, min(x.n_lay, 4) + 1))
            if key in joint:
                classes = DEMAG
            best, best_v = None, x.get_tune(key)
            for v in values:
                try:
                    x.tune(key, v)
                except Exception:
                    continue
                x.step(2, philox_seed=philox_seed, diag=False)   # warm the shape
                c = _cost(x, steps, classes, philox_seed)
                if verbose:
                    print(f"  tune {key}={v}: {c * 1e3:.2f} us", flush=True)
                if best is None or c < best:
                    best, best_v = c, v
            x.tune(key, best_v)
            chosen[key] = best_v
    for k, m in enumerate(saved):
        x.set_state(k, m)
        x.set_mag(k, saved_mag[k])
    x.step_index = step_index
    return chosen


def tune_keys(x, space, steps=3, philox_seed=12345):
    """Time each listed value of the given tunables ({key: values}) on context x, one key
    at a time over the demag classes, and keep the fastest; the state, time and thermal
    counter are restored as in autotune.  Returns {key: value}."""
    saved = [x.get_state(k)[0] for k in range(x.n_lay)]
    saved_mag = [x.get_mag(k) for k in range(x.n_lay)]
    step_index = x.step_index
    chosen = {}
    for key, values in space.items():
        best, best_v = None, x.get_tune(key)
        for v in values:
            c = _try(x, {key: v}, DEMAG, steps, philox_seed)
            if c is not None and (best is None or c < best):
                best, best_v = c, v
        x.tune(key, best_v)
        chosen[key] = best_v
    for k, m in enumerate(saved):
        x.set_state(k, m)
        x.set_mag(k, saved_mag[k])
    x.step_index = step_index
    return chosen


def state_equal(a, b):
    return all(np.array_equal(p, q) for p, q in zip(a, b))
