"""Benchmark of the granular-media time step on one MI355X.

Workload (configuration 1): 256 x 256 x 4 cells of 2 nm, two exchange-coupled 4 nm
physical layers (Ku 2e6 top, 1e7 bottom, Ms 700, Voronoi grains of 8 nm pitch with
one-cell boundaries), T = 300 K thermal field from the on-device Philox source, uniform
field tilted 5 deg from -z ramped 0 -> 20 kOe over 2 ns, Heun at dt = 50 fs, periodic FFT
demag.  A step is one full Heun step (two field evaluations, two pushes) of all cells.

value = cells * steps / wall seconds of the FP64 plain form with every input resident on
the device.  The FP32 precision form is timed on the same workload and compared with it.
The JSON line also carries the roofline of the dominant kernel (HIP-event durations on
the library's stream, algorithmic bytes per launch) and a CPU baseline (the oracle, the
C/OpenMP restatement of the reference path, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from mrm_amd import abi, media  # noqa: E402
from mrm_amd.llg import Llg  # noqa: E402
from mrm_amd.tune import autotune  # noqa: E402

METRIC = "LLG cell-updates/s FP32 vs FP64 at max|Δm|≤1e-4; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
DT = 5e-14
SEED = 1


def build_ctx(med, precision, tensor=None):
    x = Llg(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz, precision=precision)
    N = None
    if tensor is None:
        N = x.build_tensor(num_images=2, return_real=True)
    else:
        x.set_tensor(tensor)
    x.set_media(med)
    for k, u in enumerate(media.ramp_field(med.n_lay)):
        x.set_ufield(k, u)
    x.set_llg(dt=DT)
    return x, N


def timed(x, steps, t0):
    t = time.perf_counter()
    d = x.step(steps, t0=t0, philox_seed=SEED)     # returns after the device work is done
    return time.perf_counter() - t, d


def col_bytes(med, P):
    """Algorithmic bytes of one column launch: spectra in and out (12 n_lay Kc P) plus the
    unique tensor (T_u Kc 2P, T_u = 6 n_lay (n_lay+1)/2)."""
    nl = med.n_lay
    kc = (med.nx // 2 + 1) * med.ny
    tu = 6 * nl * (nl + 1) // 2
    return 12 * nl * kc * P + tu * kc * 2 * P


def kernel_bytes(med, P):
    """Algorithmic bytes per launch of each kernel class: the row r2c reads the cell
    magnetisation and writes the spectra; the column kernel see col_bytes; the row c2r
    reads the spectra and writes the cell field; a grain stage reads ~25 P per grain
    (state, constants, dm, thermal), the exchange CSRs (4 + P per entry), the cell index
    map (4 per magnetic cell) and gathers/scatters the cell field/magnetisation (6 P per
    magnetic cell)."""
    nl, C = med.n_lay, med.cells
    kc = (med.nx // 2 + 1) * med.ny
    G = med.G
    cm = sum(int(L.grain2gridindex[-1]) for L in med.layers)
    E = 0
    for L in med.layers:
        for f in ("hxgrainindex", "uhxgrainindex", "lhxgrainindex"):
            v = getattr(L, f)
            if v is not None:
                E += int(v[-1])
    grain = 25 * P * G + E * (4 + P) + 4 * cm + 6 * cm * P
    tu = 6 * nl * (nl + 1) // 2
    return {
        "row_fwd": 3 * C * P + 6 * nl * kc * P,
        "col": col_bytes(med, P),
        "row_inv": 6 * nl * kc * P + 3 * C * P,
        "grain1": grain,
        "grain2": grain,
        # split column path: column FFTs in place, product streaming spectra and tensor
        "col_fwd": 12 * nl * kc * P,
        "prod": 12 * nl * kc * P + 2 * tu * kc * P,
        "col_inv": 12 * nl * kc * P,
    }


def pmc_traffic(kernel, tag):
    """HBM bytes per launch of a kernel class from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{tag}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == kernel:
            try:
                return int(float(f[3]))
            except ValueError:
                return None
    return None


def max_rms(a, b, n_lay):
    dmax, ss, n = 0.0, 0.0, 0
    for k in range(n_lay):
        d = np.abs(a.get_state(k)[0] - b.get_state(k)[0])
        dmax = max(dmax, float(d.max()))
        ss += float((d ** 2).sum())
        n += d.size
    return dmax, (ss / max(n, 1)) ** 0.5


def accuracy(med, N, steps, log):
    """The FP32 form against the FP64 form after `steps` steps from the same start:
    deterministic (T = 0, same media) and thermal (T = 300 K, the same Philox noise).  The
    thermal runs switch grains by activation, which is chaotic: the control compares two
    FP64 runs that differ only in the summation order of the cell gathers (4 instead of 8
    lanes per grain), and `switched` counts grains whose final m_z sign differs."""
    out = {"steps": steps}

    def run(prec, thermal, lanes=None):
        m = media.config_media("cfg1", thermal=thermal, temperature=300.0 if thermal else 0.0)
        x = build_ctx(m, prec, tensor=N)[0]
        if lanes:
            x.tune("lanes_per_grain", lanes)
        x.step(steps, philox_seed=SEED, diag=False)
        return x, m

    def cmp(a, b, m):
        dmax, drms = max_rms(a, b, m.n_lay)
        sw = sum(int((np.sign(a.get_state(k)[0][:, 2]) != np.sign(b.get_state(k)[0][:, 2])).sum())
                 for k in range(m.n_lay))
        return {"max_dm": dmax, "rms_dm": drms, "switched": sw, "grains": m.G}

    a, m = run(abi.MRMG_FP64, False)
    b, _ = run(abi.MRMG_FP32, False)
    out["T0_fp32_vs_fp64"] = cmp(a, b, m)
    for x in (a, b):
        x.close()
    a, m = run(abi.MRMG_FP64, True)
    b, _ = run(abi.MRMG_FP32, True)
    c, _ = run(abi.MRMG_FP64, True, lanes=4)
    out["T300_fp32_vs_fp64"] = cmp(a, b, m)
    out["T300_fp64_vs_fp64_reordered"] = cmp(a, c, m)
    for x in (a, b, c):
        x.close()
    log(f"accuracy after {steps} steps: T=0 fp32 vs fp64 max|dm| {out['T0_fp32_vs_fp64']['max_dm']:.2e}; "
        f"T=300 fp32 vs fp64 {out['T300_fp32_vs_fp64']['max_dm']:.2e} "
        f"({out['T300_fp32_vs_fp64']['switched']} grains switched differently), "
        f"fp64 vs fp64 reordered {out['T300_fp64_vs_fp64_reordered']['max_dm']:.2e} "
        f"({out['T300_fp64_vs_fp64_reordered']['switched']})")
    return out


def track_form(log, steps=100, warmup=10):
    """The 512 x 4096 x 4 recording-track grid (1.5 nm cells, the two-layer media, T=300 K):
    both forms timed, and compared after the same steps."""
    med = media.config_media("cfg2")
    x64, N = build_ctx(med, abi.MRMG_FP64)
    x32, _ = build_ctx(med, abi.MRMG_FP32, tensor=N)
    del N
    out = {"workload": f"{med.nx}x{med.ny}x{med.n_lay} cells of 1.5 nm, {med.G} grains, T=300 K, Heun dt=50 fs",
           "steps": steps}
    for name, x in (("fp64", x64), ("fp32", x32)):
        shapes = autotune(x, steps=3)
        x.step(warmup, philox_seed=SEED, diag=False)
        el, _ = timed(x, steps, warmup * DT)
        out[name] = {"value": med.cells * steps / el, "ms_per_step": el * 1e3 / steps, "launch_shapes": shapes}
        log(f"track grid {name}: {el * 1e3 / steps:.3f} ms/step, {med.cells * steps / el:.4e} cell-updates/s")
        # per-kernel event-pair durations and algorithmic bandwidth in this bandwidth-bound regime
        x.profile(True)
        x.step(20, t0=(warmup + steps) * DT, philox_seed=SEED, diag=False)
        st = x.kernel_stats()
        x.profile(False)
        kb = kernel_bytes(med, 8 if name == "fp64" else 4)
        ks = {}
        for k, (n, ms) in st.items():
            if n > 0:
                us = 1e3 * ms / n
                ks[k] = {"avg_us": round(us, 2), "gbs": round(kb[k] / (us * 1e-6) / 1e9, 1),
                         "frac": round(kb[k] / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 3)}
        out[name]["kernels"] = ks
        log(f"track grid {name} kernels: " + ", ".join(f"{k} {v['avg_us']}us {v['frac']}" for k, v in ks.items()))
    out["speedup_fp32"] = out["fp32"]["value"] / out["fp64"]["value"]
    dmax, drms = max_rms(x64, x32, med.n_lay)
    out["max_dm_fp32_vs_fp64"], out["rms_dm_fp32_vs_fp64"] = dmax, drms
    out["steps_compared"] = warmup + steps
    log(f"track grid fp32 x{out['speedup_fp32']:.2f}, max|dm| {dmax:.2e} rms {drms:.2e}")
    x64.close()
    x32.close()
    return out


def serving(log, n_requests=40, rate=40.0):
    """The request loop under a Poisson load of config-1-style requests (FP32 form)."""
    from mrm_amd import serve
    reqs = serve.build_load(n_requests, rate=rate, sizes=(64, 384), steps=(200, 2000), seed=11)
    srv = serve.Server(slots=8)
    wall = srv.run(reqs)
    st = serve.latency_stats(reqs, wall)
    st["peak_admitted_bytes"] = srv.peak
    st["load"] = (f"{n_requests} requests, Poisson {rate}/s, Nx, Ny log-uniform 64-384 (2^a 3^b 5^c), n_lay 1/2/4, "
                  "T 0/300/400 K, 200-2000 steps, 8 slots, FP32 form")
    srv.close()
    log(f"serving: p50 {st['p50_s']:.3f} s, p99 {st['p99_s']:.3f} s, {st['cell_updates_per_s']:.3e} cell-updates/s")
    return st


def cpu_baseline(med, tensor, target_s=15.0):
    """The oracle (C restatement, OpenMP) on the same workload: a bounded sample of steps."""
    from oracle import oracle as O
    g = abi.make_geom(med.nx, med.ny, med.n_lay, med.dx, med.dy, med.dz)
    s = O.Sim(g)
    s.set_tensor(tensor)
    for k, L in enumerate(med.layers):
        s.set_layer(k, L)
    for k, u in enumerate(media.ramp_field(med.n_lay)):
        s.set_ufield(k, u)
    rng = np.random.default_rng(3)
    probe = 2
    noise = rng.standard_normal((probe, med.G, 3))
    t = time.perf_counter()
    s.step(probe, DT, noise=noise)
    per = (time.perf_counter() - t) / probe
    n = int(max(2, min(400, target_s / max(per, 1e-6))))
    noise = rng.standard_normal((n, med.G, 3))
    t = time.perf_counter()
    s.step(n, DT, t0=probe * DT, noise=noise)
    el = time.perf_counter() - t
    return {"value": med.cells * n / el, "unit": "cell-updates/s", "cores": int(O.lib().ora_num_threads()),
            "kind": "port",
            "sample": f"{n} FP64 Heun steps of the same 256x256x4 workload (injected normals), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=10000)
    ap.add_argument("--no-track", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2
    log = (lambda *x: None) if a.quiet else (lambda *x: print(*x, file=sys.stderr, flush=True))

    med = media.config_media("cfg1")
    log(f"workload cfg1: {med.nx}x{med.ny}x{med.n_lay} cells, {med.G} grains")
    x64, N = build_ctx(med, abi.MRMG_FP64)
    shapes64 = autotune(x64)
    log(f"fp64 launch shapes: {shapes64}")
    x64.step(a.warmup, philox_seed=SEED)
    t0 = a.warmup * DT
    el64, _ = timed(x64, a.steps, t0)
    v64 = med.cells * a.steps / el64
    log(f"fp64: {el64 * 1e3 / a.steps:.4f} ms/step, {v64:.4e} cell-updates/s")

    # per-kernel durations with HIP events on the library's stream, same workload
    x64.profile(True)
    x64.step(a.steps, t0=t0 + a.steps * DT, philox_seed=SEED)
    st = x64.kernel_stats()
    x64.profile(False)
    kb = kernel_bytes(med, 8)
    kern = {k: {"launches": n, "avg_us": 1e3 * ms / max(n, 1), "share": 0.0} for k, (n, ms) in st.items() if n > 0}
    tot = sum(ms for n, ms in st.values() if n > 0)
    for k in kern:
        kern[k]["share"] = st[k][1] / tot if tot else 0.0
        kern[k]["gbs"] = kb[k] / (kern[k]["avg_us"] * 1e-6) / 1e9 if kern[k]["avg_us"] > 0 else 0.0
    dom = max(kern, key=lambda k: kern[k]["share"])
    log("kernels: " + ", ".join(f"{k} {v['avg_us']:.2f}us {v['share'] * 100:.0f}% {v['gbs']:.0f}GB/s" for k, v in kern.items()))
    # avg_launch_us: the event-pair average over the profiled steps of the workload (it
    # includes the pairs' own overhead, so it errs long).  back_to_back_us: the same kernel
    # launched 200 times between two events, its operands warm in the caches (errs short).
    avg_us = kern[dom]["avg_us"]
    b2b = None
    try:
        b2b = round(1e3 * x64.time_kernel(dom, reps=200), 3)
    except Exception as e:  # a grain class is not re-launchable on its own
        log("back-to-back timing unavailable for", dom, e)
    ach = kb[dom] / (avg_us * 1e-6) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(dom, "cfg1_fp64"),
            "bytes_per_launch": kb[dom], "avg_launch_us": round(avg_us, 3), "back_to_back_us": b2b,
            "copy_ceiling_gbs": None}
    try:
        nb = 1 << 30
        ms = x64.copy_probe(nb, reps=10)
        roof["copy_ceiling_gbs"] = round(2 * nb / (ms * 1e-3) / 1e9, 1)
    except Exception as e:  # pragma: no cover
        log("copy probe failed:", e)

    out = {"metric": METRIC, "value": v64, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": el64 * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded Voronoi media, Philox thermal field)",
           "config": {"workload": "cfg1: 256x256x4 cells of 2 nm, two coupled 4 nm layers, T=300 K, "
                                  "Heun dt=50 fs, periodic FFT demag, FP64 plain form",
                      "grid": [med.nx, med.ny, med.n_lay], "cells": med.cells, "grains": med.G,
                      "launch_shapes": shapes64,
                      "grain_updates_per_s": med.G * a.steps / el64},
           "roofline": roof,
           "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in kern.items()}}

    if not a.no_fp32:
        x32, _ = build_ctx(med, abi.MRMG_FP32, tensor=N)
        shapes32 = autotune(x32)
        log(f"fp32 launch shapes: {shapes32}")
        x32.step(a.warmup, philox_seed=SEED)
        el32, _ = timed(x32, a.steps, t0)
        v32 = med.cells * a.steps / el32
        # accuracy of the precision form: same start, same noise, same number of steps
        x32.close()
        log(f"fp32: {el32 * 1e3 / a.steps:.4f} ms/step, {v32:.4e} cell-updates/s, x{v32 / v64:.2f}")
        acc = accuracy(med, N, a.accuracy_steps, log)
        out["fp32_form"] = {"value": v32, "unit": "cell-updates/s", "ms_per_step": el32 * 1e3 / a.steps,
                            "speedup_vs_fp64": v32 / v64, "max_dm_vs_fp64_T0": acc["T0_fp32_vs_fp64"]["max_dm"],
                            "accuracy": acc, "launch_shapes": shapes32}
    if not a.no_track:
        out["track_grid"] = track_form(log)
    if not a.no_serve:
        out["serving"] = serving(log)

    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(med, N)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
