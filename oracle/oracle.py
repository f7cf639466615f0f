"""Python handle on the oracle (test infrastructure only).

Loads oracle/liboracle.so (the C restatement of the reference path, mrm_oracle.c) and,
where it was built, oracle/_ref/libmrmref.so (the reference's own tensor_element, random,
io and transform modules compiled in place with flang plus ref_shim.f90).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from mrm_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmrmref.so")

_dp = C.POINTER(C.c_double)
_cp = C.c_void_p


def build():
    """Compile the oracle (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib(path=None):
    """The oracle library (cached), or the one built at path."""
    global _lib
    if _lib is None or path is not None:
        if path is None and not os.path.exists(LIB):
            build()
        L = C.CDLL(path or LIB)
        sig = {
            "ora_mt_size": ([], C.c_size_t),
            "ora_mt_seed": ([_cp, C.c_int], None),
            "ora_mt_fill_uniform": ([_cp, C.c_int, _dp], None),
            "ora_mt_fill_normal3d": ([_cp, C.c_int, _dp], None),
            "ora_volume_averaged_n": ([C.c_double] * 9 + [_dp], C.c_double),
            "ora_plane_to_plane": ([C.c_int, C.c_int] + [C.c_double] * 5 + [C.c_int, C.c_int, _dp], None),
            "ora_calculate_tensor": ([C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp,
                                      C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp], None),
            "ora_fft": ([_cp, C.c_int, C.c_int, C.c_int], None),
            "ora_rfft2": ([_dp, _cp, C.c_int, C.c_int], None),
            "ora_irfft2": ([_cp, _dp, C.c_int, C.c_int], None),
            "ora_create": ([C.POINTER(abi.Geom)], _cp),
            "ora_destroy": ([_cp], None),
            "ora_set_tensor": ([_cp, _dp], C.c_int),
            "ora_set_layer": ([_cp, C.c_int, C.POINTER(abi.Layer)], C.c_int),
            "ora_set_ufield": ([_cp, C.c_int, C.POINTER(abi.UField)], C.c_int),
            "ora_set_source": ([_cp, C.c_int, C.c_int, C.POINTER(abi.Source)], C.c_int),
            "ora_source_cells": ([_cp, C.c_int, C.c_int, C.c_double, _dp], C.c_int),
            "ora_set_max_dm": ([_cp, C.c_int, C.c_double], None),
            "ora_set_state": ([_cp, C.c_int, _dp], C.c_int),
            "ora_get_state": ([_cp, C.c_int, _dp, _dp], C.c_int),
            "ora_set_mag": ([_cp, C.c_int, _dp], C.c_int),
            "ora_get_mag": ([_cp, C.c_int, _dp], C.c_int),
            "ora_step": ([_cp, C.POINTER(abi.StepParams), _dp, C.POINTER(abi.Diag)], C.c_int),
            "ora_demag_field": ([_cp, _dp, _dp], C.c_int),
            "ora_num_threads": ([], C.c_int),
            "ora_lv": ([C.c_double], C.c_double),
            "ora_dlv": ([C.c_double], C.c_double),
            "ora_funct_value": ([C.c_int] + [C.c_double] * 6, C.c_double),
            "ora_get_scales": ([_cp, C.c_int, _dp, _dp, _dp], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        if path is not None:
            return L
        _lib = L
    return _lib


def planted(define, directory):
    """The oracle compiled with -D<define> into directory: a copy with a planted mistake
    (the ORA_PLANT_* switches of mrm_oracle.c), which a pin must tell from the real one."""
    out = os.path.join(directory, f"liboracle_{define.lower()}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-fno-fast-math", "-std=c11", "-D_GNU_SOURCE", "-D" + define,
                    "-shared", "-o", out, os.path.join(HERE, "mrm_oracle.c"), "-lm"], check=True)
    return lib(out)


def ref_available():
    return os.path.exists(REF_LIB)


_ref = None


def ref():
    """The flang-built reference modules (only where /root/reference was present)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        R.ref_volume_averaged_n.argtypes = [C.c_double] * 9 + [_dp]
        R.ref_volume_averaged_n.restype = C.c_double
        R.ref_plane_to_plane.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [C.c_int, C.c_int, _dp]
        R.ref_random_create.argtypes = [C.c_int, C.c_int]
        R.ref_random_uniform.argtypes = [C.c_int, C.c_int, _dp]
        R.ref_random_normal3d.argtypes = [C.c_int, C.c_int, _dp]
        _ref = R
    return _ref


# ---------------------------------------------------------------- MT19937 stream
class MT:
    """random_m generator (random.f90:166-249), one per layer in the reference."""

    def __init__(self, seed):
        L = lib()
        self._buf = C.create_string_buffer(L.ora_mt_size())
        L.ora_mt_seed(self._buf, int(seed))

    def uniform(self, n):
        out = np.empty(n)
        lib().ora_mt_fill_uniform(self._buf, n, abi.dptr(out))
        return out

    def normal3d(self, n):
        out = np.empty((n, 3))
        lib().ora_mt_fill_normal3d(self._buf, n, abi.dptr(out))
        return out


def mt_noise(seeds, ngs, thermal, n_steps):
    """Standard normals (n_steps, G, 3) in get_thermal's order: per step, per layer with
    thermal_p, per grain, one random_normal3d from that layer's stream."""
    gens = [MT(s) for s in seeds]
    G = sum(ngs)
    out = np.zeros((n_steps, G, 3))
    for it in range(n_steps):
        off = 0
        for k, ng in enumerate(ngs):
            if thermal[k]:
                out[it, off:off + ng] = gens[k].normal3d(ng)
            off += ng
    return out


# ---------------------------------------------------------------- tensor
def volume_averaged_n(rx, ry, rz, xs, ys, zs, xo, yo, zo):
    n = np.zeros(9)
    vol = lib().ora_volume_averaged_n(rx, ry, rz, xs, ys, zs, xo, yo, zo, abi.dptr(n))
    return n.reshape(3, 3).T.copy(), vol   # [a][b] = n(a,b)


def plane_to_plane(nx, ny, dx, dy, dz, dzo, rz, imx, imy):
    out = np.zeros(9 * nx * ny)
    lib().ora_plane_to_plane(nx, ny, dx, dy, dz, dzo, rz, imx, imy, abi.dptr(out))
    return out  # Fortran (nx,ny,3,3)


def calculate_tensor(nx, ny, n_lay, dx, dy, dz, il_space=None, num_images=2, sul_mu=1.0, head_mu=1.0,
                     fly_height=8e-7, space_to_sul=40e-7):
    dz = np.ascontiguousarray(np.broadcast_to(np.asarray(dz, float), (n_lay,)).copy())
    il = np.zeros(n_lay) if il_space is None else np.ascontiguousarray(il_space, float)
    out = np.zeros(nx * ny * 9 * n_lay * n_lay)
    lib().ora_calculate_tensor(nx, ny, n_lay, dx, dy, abi.dptr(dz), abi.dptr(il), num_images, num_images,
                               sul_mu, head_mu, fly_height, space_to_sul, abi.dptr(out))
    return out  # Fortran (nx,ny,9,n_lay,n_lay)


# ---------------------------------------------------------------- FFT
def rfft2(r):
    """r: numpy (ny,nx) (= Fortran (nx,ny)); returns (ny, nx//2+1) complex."""
    ny, nx = r.shape
    r = np.ascontiguousarray(r, dtype=np.float64)
    c = np.zeros((ny, nx // 2 + 1), dtype=np.complex128)
    lib().ora_rfft2(abi.dptr(r), c.ctypes.data_as(C.c_void_p), nx, ny)
    return c


def irfft2(c, nx):
    ny = c.shape[0]
    c = np.ascontiguousarray(c, dtype=np.complex128).copy()
    r = np.zeros((ny, nx))
    lib().ora_irfft2(c.ctypes.data_as(C.c_void_p), abi.dptr(r), nx, ny)
    return r


# ---------------------------------------------------------------- simulation
def funct(func, t, gamma, phase=0.0, param=0.0, duty=0.0, moddepth=0.0):
    """funct (applied.f90:328-396) as restated in the oracle."""
    return lib().ora_funct_value(int(func), t, gamma, phase, param, duty, moddepth)


class Sim:
    """The restated llg/field/demag path on the CPU, driven like the C ABI."""

    def __init__(self, geom, library=None):
        self._g = geom
        self.L = library if library is not None else lib()
        self.h = self.L.ora_create(C.byref(geom))
        self.nx, self.ny, self.nl = geom.nx, geom.ny, geom.n_lay
        self.ngs = [0] * self.nl
        self._keep = []

    def close(self):
        if self.h:
            self.L.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_tensor(self, n_real):
        n_real = np.ascontiguousarray(n_real, dtype=np.float64)
        self.L.ora_set_tensor(self.h, abi.dptr(n_real))

    def set_layer(self, k, la):
        s = la.struct()
        self._keep.append(la)
        assert self.L.ora_set_layer(self.h, k, C.byref(s)) == 0
        self.ngs[k] = la.ng

    def set_ufield(self, k, uf):
        assert self.L.ora_set_ufield(self.h, k, C.byref(uf)) == 0

    def set_source(self, k, slot, src):
        self._keep.append(src)
        st = src.struct() if src is not None else abi.Source(kind=abi.MRMG_SRC_NONE)
        assert self.L.ora_set_source(self.h, k, slot, C.byref(st)) == 0

    def source_cells(self, k, slot, t):
        nv = 3 if slot == 0 else 1
        out = np.zeros((nv, self.ny, self.nx))
        assert self.L.ora_source_cells(self.h, k, slot, t, abi.dptr(out)) == 0
        return out

    def set_max_dm(self, k, deg):
        self.L.ora_set_max_dm(self.h, k, deg)

    def set_state(self, k, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        self.L.ora_set_state(self.h, k, abi.dptr(m))

    def get_scales(self, k):
        """(ms_scale, alpha_scale, hk_scale) of layer k as find_temp_scaling last set them."""
        a, b, c = (np.zeros(self.ngs[k]) for _ in range(3))
        assert self.L.ora_get_scales(self.h, k, abi.dptr(a), abi.dptr(b), abi.dptr(c)) == 0
        return a, b, c

    def get_state(self, k):
        m = np.zeros((self.ngs[k], 3))
        h = np.zeros((self.ngs[k], 3))
        self.L.ora_get_state(self.h, k, abi.dptr(m), abi.dptr(h))
        return m, h

    def set_mag(self, k, mag):
        mag = np.ascontiguousarray(mag, dtype=np.float64)
        self.L.ora_set_mag(self.h, k, abi.dptr(mag))

    def get_mag(self, k):
        mag = np.zeros(self.ngs[k])
        self.L.ora_get_mag(self.h, k, abi.dptr(mag))
        return mag

    def step(self, n_steps, dt, t0=0.0, type=abi.MRMG_TYPE_HEUN, noise=None, dt_frac=0.0, dt_max=1e22,
             max_dm=180.0, llb=False):
        p = abi.StepParams(type=type, n_steps=n_steps, t0=t0, dt=dt,
                           noise=abi.MRMG_NOISE_INJECT if noise is not None else abi.MRMG_NOISE_NONE,
                           dt_frac=dt_frac, dt_max=dt_max, max_dm=max_dm, llb=int(bool(llb)))
        d = abi.Diag()
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        self.L.ora_step(self.h, C.byref(p), abi.dptr(nz), C.byref(d))
        return d

    def demag_field(self):
        G = sum(self.ngs)
        hc = np.zeros((3, self.nl, self.ny, self.nx))
        hg = np.zeros((G, 3))
        self.L.ora_demag_field(self.h, abi.dptr(hc), abi.dptr(hg))
        return hc, hg
