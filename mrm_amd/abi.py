"""ctypes mirror of include/mrmg.h and the loader of the HIP library libmrmg.so.

The structures here are the C ABI of the library (the same layouts a Fortran caller
binds through iso_c_binding).  A CPU test compiles the header with the host compiler and
checks that every size and field offset below matches it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrmg.so")

MRMG_OK = 0
MRMG_FP64, MRMG_FP32 = 0, 1
MRMG_NOISE_NONE, MRMG_NOISE_INJECT, MRMG_NOISE_PHILOX = 0, 1, 2
MRMG_TYPE_HEUN = -3
MRMG_TYPE_BORIS = 2
MRMG_TYPE_BORIS_MID = -2
MRMG_TYPE_QUAT = 1
MRMG_TYPE_QUAT_MID = -1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Geom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("n_lay", C.c_int),
                ("precision", C.c_int), ("demag", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double),
                ("dz", _dp), ("il_space", _dp)]


class Layer(C.Structure):
    _fields_ = [("ng", C.c_int),
                ("grid2grain", _ip), ("grain2gridindex", _ip), ("grain2grid", _ip),
                ("grain_vol", _dp), ("ms", _dp), ("hk", _dp),
                ("uni_axis", _dp), ("cubic_axis", _dp),
                ("hxgrainindex", _ip), ("hxgrain", _ip), ("hx", _dp),
                ("uhxgrainindex", _ip), ("uhxgrain", _ip), ("uhx", _dp),
                ("lhxgrainindex", _ip), ("lhxgrain", _ip), ("lhx", _dp),
                ("ms_scale", _dp), ("alpha_scale", _dp), ("hk_scale", _dp),
                ("temp", _dp), ("h_app", _dp),
                ("alpha", C.c_double), ("gamma", C.c_double),
                ("thm_alpha", C.c_double), ("floor_ms", C.c_double),
                ("thermal", C.c_int), ("use_mag", C.c_int),
                ("tc", _dp), ("lambda_ms", _dp), ("spin", C.c_double), ("tc_nom", C.c_double),
                ("msexp", C.c_double), ("hkexp", C.c_double),
                ("n_ms_tab", C.c_int), ("n_alpha_tab", C.c_int), ("n_hk_tab", C.c_int), ("pad3_", C.c_int),
                ("ms_tab_t", _dp), ("ms_tab_s", _dp), ("alpha_tab_t", _dp), ("alpha_tab_s", _dp),
                ("hk_tab_t", _dp), ("hk_tab_s", _dp)]


MRMG_FN_RAMP, MRMG_FN_SIN, MRMG_FN_EXP, MRMG_FN_COS, MRMG_FN_NEW, MRMG_FN_STE = 0, 1, 2, 3, 4, 5
_TFUN = [("func", C.c_int), ("pad_", C.c_int), ("amp", C.c_double), ("gamma", C.c_double), ("phase", C.c_double),
         ("param", C.c_double), ("duty", C.c_double), ("moddepth", C.c_double), ("dc", C.c_double)]
TFUN_KEYS = ("func", "amp", "gamma", "phase", "param", "duty", "moddepth", "dc")


class UField(C.Structure):
    _fields_ = [("dir", C.c_double * 3), ("t_start", C.c_double), ("t_end", C.c_double),
                ("s_start", C.c_double), ("s_end", C.c_double), ("scale", C.c_double)] + _TFUN


MRMG_SRC_NONE, MRMG_SRC_FIELD_TABLE, MRMG_SRC_TEMP_TABLE, MRMG_SRC_TEMP_SG = 0, 1, 2, 3


class Source(C.Structure):
    _fields_ = [("kind", C.c_int), ("nxp", C.c_int), ("nyp", C.c_int), ("expn", C.c_int),
                ("xpos", _dp), ("ypos", _dp), ("vals", _dp),
                ("vel_x", C.c_double), ("vel_y", C.c_double),
                ("offset_x", C.c_double), ("offset_y", C.c_double), ("ref_x", C.c_double), ("ref_y", C.c_double),
                ("t_origin", C.c_double), ("isigx", C.c_double), ("isigy", C.c_double), ("dtemp", C.c_double),
                ("t_start", C.c_double), ("t_end", C.c_double), ("s_start", C.c_double), ("s_end", C.c_double),
                ("scale", C.c_double)] + _TFUN


class SourceArrays:
    """A moving source with its tables kept alive (mrmg_source).  Tables: xpos (nxp),
    ypos (nyp), vals (nyp, nxp, 3) for a field or (nyp, nxp) for a temperature (x fastest,
    the Fortran (3,nxp,nyp) / (nxp,nyp) layouts)."""

    def __init__(self, kind, xpos=None, ypos=None, vals=None, vel=(0.0, 0.0), offset=(0.0, 0.0), ref=(0.0, 0.0),
                 t_origin=0.0, sigma=(0.0, 0.0), expn=2, dtemp=0.0, t_start=0.0, t_end=1.0, s_start=1.0, s_end=1.0,
                 scale=1.0, **tfun):
        self.kind = int(kind)
        self.xpos = None if xpos is None else np.ascontiguousarray(xpos, dtype=np.float64)
        self.ypos = None if ypos is None else np.ascontiguousarray(ypos, dtype=np.float64)
        self.vals = None if vals is None else np.ascontiguousarray(vals, dtype=np.float64)
        self.kw = dict(vel_x=vel[0], vel_y=vel[1], offset_x=offset[0], offset_y=offset[1], ref_x=ref[0], ref_y=ref[1],
                       t_origin=t_origin, isigx=(1.0 / sigma[0]) if sigma[0] else 0.0,
                       isigy=(1.0 / sigma[1]) if sigma[1] else 0.0, dtemp=dtemp, t_start=t_start, t_end=t_end,
                       s_start=s_start, s_end=s_end, scale=scale)
        for key, v in tfun.items():
            if key not in TFUN_KEYS:
                raise TypeError(f"unknown time-function argument {key}")
            self.kw[key] = int(v) if key == "func" else float(v)
        self.expn = int(expn)

    def struct(self):
        s = Source(kind=self.kind, expn=self.expn, **self.kw)
        if self.xpos is not None:
            s.nxp, s.nyp = len(self.xpos), len(self.ypos)
            s.xpos, s.ypos, s.vals = dptr(self.xpos), dptr(self.ypos), dptr(self.vals)
        return s


class StepParams(C.Structure):
    _fields_ = [("type", C.c_int), ("n_steps", C.c_int), ("t0", C.c_double), ("dt", C.c_double),
                ("noise", C.c_int), ("pad_", C.c_int), ("seed", C.c_uint64), ("step0", C.c_uint64),
                ("dt_frac", C.c_double), ("dt_max", C.c_double), ("max_dm", C.c_double),
                ("llb", C.c_int), ("pad2_", C.c_int)]


class Diag(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("steps", C.c_int), ("n_lay", C.c_int),
                ("errmx", C.c_double), ("errmx1", C.c_double), ("mmag", C.c_double),
                ("mave", (C.c_double * 3) * 8), ("rejected", C.c_int), ("pad_", C.c_int),
                ("max_dmdt", C.c_double), ("mvar", (C.c_double * 3) * 8), ("have", (C.c_double * 3) * 8),
                ("mave_all", C.c_double * 3), ("mvar_all", C.c_double * 3), ("have_all", C.c_double * 3)]


class KStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int), ("total_ms", C.c_double)]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array, or NULL for None."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "need contiguous int32"
    return a.ctypes.data_as(_ip)


class LayerArrays:
    """Holds the numpy arrays behind one Layer struct (keeps them alive).

    Arrays are given row-per-grain in C order, which is the memory of the reference's
    Fortran arrays: (ng,3) is Fortran (3,ng), (ng,2) is (2,ng), (ng,3,3) indexed [n][c][r]
    is cubic_axis(r,c,n), (ny,nx) is grid2grain(nx,ny), (ncell,2) is grain2grid(2,ncell).
    """

    FIELDS_D = ("grain_vol", "ms", "hk", "uni_axis", "cubic_axis", "hx", "uhx", "lhx",
                "ms_scale", "alpha_scale", "hk_scale", "temp", "h_app", "tc", "lambda_ms")
    FIELDS_I = ("grid2grain", "grain2gridindex", "grain2grid", "hxgrainindex", "hxgrain",
                "uhxgrainindex", "uhxgrain", "lhxgrainindex", "lhxgrain")

    def __init__(self, **kw):
        self.ng = int(kw["ng"])
        for f in self.FIELDS_D:
            v = kw.get(f)
            setattr(self, f, None if v is None else np.ascontiguousarray(v, dtype=np.float64).ravel())
        for f in self.FIELDS_I:
            v = kw.get(f)
            setattr(self, f, None if v is None else np.ascontiguousarray(v, dtype=np.int32).ravel())
        self.alpha = float(kw.get("alpha", 1.0))
        self.gamma = float(kw.get("gamma", 1.76e7))
        self.thm_alpha = float(kw.get("thm_alpha", 1.0))
        self.floor_ms = float(kw.get("floor_ms", 0.01))
        self.thermal = int(bool(kw.get("thermal", False)))
        self.use_mag = int(bool(kw.get("use_mag", False)))
        self.spin = float(kw.get("spin", 0.5))
        self.tc_nom = float(kw.get("tc_nom", 1e22))
        self.msexp = float(kw.get("msexp", -1.0))
        self.hkexp = float(kw.get("hkexp", -1.0))
        # READ_SCALE tables: {"ms"|"alpha"|"hk": (T/Tc nodes, scalings)}
        self.tables = {}
        for name, tab in (kw.get("tables") or {}).items():
            if name not in ("ms", "alpha", "hk"):
                raise ValueError(f"unknown scaling table {name}")
            t = np.ascontiguousarray(tab[0], dtype=np.float64)
            v = np.ascontiguousarray(tab[1], dtype=np.float64)
            self.tables[name] = (t, v)

    def struct(self):
        s = Layer()
        s.ng = self.ng
        for f in self.FIELDS_D:
            setattr(s, f, dptr(getattr(self, f)))
        for f in self.FIELDS_I:
            setattr(s, f, iptr(getattr(self, f)))
        s.alpha, s.gamma, s.thm_alpha, s.floor_ms = self.alpha, self.gamma, self.thm_alpha, self.floor_ms
        s.thermal = self.thermal
        s.use_mag, s.spin, s.tc_nom, s.msexp, s.hkexp = self.use_mag, self.spin, self.tc_nom, self.msexp, self.hkexp
        for name, (t, v) in self.tables.items():
            setattr(s, f"n_{name}_tab", len(t))
            setattr(s, f"{name}_tab_t", dptr(t))
            setattr(s, f"{name}_tab_s", dptr(v))
        return s


class Sensor(C.Structure):
    """mrmg_sensor: Gaussian reciprocity sensitivity of a read head."""
    _fields_ = [("sigma_x", C.c_double), ("sigma_y", C.c_double), ("dir", C.c_double * 3), ("w", _dp)]


def make_geom(nx, ny, n_lay, dx, dy, dz, il_space=None, precision=MRMG_FP64, demag=True):
    dz = np.ascontiguousarray(np.broadcast_to(np.asarray(dz, dtype=np.float64), (n_lay,)).copy())
    il = np.zeros(n_lay) if il_space is None else np.ascontiguousarray(il_space, dtype=np.float64)
    g = Geom(nx=nx, ny=ny, n_lay=n_lay, precision=precision, demag=int(bool(demag)),
             dx=dx, dy=dy, dz=dptr(dz), il_space=dptr(il))
    g._keep = (dz, il)
    return g


def make_ufield(direction, t_start=-1e22, t_end=1e22, s_start=1.0, s_end=1.0, scale=1.0, **tfun):
    """mrmg_ufield; tfun: func (MRMG_FN_*), amp, gamma, phase, param, duty, moddepth, dc."""
    u = UField()
    for i in range(3):
        u.dir[i] = float(direction[i])
    u.t_start, u.t_end, u.s_start, u.s_end, u.scale = t_start, t_end, s_start, s_end, scale
    for key, v in tfun.items():
        if key not in TFUN_KEYS:
            raise TypeError(f"unknown time-function argument {key}")
        setattr(u, key, int(v) if key == "func" else float(v))
    return u


class MrmgError(RuntimeError):
    pass


_lib = None


def load(path=None):
    """Load libmrmg.so.  Raises if it is missing: there is no fallback path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise MrmgError(f"{p} not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(p)
    lib.mrmg_last_error.restype = C.c_char_p
    for name, args in [
        ("mrmg_device_count", [_ip]),
        ("mrmg_create", [C.POINTER(Geom), C.POINTER(C.c_void_p)]),
        ("mrmg_create_open", [C.POINTER(Geom), C.POINTER(C.c_void_p)]),
        ("mrmg_destroy", [C.c_void_p]),
        ("mrmg_set_tensor", [C.c_void_p, _dp]),
        ("mrmg_build_tensor", [C.c_void_p, C.c_int, _dp]),
        ("mrmg_build_tensor_images", [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
        ("mrmg_set_layer", [C.c_void_p, C.c_int, C.POINTER(Layer)]),
        ("mrmg_set_ufield", [C.c_void_p, C.c_int, C.POINTER(UField)]),
        ("mrmg_set_source", [C.c_void_p, C.c_int, C.c_int, C.POINTER(Source)]),
        ("mrmg_set_state", [C.c_void_p, C.c_int, _dp]),
        ("mrmg_get_state", [C.c_void_p, C.c_int, _dp, _dp]),
        ("mrmg_set_mag", [C.c_void_p, C.c_int, _dp]),
        ("mrmg_get_mag", [C.c_void_p, C.c_int, _dp]),
        ("mrmg_step", [C.c_void_p, C.POINTER(StepParams), _dp, C.POINTER(Diag)]),
        ("mrmg_step_batch", [C.POINTER(C.c_void_p), C.POINTER(StepParams), C.c_int, C.c_int, _dp]),
        ("mrmg_copy_tensor", [C.c_void_p, C.c_void_p]),
        ("mrmg_philox_normals", [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, _dp]),
        ("mrmg_demag_field", [C.c_void_p, _dp, _dp]),
        ("mrmg_footprint", [C.c_void_p, C.POINTER(C.c_size_t)]),
        ("mrmg_profile", [C.c_void_p, C.c_int]),
        ("mrmg_kernel_stats", [C.c_void_p, C.POINTER(KStat), C.c_int, _ip]),
        ("mrmg_copy_probe", [C.c_void_p, C.c_size_t, C.c_int, _dp]),
        ("mrmg_memcpy_probe", [C.c_void_p, C.c_size_t, C.c_int, _dp]),
        ("mrmg_time_kernel", [C.c_void_p, C.c_char_p, C.c_int, _dp]),
        ("mrmg_tune", [C.c_void_p, C.c_char_p, C.c_int]),
        ("mrmg_get_tune", [C.c_void_p, C.c_char_p, _ip]),
        ("mrmg_readback", [C.c_void_p, C.POINTER(Sensor), C.c_int, _dp, _dp, _dp]),
        ("mrmg_readback_noalloc", [C.c_void_p, C.POINTER(Sensor), C.c_int, _dp, _dp, _dp]),
        ("mrmg_switch_arm", [C.c_void_p, C.c_int, _dp, C.c_double]),
        ("mrmg_switch_read", [C.c_void_p, C.c_int, C.POINTER(C.c_int64), _dp, _dp]),
    ]:
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def check(lib, rc):
    if rc != MRMG_OK:
        msg = lib.mrmg_last_error()
        raise MrmgError(f"mrmg error {rc}: {msg.decode() if msg else ''}")
