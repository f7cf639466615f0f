"""The reference's binary files, read and written on the host.

All of them are Fortran unformatted sequential files as gfortran lays them out: each
record is its bytes between two 4-byte little-endian lengths; a record longer than
2**31 - 9 bytes is split into subrecords whose markers carry a negative length on every
subrecord but the last.  Default INTEGER and LOGICAL are 4 bytes (.true. written as 1,
any nonzero read as true), REAL(8) is 8 bytes, arrays are in Fortran order.

* Grain microstructure (READ_GRAIN / WRITE_GRAIN, grain.f90:40-70): per layer the three
  records of file_grain (data.f90:306-312) and the records of file_field
  (field.f90:983-1001).  The arrays map onto mrm_amd.abi.LayerArrays unchanged
  (grid2grain (ny,nx) is Fortran (nx,ny), grain2grid (ncell,2) is (2,ncell), hk (ng,2)
  is (2,ng), uni_axis (ng,3) is (3,ng), cubic_axis (ng,3,3) is (3,3,ng)); the grain volume
  is grain_area * dz (data.f90:205).
* The magnetostatic tensor cache mrm_N.dat (tensor.f90:52-80): a header record with the
  geometry and image options, then n2d (nx,ny,9,n_lay,n_lay).
* The binary magnetisation of WRITE_MAG (mag.f90:285-290): per layer (nx, ny, n, ms_nom)
  and Ms * m as (3,n); its formatted (tecplot) form and the old REAL(4) per-cell form
  (mag.f90:254-283).
* The checkpoint trio (checkpoint.f90:14-149): mrm_grain.chk (the grain microstructure),
  mrm_0.chk / mrm_1.chk written in turn ((time, iline), the magnetisation, the random
  generators of save_random, random.f90:38-64), and the formatted mrm.chk naming the
  slot last written.
"""
import os
import struct

import numpy as np

from mrm_amd import abi

_MAXSUB = 2 ** 31 - 9     # gfortran's largest subrecord


class RecordWriter:
    """Writes Fortran sequential unformatted records to a binary file object."""

    def __init__(self, f):
        self.f = f

    def write(self, *items):
        """One record of the items in order: ints (INTEGER), floats (REAL(8)), bools
        (LOGICAL), or numpy arrays written in their dtype (flattened in C order, which is
        the Fortran order of the reference's arrays as this module passes them)."""
        parts = []
        for it in items:
            if isinstance(it, (bool, np.bool_)):
                parts.append(struct.pack("<i", 1 if it else 0))
            elif isinstance(it, (int, np.integer)):
                parts.append(struct.pack("<i", int(it)))
            elif isinstance(it, (float, np.floating)):
                parts.append(struct.pack("<d", float(it)))
            else:
                a = np.asarray(it)
                if a.dtype.kind == "f":
                    a = a.astype("<f8", copy=False)
                elif a.dtype.kind in "iub":
                    a = a.astype("<i4", copy=False)
                else:
                    raise TypeError(f"cannot write {a.dtype}")
                parts.append(np.ascontiguousarray(a).tobytes())
        self.write_bytes(b"".join(parts))

    def write_bytes(self, data):
        n = len(data)
        if n <= _MAXSUB:
            m = struct.pack("<i", n)
            self.f.write(m + data + m)
            return
        pos = 0
        while pos < n:
            k = min(_MAXSUB, n - pos)
            last = pos + k == n
            m = struct.pack("<i", k if last else -k)
            self.f.write(m + data[pos:pos + k] + m)
            pos += k


class RecordReader:
    """Reads Fortran sequential unformatted records from a binary file object."""

    def __init__(self, f):
        self.f = f

    def read_bytes(self):
        out = []
        while True:
            h = self.f.read(4)
            if len(h) < 4:
                raise EOFError("end of file inside a record marker")
            n = struct.unpack("<i", h)[0]
            data = self.f.read(abs(n))
            t = self.f.read(4)
            if len(data) < abs(n) or len(t) < 4 or struct.unpack("<i", t)[0] != n:
                raise ValueError("damaged record (length markers do not match)")
            out.append(data)
            if n >= 0:
                return b"".join(out)

    def read(self, *spec):
        """One record split by spec: ('i', count), ('d', count) or ('l', count) per item;
        count 1 gives a scalar.  The record must hold exactly these bytes."""
        data = self.read_bytes()
        pos, out = 0, []
        for kind, count in spec:
            if kind == "d":
                a = np.frombuffer(data, "<f8", count, pos)
                pos += 8 * count
            else:
                a = np.frombuffer(data, "<i4", count, pos)
                pos += 4 * count
                if kind == "l":
                    a = a != 0
            out.append(a[0] if count == 1 else a.copy())
        if pos != len(data):
            raise ValueError(f"record holds {len(data)} bytes, expected {pos}")
        return out


# ------------------------------------------------------------------ grain microstructure
def write_grain(path, layers):
    """WRITE_GRAIN: layers is a list of dicts with nx, ny, il (1-based layer number), dx,
    dy, grid2grain (ny,nx), grain2grid (ncell,2), grain2gridindex (ng+1), grain_area (ng),
    ms, tc, tau (ng), hk (ng,2) and optionally uni_axis (ng,3), cubic_axis (ng,3,3) and the
    exchange CSRs hx/lhx/uhx as (index (ng+1), neighbour (n2), value (n2))."""
    with open(path, "wb") as f:
        w = RecordWriter(f)
        for L in layers:
            ng = len(L["grain_area"])
            g2g = np.asarray(L["grain2grid"], np.int32).reshape(-1, 2)
            # file_grain (data.f90:306-312)
            w.write(int(L["nx"]), int(L["ny"]), int(L["il"]), float(L["dx"]), float(L["dy"]), ng, len(g2g))
            w.write(np.asarray(L["grid2grain"], np.int32), g2g, np.asarray(L["grain2gridindex"], np.int32),
                    np.asarray(L["grain_area"], np.float64))
            w.write(np.asarray(L["ms"], np.float64), np.asarray(L["tc"], np.float64),
                    np.asarray(L.get("tau", np.zeros(ng)), np.float64))
            # file_field (field.f90:983-1001): flags, hk; axes; then hx, lhx, uhx
            uni, cub = L.get("uni_axis"), L.get("cubic_axis")
            csr = [L.get(k) for k in ("hx", "lhx", "uhx")]
            w.write(uni is not None, cub is not None, csr[0] is not None, csr[1] is not None, csr[2] is not None,
                    np.asarray(L["hk"], np.float64).reshape(ng, 2))
            if uni is not None:
                w.write(np.asarray(uni, np.float64).reshape(ng, 3))
            if cub is not None:
                w.write(np.asarray(cub, np.float64).reshape(ng, 3, 3))
            for c in csr:
                if c is None:
                    continue
                idx, nb, val = (np.asarray(c[0], np.int32), np.asarray(c[1], np.int32), np.asarray(c[2], np.float64))
                w.write(idx, len(nb))
                w.write(nb, val)


def read_grain(path, n_lay):
    """READ_GRAIN (new format) of n_lay layers; returns a list of dicts as write_grain
    takes them."""
    out = []
    with open(path, "rb") as f:
        r = RecordReader(f)
        for _ in range(n_lay):
            hd = r.read_bytes()
            nx, ny, il = struct.unpack_from("<3i", hd, 0)
            dx, dy = struct.unpack_from("<2d", hd, 12)
            ng, n2 = struct.unpack_from("<2i", hd, 28)
            if len(hd) != 36:
                raise ValueError("grain header record of unexpected length")
            g2g, cells, idx, area = r.read(("i", nx * ny), ("i", 2 * n2), ("i", ng + 1), ("d", ng))
            ms, tc, tau = r.read(("d", ng), ("d", ng), ("d", ng))
            u_p, c_p, x_p, lx_p, ux_p, hk = r.read(("l", 1), ("l", 1), ("l", 1), ("l", 1), ("l", 1), ("d", 2 * ng))
            L = dict(nx=nx, ny=ny, il=il, dx=dx, dy=dy, grid2grain=np.asarray(g2g).reshape(ny, nx),
                     grain2grid=np.asarray(cells).reshape(n2, 2), grain2gridindex=np.atleast_1d(idx),
                     grain_area=np.atleast_1d(area), ms=np.atleast_1d(ms), tc=np.atleast_1d(tc),
                     tau=np.atleast_1d(tau), hk=np.asarray(hk).reshape(ng, 2), uni_axis=None, cubic_axis=None,
                     hx=None, lhx=None, uhx=None)
            if u_p:
                L["uni_axis"] = r.read(("d", 3 * ng))[0].reshape(ng, 3)
            if c_p:
                L["cubic_axis"] = r.read(("d", 9 * ng))[0].reshape(ng, 3, 3)
            for key, on in (("hx", x_p), ("lhx", lx_p), ("uhx", ux_p)):
                if not on:
                    continue
                idx_e, n_e = r.read(("i", ng + 1), ("i", 1))
                nb, val = r.read(("i", int(n_e)), ("d", int(n_e)))
                L[key] = (np.atleast_1d(idx_e), np.atleast_1d(nb), np.atleast_1d(val))
            out.append(L)
    return out


def layer_arrays(L, dz, **physics):
    """A read layer as mrm_amd.abi.LayerArrays: grain volume = grain_area * dz
    (data.f90:205); physics: alpha, gamma, thermal, temp, ... as LayerArrays takes them."""
    kw = dict(ng=len(L["grain_area"]), grid2grain=L["grid2grain"], grain2gridindex=L["grain2gridindex"],
              grain2grid=L["grain2grid"], grain_vol=np.asarray(L["grain_area"]) * dz, ms=L["ms"], hk=L["hk"],
              tc=L["tc"], uni_axis=L.get("uni_axis"), cubic_axis=L.get("cubic_axis"))
    for key, pre in (("hx", ""), ("uhx", "u"), ("lhx", "l")):
        c = L.get(key)
        if c is not None:
            kw[f"{pre}hxgrainindex"], kw[f"{pre}hxgrain"], kw[f"{pre}hx"] = c
    kw.update(physics)
    return abi.LayerArrays(**kw)


# ------------------------------------------------------------------ tensor cache
def write_tensor_cache(path, n2d, nx, ny, n_lay, dx, dy, dz, il_space, num_images=2, head_mu=1.0, sul_mu=1.0,
                       fly_height=8e-7, space_to_sul=40e-7):
    """mrm_N.dat (tensor.f90:69-70) for the flat Fortran (nx,ny,9,n_lay,n_lay) tensor."""
    head_p, sul_p = head_mu != 1.0, sul_mu != 1.0
    with open(path, "wb") as f:
        w = RecordWriter(f)
        w.write(bool(head_p), bool(sul_p), nx, ny, n_lay, num_images, num_images, float(dx), float(dy),
                float(head_mu), float(sul_mu), float(fly_height), float(space_to_sul),
                np.asarray(dz, np.float64), np.asarray(il_space, np.float64))
        w.write(np.asarray(n2d, np.float64).ravel())


def read_tensor_cache(path, nx, ny, n_lay, dx, dy, dz, il_space, num_images=2, head_mu=1.0, sul_mu=1.0,
                      fly_height=8e-7, space_to_sul=40e-7):
    """The cached tensor if every header value equals the requested geometry, as the
    reference's file_tensor demands (tensor.f90:52-58), else None."""
    with open(path, "rb") as f:
        r = RecordReader(f)
        h = r.read(("l", 1), ("l", 1), ("i", 1), ("i", 1), ("i", 1), ("i", 1), ("i", 1), ("d", 1), ("d", 1),
                   ("d", 1), ("d", 1), ("d", 1), ("d", 1), ("d", n_lay), ("d", n_lay))
        want = [head_mu != 1.0, sul_mu != 1.0, nx, ny, n_lay, num_images, num_images, dx, dy, head_mu, sul_mu,
                fly_height, space_to_sul]
        if any(bool(a != b) for a, b in zip(h[:13], want)):
            return None
        if not (np.array_equal(np.atleast_1d(h[13]), np.asarray(dz, float))
                and np.array_equal(np.atleast_1d(h[14]), np.asarray(il_space, float))):
            return None
        return r.read(("d", nx * ny * 9 * n_lay * n_lay))[0]


# ------------------------------------------------------------------ magnetisation
def _write_mag_records(w, layers):
    for nx, ny, ms_nom, m, ms in layers:
        m = np.asarray(m, np.float64).reshape(-1, 3)
        w.write(int(nx), int(ny), len(m), float(ms_nom))
        w.write(m * np.asarray(ms, np.float64)[:, None])


def _read_mag_records(r, n_lay):
    out = []
    for _ in range(n_lay):
        hd = r.read_bytes()
        if len(hd) != 20:
            raise ValueError("magnetisation header record of unexpected length")
        nx, ny, n = struct.unpack_from("<3i", hd, 0)
        ms_nom = struct.unpack_from("<d", hd, 12)[0]
        mm = r.read(("d", 3 * n))[0].reshape(n, 3)
        mag = np.sqrt((mm * mm).sum(1))
        out.append((nx, ny, ms_nom, mm / np.where(mag > 0, mag, 1.0)[:, None], mag))
    return out


def write_mag(path, layers):
    """WRITE_MAG binary form (mag.f90:285-290): layers is a list of (nx, ny, ms_nom, m
    (ng,3), ms (ng)); each writes (nx, ny, ng, ms_nom) and Ms * m as (3,ng)."""
    with open(path, "wb") as f:
        _write_mag_records(RecordWriter(f), layers)


def read_mag(path, n_lay):
    """The binary magnetisation of n_lay layers: per layer (nx, ny, ms_nom, unit m (ng,3),
    |Ms m| (ng)).  The reference's reader normalises m the same way; its Ms update writes
    ms(n) with n the grain count (mag.f90:246), i.e. only the last grain, which is not
    reproduced here: the per-grain magnitude is returned instead."""
    with open(path, "rb") as f:
        return _read_mag_records(RecordReader(f), n_lay)


def write_mag_ascii(path, layers, old_tran=0):
    """WRITE_MAG to a formatted file (mag.f90:254-270): per layer a tecplot ZONE header
    and one line per cell, x fastest: i, j, layer, then Ms*ms_scale*m of the cell's grain
    (zeros in non-magnetic cells) in 1PE12.5.  layers: list of (grid2grain (ny,nx), 1-based
    grain ids with 0 for none; m (ng,3); ms_eff (ng) = Ms * ms_scale)."""
    with open(path, "w") as f:
        for k, (g2g, m, ms_eff) in enumerate(layers, start=1):
            g2g = np.asarray(g2g)
            ny, nx = g2g.shape
            v = np.asarray(m, np.float64).reshape(-1, 3) * np.asarray(ms_eff, np.float64)[:, None]
            f.write(f'ZONE T="{float(old_tran):9.2f}", I={nx:5d}, J={ny:5d}, K={1:3d} F=POINT\n')
            for j in range(ny):
                for i in range(nx):
                    n = int(g2g[j, i])
                    c = v[n - 1] if n > 0 else (0.0, 0.0, 0.0)
                    f.write(f"{i + 1:5d}{j + 1:5d}{k:5d}" + "".join(f" {float(x):12.5E}" for x in c) + "\n")


def read_mag_ascii(path, n_lay):
    """The formatted magnetisation as the reference reads it (mag.f90:191-209): per layer
    (nx, ny, cells (ny,nx,3)), the three values of each cell line as written."""
    out = []
    with open(path) as f:
        for _ in range(n_lay):
            hd = f.readline()
            nx, ny = int(hd[22:27]), int(hd[31:36])     # "(22x,i5,4x,i5)"
            cells = np.zeros((ny, nx, 3))
            for j in range(ny):
                for i in range(nx):
                    line = f.readline()
                    cells[j, i] = [float(line[15 + 13 * c + 1:15 + 13 * (c + 1)]) for c in range(3)]
            out.append((nx, ny, cells))
    return out


def grain_directions(g2g, cells, ng):
    """Unit m per grain from a cell array, as the reference's ASCII reader sets it: every
    cell of a grain overwrites the grain's m with its normalised value, so the last cell
    in reading order (x fastest) wins.  A zero vector on a magnetic cell, or a grain number
    above ng, fails the read as the reference's does (ok_p false, mag.f90:203-205, which
    then takes the error exit); a grain with no cells keeps NaN."""
    m = np.full((ng, 3), np.nan)
    g2g = np.asarray(g2g)
    if g2g.size and int(g2g.max()) > ng:
        raise ValueError(f"grain number {int(g2g.max())} above the layer's {ng} grains")
    for j in range(g2g.shape[0]):
        for i in range(g2g.shape[1]):
            n = int(g2g[j, i])
            if n > 0:
                v = cells[j, i]
                mg = float(v @ v)
                if mg == 0.0:
                    raise ValueError(f"zero magnetisation on magnetic cell ({i + 1},{j + 1}) of grain {n}")
                m[n - 1] = v / np.sqrt(mg)
    return m


def write_mag_old(path, layers):
    """The old binary magnetisation (mag.f90:272-283): per layer one record of REAL(4)
    (3,nx,ny), the grain's unit m in every cell and zeros elsewhere.  layers: list of
    (grid2grain (ny,nx), m (ng,3))."""
    with open(path, "wb") as f:
        w = RecordWriter(f)
        for g2g, m in layers:
            g2g = np.asarray(g2g)
            m = np.asarray(m, np.float64).reshape(-1, 3)
            a = np.zeros(g2g.shape + (3,), np.float32)
            on = g2g > 0
            a[on] = m[g2g[on] - 1].astype(np.float32)
            w.write_bytes(a.astype("<f4").tobytes())


def read_mag_old(path, grid2grains, ngs=None):
    """The old binary magnetisation as the reference reads it (mag.f90:212-229): each
    magnetic cell sets its grain's m to the REAL(4) vector divided by its squared norm
    (dot_product, not its norm: a quirk of the reference, harmless for unit vectors to
    within REAL(4) rounding); the last cell of a grain wins.  ngs: the grain count of each
    layer (size(m,2)); default the largest grain number of its map.  A grain number above
    it, or a zero vector on a magnetic cell, fails the read as in the reference (ok_p
    false, then its error exit, mag.f90:219-226): ValueError.  Returns per layer m (ng,3)."""
    out = []
    with open(path, "rb") as f:
        r = RecordReader(f)
        for il, g2g in enumerate(grid2grains):
            g2g = np.asarray(g2g)
            ny, nx = g2g.shape
            data = r.read_bytes()
            if len(data) != 12 * nx * ny:
                raise ValueError("old magnetisation record of unexpected length")
            a = np.frombuffer(data, "<f4").reshape(ny, nx, 3).astype(np.float64)
            ng = int(g2g.max()) if ngs is None else int(ngs[il])
            m = np.full((ng, 3), np.nan)
            for j in range(ny):
                for i in range(nx):
                    n = int(g2g[j, i])
                    if n > ng:
                        raise ValueError(f"layer {il + 1}: grain number {n} above its {ng} grains")
                    if n > 0:
                        v = a[j, i]
                        mg = float(v @ v)
                        if not mg > 0.0:
                            raise ValueError(f"layer {il + 1}: zero magnetisation on magnetic cell ({i + 1},{j + 1})")
                        m[n - 1] = v / mg
            out.append(m)
    return out


# ------------------------------------------------------------------ random generators
MT_N = 624                 # MT19937 state words (random.f90:10)
NAME_LEN = 40              # random_s%name, character(40)
PHILOX_NAME = "philox4x32-10"


def philox_generator(seed, counter):
    """The device's Philox4x32-10 stream as a save_random entry: the 64-bit key in mt(0:1)
    and the 64-bit step counter in mt(2:3), low word first; mti = 625 marks it as no
    MT19937 state.  A reference run reading it finds no generator of that name and skips
    it, as restore_random does for any unknown name (random.f90:100-103)."""
    mt = np.zeros(MT_N, np.uint32)
    mt[0], mt[1] = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    mt[2], mt[3] = counter & 0xFFFFFFFF, (counter >> 32) & 0xFFFFFFFF
    return dict(name=PHILOX_NAME, mti=MT_N + 1, mt=mt.view(np.int32), iset=0, gset=0.0)


def philox_state(gen):
    """(seed, counter) of a philox_generator entry."""
    mt = np.asarray(gen["mt"], np.int32).view(np.uint32).astype(np.uint64)
    return int(mt[0] | (mt[1] << np.uint64(32))), int(mt[2] | (mt[3] << np.uint64(32)))


def _write_generators(w, gens):
    # save_random (random.f90:38-64): the count, then one record per generator
    w.write(len(gens))
    for g in gens:
        mt = np.asarray(g["mt"]).astype(np.int64)
        if mt.shape != (MT_N,):
            raise ValueError(f"generator state must hold {MT_N} words")
        name = g["name"].encode("ascii")
        if len(name) > NAME_LEN:
            raise ValueError(f"generator name longer than {NAME_LEN} characters")
        w.write_bytes(struct.pack("<i", int(g["mti"])) + (mt & 0xFFFFFFFF).astype("<u4").tobytes()
                      + struct.pack("<id", int(g["iset"]), float(g["gset"])) + name.ljust(NAME_LEN))


def _read_generators(r):
    # restore_random (random.f90:66-107)
    n = r.read(("i", 1))[0]
    out = []
    for _ in range(n):
        data = r.read_bytes()
        if len(data) != 4 + 4 * MT_N + 12 + NAME_LEN:
            raise ValueError("random generator record of unexpected length")
        mti = struct.unpack_from("<i", data, 0)[0]
        mt = np.frombuffer(data, "<i4", MT_N, 4).copy()
        iset, gset = struct.unpack_from("<id", data, 4 + 4 * MT_N)
        name = data[4 + 4 * MT_N + 12:].decode("ascii").strip()
        out.append(dict(name=name, mti=mti, mt=mt, iset=iset, gset=gset))
    return out


# ------------------------------------------------------------------ checkpoints
CHK_FILES = ("mrm.chk", "mrm_grain.chk", "mrm_0.chk", "mrm_1.chk")   # checkpoint.f90:8
_EXACT_TAG = b"mrmgexct"


def _last_slot(directory):
    p = os.path.join(directory, CHK_FILES[0])
    if not os.path.exists(p):
        return 1                                   # idd starts at 1 (checkpoint.f90:10)
    try:
        with open(p) as f:
            return int(f.read().split()[2])
    except (IndexError, ValueError):
        return 1


def write_checkpoint(directory, time, iline, mag_layers, generators, grain_layers=None, exact=None):
    """One checkpoint save (checkpoint.f90:106-145).  grain_layers (write_grain's dicts)
    writes mrm_grain.chk, which the reference saves once per run.  The slot after the
    one mrm.chk names gets (time, iline), the magnetisation of mag_layers (write_mag's
    tuples) and the generators; only then is mrm.chk rewritten, so a save cut short
    leaves the previous slot valid.  exact: optional dict(layers=[(m (ng,3), |m| (ng))
    per layer], dt=the integrator's current step) kept bit for bit in trailing records,
    which the reference never reads (it closes the file after the generators); the
    magnetisation records alone renormalise m and carry no step size.  Returns the slot
    written (0 or 1)."""
    if grain_layers is not None:
        write_grain(os.path.join(directory, CHK_FILES[1]), grain_layers)
    idd = 1 - _last_slot(directory)
    with open(os.path.join(directory, CHK_FILES[2 + idd]), "wb") as f:
        w = RecordWriter(f)
        w.write(float(time), int(iline))
        _write_mag_records(w, mag_layers)
        _write_generators(w, generators)
        if exact is not None:
            w.write_bytes(_EXACT_TAG + struct.pack("<id", len(exact["layers"]), float(exact["dt"])))
            for m, mag in exact["layers"]:
                m = np.asarray(m, np.float64).reshape(-1, 3)
                w.write(len(m), m, np.asarray(mag, np.float64).reshape(len(m)))
    # list-directed time, iline, idd; replaced in one rename so it never names a slot half-written
    tmp = os.path.join(directory, CHK_FILES[0] + ".tmp")
    with open(tmp, "w") as f:
        f.write(f" {float(time):.17E} {int(iline):11d} {idd:11d}\n")
    os.replace(tmp, os.path.join(directory, CHK_FILES[0]))
    return idd


def _read_slot(path, n_lay):
    with open(path, "rb") as f:
        r = RecordReader(f)
        time, iline = r.read(("d", 1), ("i", 1))
        mag = _read_mag_records(r, n_lay)
        gens = _read_generators(r)
        exact = None
        try:
            hd = r.read_bytes()
        except EOFError:
            hd = b""
        if hd[:8] == _EXACT_TAG:
            n, dt = struct.unpack_from("<id", hd, 8)
            exact = dict(layers=[], dt=dt)
            for _ in range(n):
                data = r.read_bytes()
                ng = struct.unpack_from("<i", data, 0)[0]
                if len(data) != 4 + 32 * ng:
                    raise ValueError("exact-state record of unexpected length")
                m = np.frombuffer(data, "<f8", 3 * ng, 4).reshape(ng, 3).copy()
                exact["layers"].append((m, np.frombuffer(data, "<f8", ng, 4 + 24 * ng).copy()))
    return dict(time=float(time), iline=int(iline), mag=mag, generators=gens, exact=exact)


def read_checkpoint(directory, n_lay):
    """Checkpoint read as the reference does it (checkpoint.f90:58-99): None unless both
    mrm.chk and mrm_grain.chk exist; the slot mrm.chk names is tried first, the other slot
    if that one is missing or damaged.  Returns a dict with time, iline, slot, grain
    (read_grain's layers), mag (read_mag's tuples), generators (dicts of name, mti, mt,
    iset, gset) and exact (write_checkpoint's exact, or None)."""
    if not all(os.path.exists(os.path.join(directory, n)) for n in CHK_FILES[:2]):
        return None
    grain = read_grain(os.path.join(directory, CHK_FILES[1]), n_lay)
    first = _last_slot(directory)
    for idd in (first, 1 - first):
        p = os.path.join(directory, CHK_FILES[2 + idd])
        if not os.path.exists(p):
            continue
        try:
            out = _read_slot(p, n_lay)
        except (ValueError, EOFError, struct.error):
            continue
        out.update(slot=idd, grain=grain)
        return out
    return None


def grain_dicts(layers, nx, ny, dx, dy, dz, tau=None):
    """write_grain's layer dicts for mrm_amd.abi.LayerArrays layers (the inverse of
    layer_arrays): grain_area = grain volume / dz; tc falls back to the layer's tc_nom."""
    out = []
    for k, la in enumerate(layers):
        ng = la.ng
        d = dict(nx=nx, ny=ny, il=k + 1, dx=dx, dy=dy, grid2grain=la.grid2grain.reshape(ny, nx),
                 grain2grid=la.grain2grid.reshape(-1, 2), grain2gridindex=la.grain2gridindex,
                 grain_area=la.grain_vol / np.broadcast_to(np.asarray(dz, float), (len(layers),))[k], ms=la.ms,
                 tc=la.tc if la.tc is not None else np.full(ng, la.tc_nom),
                 tau=np.zeros(ng) if tau is None else tau[k], hk=la.hk.reshape(ng, 2),
                 uni_axis=None if la.uni_axis is None else la.uni_axis.reshape(ng, 3),
                 cubic_axis=None if la.cubic_axis is None else la.cubic_axis.reshape(ng, 3, 3))
        for key, pre in (("hx", ""), ("uhx", "u"), ("lhx", "l")):
            idx = getattr(la, f"{pre}hxgrainindex")
            if idx is not None:
                d[key] = (idx, getattr(la, f"{pre}hxgrain"), getattr(la, f"{pre}hx"))
        out.append(d)
    return out
