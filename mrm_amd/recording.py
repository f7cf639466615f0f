"""The recording workload on a track: a Karlqvist write head moving down-track, a seeded
NRZ bit sequence written by flipping the head current at bit boundaries, and a
reciprocity readback scan of the written track.

The reference has no head model: its write fields are tables read with READ_FIELD
(columns x, -, y, Hx, Hz, Hy, applied.f90:711-813) that FLD_VELOCITY moves and `set` /
`interpolate` sample bilinearly (applied.f90:1471-1522, 1757-1777), and a bit pattern is
a script of `SET_1FIELD 1 +-1` lines, each followed by an LLG command.  This module writes
the table (a closed-form Karlqvist field on a non-uniform mesh) and drives the per-bit
commands through the same entry points (mrmg_set_source, mrmg_step); the readback is
mrmg_readback with a Gaussian sensitivity.
"""
import numpy as np

from . import abi


def karlqvist(u, z, gap, h_gap):
    """The 2D Karlqvist field of a ring head with gap `gap` and deep-gap field `h_gap` at
    down-track offset u from the gap centre and distance z > 0 below the pole faces:
    longitudinal (along u) and perpendicular (toward the head) components."""
    a = 0.5 * gap
    hl = (h_gap / np.pi) * (np.arctan((u + a) / z) - np.arctan((u - a) / z))
    hp = (h_gap / (2.0 * np.pi)) * np.log(((u - a) ** 2 + z * z) / ((u + a) ** 2 + z * z))
    return hl, hp


def deep_gap_field(peak, gap, fly_height):
    """The deep-gap field whose longitudinal field at the gap centre, fly_height below the
    head, is `peak`: H_l(0, z) = (2 H_g/pi) atan(gap/(2 z))."""
    return peak * np.pi / (2.0 * np.arctan(gap / (2.0 * fly_height)))


def head_nodes(span, fine=0.5e-7, fine_half=60e-7, mid=5e-7, mid_half=600e-7, coarse=50e-7):
    """Down-track table nodes (cm) centred on the gap: `fine` spacing within +-fine_half,
    `mid` spacing to +-mid_half, `coarse` beyond, out to +-span."""
    a = np.arange(0.0, fine_half + 0.5 * fine, fine)
    b = np.arange(a[-1] + mid, mid_half + 0.5 * mid, mid)
    c = np.arange(b[-1] + coarse, span + coarse, coarse)
    half = np.concatenate([a, b, c])
    return np.concatenate([-half[:0:-1], half])


def layer_depths(dz, il_space, fly_height):
    """Distance from the head to the centre of each layer (layer 0 = top)."""
    out, z = [], fly_height
    for k, d in enumerate(dz):
        if k > 0:
            z += dz[k - 1] + il_space[k]
        out.append(z + 0.5 * d)
    return out


def head_sources(nx, ny, dx, dy, dz, il_space=None, gap=20e-7, fly_height=5e-7, peak=15000.0, velocity=2000.0,
                 start=None, polarity=1.0):
    """One READ_FIELD-style table per layer: the Karlqvist field evaluated at the layer's
    depth on `head_nodes`, longitudinal along y (down-track), perpendicular along -z
    (toward the media), uniform across the track (two x nodes spanning it), scaled by the
    write polarity (+1/-1).  The gap starts at y = start (default a quarter of the track)
    and moves at velocity (cm/s) along +y.  Returns a list of abi.SourceArrays."""
    il = np.zeros(len(dz)) if il_space is None else np.asarray(il_space, float)
    hg = deep_gap_field(peak, gap, fly_height)
    yp = head_nodes(ny * dy + 100e-7)
    xp = np.array([-1e-7, nx * dx + 1e-7])
    y0 = 0.25 * ny * dy if start is None else start
    out = []
    for z in layer_depths(list(dz), il, fly_height):
        hl, hp = karlqvist(yp, z, gap, hg)
        vals = np.zeros((len(yp), 2, 3))       # (nyp, nxp, 3): Hx, Hy, Hz at x fastest
        vals[:, :, 1] = hl[:, None]
        vals[:, :, 2] = -hp[:, None]
        out.append(abi.SourceArrays(abi.MRMG_SRC_FIELD_TABLE, xpos=xp, ypos=yp, vals=vals, vel=(0.0, velocity),
                                    offset=(0.0, y0), t_start=-1e22, t_end=1e22, s_start=1.0, s_end=1.0,
                                    scale=float(polarity)))
    return out


def nrz_bits(n, seed=7):
    """n bits of a 64-bit Galois LFSR (x^64 + x^63 + x^61 + x^60 + 1, a maximal-length
    polynomial) started from `seed`: the least significant bit of the state after each
    shift."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    if s == 0:
        raise ValueError("the LFSR seed must be non-zero")
    out = []
    for _ in range(n):
        lsb = s & 1
        s >>= 1
        if lsb:
            s ^= 0xD800000000000000
        out.append(s & 1)
    return np.array(out, dtype=np.int8)


def write_nrz(x, bits, bit_length, velocity, dt, sources_of, t0=0.0, philox_seed=None, noise=None):
    """Write `bits` (0/1) with the moving head: for each bit the head polarity is set
    (+1 for 1, -1 for 0; the SET_1FIELD +-1 of a bit script) and the head moves one bit
    length, round(bit_length/velocity/dt) steps.  sources_of(polarity) gives the per-layer
    tables; only a polarity change sends new tables.  Returns the end time."""
    steps = max(1, int(round(bit_length / velocity / dt)))
    t, last = t0, None
    x.set_llg(dt=dt)
    for b in bits:
        pol = 1.0 if b else -1.0
        if pol != last:
            for k, src in enumerate(sources_of(pol)):
                x.set_source(k, 0, src)
            last = pol
        d = x.step(steps, t0=t, philox_seed=philox_seed, noise=None if noise is None else noise(steps))
        t = d.t
    return t


def readback_scan(x, y_from, y_to, step, x_centre, fwhm=15e-7, direction=(0.0, 0.0, 1.0), weights=None):
    """The reciprocity readback along the track at x_centre from y_from to y_to: a Gaussian
    sensitivity of the given FWHM in both directions.  Returns (positions, signal in emu)."""
    ys = np.arange(y_from, y_to + 0.5 * step, step)
    sigma = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    v = x.readback(np.full(ys.size, x_centre), ys, (sigma, sigma), direction=direction, weights=weights)
    return ys, v
