"""Synthetic granular media in the reference's grain-file layout.

The reference builds media with its own grain-growth generator (construct_grain.f90),
which is outside the time-stepping path; here the same data structures are produced
from a seeded periodic Voronoi tessellation so that both the HIP path and the oracle get
identical inputs:

* grid2grain (ny,nx) 1-based grain numbers, 0 for the one-cell non-magnetic boundaries;
* grain2gridindex / grain2grid in the order file_grain builds them (data.f90:276-284):
  grain by grain, cells in x-fastest row-major order;
* lateral exchange CSR with hx_ij = ex * L_ij / (Ms_i * A_i), the reciprocal form of
  construct_hx_transverse (construct_grain.f90:1355-1370).  Grains separated by a
  one-cell boundary are coupled through it (L_ij = dx per boundary cell touching both),
  which the reference's direct-contact builder would not find;
* vertical exchange of columnar grains, hx = ex_upper * A_overlap / (Ms * dz * A)
  (construct_grain.f90:1144-1145), with ex_upper = 2 * A_vertical / dz inside a physical
  layer (input.f90:328).
"""
import numpy as np

try:
    from scipy.spatial import cKDTree
except Exception:  # pragma: no cover - scipy is part of the image
    cKDTree = None

from . import abi


def voronoi_grid(nx, ny, dx, dy, pitch, seed, boundary=True):
    """Periodic Voronoi map of ~ (nx*dx*ny*dy)/pitch^2 grains; returns grid2grain (ny,nx)."""
    rng = np.random.default_rng(seed)
    lx, ly = nx * dx, ny * dy
    nseed = max(1, int(round(lx * ly / (pitch * pitch))))
    pts = rng.random((nseed, 2)) * [lx, ly]
    xs = (np.arange(nx) + 0.5) * dx
    ys = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(xs, ys)
    cells = np.stack([X.ravel(), Y.ravel()], axis=1)
    tree = cKDTree(pts, boxsize=[lx, ly])
    _, owner = tree.query(cells)
    owner = owner.reshape(ny, nx)
    if boundary:
        right = np.roll(owner, -1, axis=1)
        up = np.roll(owner, -1, axis=0)
        nonmag = (owner != right) | (owner != up)
        owner = np.where(nonmag, -1, owner)
    # renumber 1..ng in order of first appearance (x fastest)
    flat = owner.ravel()
    uniq, first = np.unique(flat[flat >= 0], return_index=True)
    order = uniq[np.argsort(np.flatnonzero(flat >= 0)[first])]
    remap = -np.ones(nseed, dtype=np.int64)
    remap[order] = np.arange(1, len(order) + 1)
    g2g = np.where(owner >= 0, remap[np.maximum(owner, 0)], 0).astype(np.int32)
    return g2g


def grain_cells(grid2grain):
    """grain2gridindex (ng+1) and grain2grid (ncell,2) with 1-based (i,j)."""
    ny, nx = grid2grain.shape
    ng = int(grid2grain.max())
    flat = grid2grain.ravel()
    idx = np.flatnonzero(flat > 0)
    gr = flat[idx]
    order = np.argsort(gr, kind="stable")   # stable keeps x-fastest order inside a grain
    idx = idx[order]
    gr = gr[order]
    counts = np.bincount(gr, minlength=ng + 1)[1:]
    g2gi = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    j, i = np.divmod(idx, nx)
    g2g = np.stack([i + 1, j + 1], axis=1).astype(np.int32)
    return g2gi, g2g, counts


def lateral_pairs(grid2grain):
    """Shared boundary count between grains: across 1-cell boundaries and direct contact."""
    ny, nx = grid2grain.shape
    pairs = {}

    def add(a, b, w):
        if a > 0 and b > 0 and a != b:
            key = (min(a, b), max(a, b))
            pairs[key] = pairs.get(key, 0.0) + w

    g = grid2grain
    left, right = np.roll(g, 1, axis=1), np.roll(g, -1, axis=1)
    down, up = np.roll(g, 1, axis=0), np.roll(g, -1, axis=0)
    # boundary cells: grains on opposite sides
    bj, bi = np.nonzero(g == 0)
    for j, i in zip(bj, bi):
        add(left[j, i], right[j, i], 1.0)
        add(down[j, i], up[j, i], 1.0)
    # direct contacts (no boundary between them)
    for a, b in ((g, right), (g, up)):
        m = (a > 0) & (b > 0) & (a != b)
        for x, y in zip(a[m], b[m]):
            add(int(x), int(y), 1.0)
    return pairs


def lateral_exchange(grid2grain, ms, area, ex, dx):
    """Intralayer CSR (hxgrainindex, hxgrain 1-based, hx Oe)."""
    ng = int(grid2grain.max())
    nb = [[] for _ in range(ng + 1)]
    for (a, b), cnt in lateral_pairs(grid2grain).items():
        L = cnt * dx
        nb[a].append((b, ex * L / (ms[a - 1] * area[a - 1])))
        nb[b].append((a, ex * L / (ms[b - 1] * area[b - 1])))
    idx = [0]
    gr, w = [], []
    for n in range(1, ng + 1):
        for b, v in sorted(nb[n]):
            gr.append(b)
            w.append(v)
        idx.append(len(gr))
    return (np.asarray(idx, np.int32), np.asarray(gr, np.int32), np.asarray(w, np.float64))


def cone_axes(ng, cone_deg, rng):
    """Easy axes uniform in solid angle inside a cone of half angle cone_deg about +z."""
    cmin = np.cos(np.radians(cone_deg))
    ct = 1.0 - rng.random(ng) * (1.0 - cmin)
    st = np.sqrt(1.0 - ct * ct)
    ph = rng.random(ng) * 2 * np.pi
    return np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)


class Media:
    """Multi-layer columnar media: one Voronoi map shared by all layers."""

    def __init__(self, nx, ny, dx, dy, dz, n_lay, grid2grain, layers):
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        self.dz = np.broadcast_to(np.asarray(dz, float), (n_lay,)).copy()
        self.n_lay = n_lay
        self.grid2grain = grid2grain
        self.layers = layers          # list of abi.LayerArrays
        self.ngs = [L.ng for L in layers]

    @property
    def G(self):
        return sum(self.ngs)

    @property
    def cells(self):
        return self.nx * self.ny * self.n_lay


def make_media(nx, ny, dx=2e-7, dy=2e-7, dz=8e-7, n_lay=1, pitch=8e-7, ms=700.0, ku=6e6, ku_sigma=0.1,
               cone_deg=3.0, a_ex=1e-6, intergrain_frac=0.05, boundary_width=2e-7, layer_ku=None,
               physical=None, alpha=0.02, gamma=1.76e7, thm_alpha=None, temperature=0.0, thermal=False,
               seed=1, boundary=True, exchange=True, tc_nom=None, tc_sigma=0.0, spin=0.5, msexp=-1.0,
               hkexp=-1.0, use_mag=False, hk_table=None):
    """Build media following the recipe of the benchmark configurations.

    layer_ku: per layer Ku (erg/cc); physical: physical-layer id of each layer (sub-layers
    of one physical layer are coupled with 2*A/dz, neighbouring physical layers with
    intergrain_frac of that).  tc_nom (K): per-grain Curie temperatures tc_nom (1 +
    tc_sigma z) for LLB and temperature scaling (Ms, Hk ~ (1-T/Tc)^exp, or Ms from |m|
    with use_mag); None leaves Tc at 1e22 and scaling off.  hk_table: (T/Tc nodes,
    Hk scalings), a READ_SCALE table for Hk (temp_scaling.f90:397-416).
    """
    g2g = voronoi_grid(nx, ny, dx, dy, pitch, seed, boundary=boundary)
    g2gi, cells, counts = grain_cells(g2g)
    ng = len(counts)
    area = counts * dx * dy
    dzs = np.broadcast_to(np.asarray(dz, float), (n_lay,)).copy()
    layer_ku = [ku] * n_lay if layer_ku is None else list(layer_ku)
    physical = list(range(n_lay)) if physical is None else list(physical)
    rng_k = np.random.default_rng(seed + 1)
    rng_a = np.random.default_rng(seed + 2)
    z = rng_k.standard_normal(ng)
    axes = cone_axes(ng, cone_deg, rng_a)
    ex = intergrain_frac * a_ex / boundary_width
    layers = []
    msv = np.full(ng, float(ms))
    lat = lateral_exchange(g2g, msv, area, ex, dx) if exchange else (None, None, None)
    llb = {}
    if tc_nom is not None:
        zt = np.random.default_rng(seed + 3).standard_normal(ng)
        llb = dict(tc=tc_nom * (1.0 + tc_sigma * zt), spin=spin, tc_nom=tc_nom, msexp=msexp, hkexp=hkexp,
                   use_mag=use_mag)
        if hk_table is not None:
            llb["tables"] = {"hk": hk_table}
    one = np.arange(0, ng + 1, dtype=np.int32)
    self_nb = np.arange(1, ng + 1, dtype=np.int32)
    for k in range(n_lay):
        kuk = layer_ku[k] * np.exp(ku_sigma * z - 0.5 * ku_sigma ** 2)   # mean-preserving lognormal
        hk = np.stack([2.0 * kuk / ms, np.zeros(ng)], axis=1)
        kw = dict(ng=ng, grid2grain=g2g, grain2gridindex=g2gi, grain2grid=cells,
                  grain_vol=area * dzs[k], ms=msv, hk=hk, uni_axis=axes,
                  alpha=alpha, gamma=gamma, thm_alpha=alpha if thm_alpha is None else thm_alpha,
                  floor_ms=0.01, thermal=thermal, temp=np.full(ng, float(temperature)), **llb)
        if exchange:
            kw.update(hxgrainindex=lat[0], hxgrain=lat[1], hx=lat[2])
            for nbk, name in ((k - 1, "u"), (k + 1, "l")):
                if 0 <= nbk < n_lay:
                    exu = 2.0 * a_ex / dzs[k]
                    if physical[nbk] != physical[k]:
                        exu *= intergrain_frac
                    # overlap area = grain area for identical columnar maps
                    kw[name + "hxgrainindex"] = one
                    kw[name + "hxgrain"] = self_nb
                    kw[name + "hx"] = np.full(ng, exu / (ms * dzs[k]))
        layers.append(abi.LayerArrays(**kw))
    return Media(nx, ny, dx, dy, dzs, n_lay, g2g, layers)


def config_media(name, **over):
    """Media of the named benchmark configuration."""
    if name == "cfg0":       # 64x64x1, 2x2x8 nm, Ku 6e6, T=0
        kw = dict(nx=64, ny=64, dx=2e-7, dy=2e-7, dz=8e-7, n_lay=1, ku=6e6)
    elif name == "cfg1":     # 256x256x4 of 2 nm, two 4 nm layers, Ku 2e6 top / 1e7 bottom, T=300
        kw = dict(nx=256, ny=256, dx=2e-7, dy=2e-7, dz=2e-7, n_lay=4, layer_ku=[2e6, 2e6, 1e7, 1e7],
                  physical=[0, 0, 1, 1], temperature=300.0, thermal=True)
    elif name == "cfg2":     # 512x4096x4 of 1.5 nm, config-1 media
        kw = dict(nx=512, ny=4096, dx=1.5e-7, dy=1.5e-7, dz=1.5e-7, n_lay=4, layer_ku=[2e6, 2e6, 1e7, 1e7],
                  physical=[0, 0, 1, 1], temperature=300.0, thermal=True)
    elif name == "cfg4":     # 2048x2048x8 of 1 nm, FePt-like LLB media, Tc = 700 K
        kw = dict(nx=2048, ny=2048, dx=1e-7, dy=1e-7, dz=1e-7, n_lay=8, pitch=8e-7, boundary_width=1e-7,
                  ms=1100.0, ku=5e7, physical=[0] * 8, alpha=0.1, temperature=300.0, thermal=True,
                  tc_nom=700.0, spin=0.5, use_mag=True, hk_table=equilibrium_table())
    else:
        raise ValueError(name)
    kw.update(over)
    return make_media(**kw)


def equilibrium_magnetisation(t_over_tc):
    """m_e at T/Tc: the zero-field fixed point m = L(3 m Tc/T) of the mean field that the
    reference's LLB relaxes |m| to (L the Langevin function of llg.f90:704-710; lambda c =
    3 Tc/T for any spin S, grain.f90:61 with llg.f90:932); 0 at and above Tc."""
    tau = float(t_over_tc)
    if tau >= 1.0:
        return 0.0
    if tau <= 0.0:
        return 1.0

    def lang(x):
        return 1.0 / np.tanh(x) - 1.0 / x if abs(x) > 1e-6 else x / 3.0

    lo, hi = 1e-12, 1.0
    for _ in range(200):            # m - L(3m/tau) < 0 below the root, > 0 above
        mid = 0.5 * (lo + hi)
        if mid - lang(3.0 * mid / tau) > 0.0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def equilibrium_table(n=41):
    """(T/Tc nodes, m_e) on [0, 1.05]: the READ_SCALE form of Hk(T)/Hk(0) = m_e(T) for
    Ku(T) ~ m_e^2 with Ms(T) ~ m_e (Hk = 2 Ku/Ms)."""
    t = np.concatenate([np.linspace(0.0, 0.95, n - 6), np.linspace(0.96, 1.0, 5), [1.05]])
    return t, np.array([equilibrium_magnetisation(x) for x in t])


def laser_write(n_lay, nx, ny, dx, dy, peak_dtemp=450.0, fwhm=40e-7, velocity=1500.0, h_write=8000.0,
                centre=None):
    """The heat-assisted write of the LLB configuration: a Gaussian laser spot (the
    super-Gaussian source with power 2, set_sg/get_sg, applied.f90:1778-1811) of the given
    FWHM and peak temperature rise over the ambient, moving down-track (y) at velocity
    (cm/s) from centre at t = 0, on every layer, and a uniform write field h_write along -z.
    Returns (sources, ufields), one per layer."""
    s = fwhm / (2.0 * np.sqrt(np.log(2.0)))       # exp(-(u/s)^2) has this FWHM
    cx, cy = centre if centre is not None else (0.5 * nx * dx, 0.5 * ny * dy)
    srcs = [abi.SourceArrays(abi.MRMG_SRC_TEMP_SG, vel=(0.0, velocity), offset=(cx, cy), sigma=(s, s), expn=2,
                             dtemp=peak_dtemp, t_start=-1e22, t_end=1e22) for _ in range(n_lay)]
    ufs = [abi.make_ufield((0.0, 0.0, -1.0), scale=h_write) for _ in range(n_lay)]
    return srcs, ufs


def ramp_field(n_lay, tilt_deg=5.0, h_max=20000.0, t_ramp=2e-9):
    """Uniform field tilted tilt_deg from -z, ramped 0 -> h_max over t_ramp (SET_1FIELD ramp)."""
    d = (np.sin(np.radians(tilt_deg)), 0.0, -np.cos(np.radians(tilt_deg)))
    return [abi.make_ufield(d, t_start=0.0, t_end=t_ramp, s_start=0.0, s_end=h_max) for _ in range(n_lay)]
