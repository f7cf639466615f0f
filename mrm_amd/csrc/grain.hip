// grain.hip - the grain kernels (field assembly of get_field, the thermal draw, the LLG and
// LLB pushes, get_mag's scatter), the loop-top statistics, the readback and switching
// record kernels, and the Engine members that launch them.
#include "engine.h"
#include "device_common.h"

__global__ void k_philox_dump(unsigned long long seed, unsigned long long step0, int nsteps, int G, double* out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nsteps * G) return;
    int it = (int)(i / G), g = (int)(i - (long long)it * G);
    double n[3];
    philox_normals3(seed, step0 + it, g, n);
    out[3 * i] = n[0]; out[3 * i + 1] = n[1]; out[3 * i + 2] = n[2];
}

// ------------------------------------------------------------------------- grain kernel
// LPG lanes cooperate on one grain: the cell gather and the exchange sums are split over
// the lanes and reduced with a fixed xor-butterfly (deterministic); the rest of the grain
// work is computed redundantly by every lane and written by lane 0.
template <int LPG, typename T>
__device__ __forceinline__ T lane_sum(T v)
{
#pragma unroll
    for (int o = LPG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPG);
    return v;
}

// neighbour terms hx(j) ms_scale(nb) m(nb) of entries [j0, j1); the Ms scale is the
// state's row ST_MSC.  Each lane takes entries j0+lane, j0+lane+LPG, ... in order; two
// per round, with all index and value loads of a round issued before they are used.
template <typename T, int LPG>
__device__ __forceinline__ void csr_sum(int j0, int j1, const int* __restrict__ nb_, const T* __restrict__ w_,
                                        const T* __restrict__ mf, int G, int lane, T* h)
{
    T s0 = 0, s1 = 0, s2 = 0;
    const T* __restrict__ msc = mf + (size_t)ST_MSC * G;
    for (int ja = j0 + lane; ja < j1; ja += 2 * LPG) {
        const int jb = ja + LPG;
        const bool hb = jb < j1;
        const int na = nb_[ja], nbb = hb ? nb_[jb] : 0;
        const T wa = w_[ja], wb = hb ? w_[jb] : (T)0;
        const T ca = msc[na], cb = msc[nbb];
        const T ax = mf[na], ay = mf[G + na], az = mf[2 * G + na];
        const T bx = mf[nbb], by = mf[G + nbb], bz = mf[2 * G + nbb];
        T w = wa * ca;
        s0 += w * ax;
        s1 += w * ay;
        s2 += w * az;
        if (hb) {
            w = wb * cb;
            s0 += w * bx;
            s1 += w * by;
            s2 += w * bz;
        }
    }
    h[0] += lane_sum<LPG>(s0);
    h[1] += lane_sum<LPG>(s1);
    h[2] += lane_sum<LPG>(s2);
}

// Value of a moving source at cell (i, j) of layer k from the stage's axis tables:
// interpolate (applied.f90:1471-1522) for tables, get_sg (1792-1811) for the super-Gaussian.
__device__ __forceinline__ double src_cell(const SrcAxes& ax, const SrcParams& sp, int k, int slot, int ts, int i,
                                           int j, int comp, int nv)
{
    const size_t bx = ((size_t)(k * 2 + slot) * 2 + ts) * ax.nx, by = ((size_t)(k * 2 + slot) * 2 + ts) * ax.ny;
    if (sp.kind == MRMG_SRC_TEMP_SG) return ax.ya[by + j] * ax.xa[bx + i];
    const int ix = ax.xi[bx + i], iy = ax.yi[by + j];
    const double fx = ax.xa[bx + i], yf = ax.ya[by + j], yfc = ax.yb[by + j];
    const double* v = ax.vals + sp.voff;
    const size_t a00 = (size_t)(ix - 1) + (size_t)sp.nxp * (iy - 1), a10 = a00 + 1, a01 = a00 + sp.nxp, a11 = a01 + 1;
    return fx * yf * v[nv * a00 + comp] + (1.0 - fx) * yf * v[nv * a10 + comp] + fx * yfc * v[nv * a01 + comp] +
           (1.0 - fx) * yfc * v[nv * a11 + comp];
}

// The super-Gaussian temperature source (slot 1) of layer k summed over the run of cells
// i0 .. i0+len-1 in row j at time ts: ya[j] times the run's share of the running sums of
// xa (the cells' own values ya[j] xa[i], get_sg, applied.f90:1792-1811, added up).
__device__ __forceinline__ double sg_run(const SrcAxes& ax, int k, int ts, int i0, int len, int j)
{
    const size_t b = (size_t)(k * 2 + 1) * 2 + ts;
    const double* xc = ax.xc + b * (ax.nx + 1);
    return ax.ya[b * ax.ny + j] * (xc[i0 + len] - xc[i0]);
}

// (1 - T/Tc)^e clamped to [0, 1] (temp_scaling.f90:390-395)
__device__ __forceinline__ double temp_power(double T, double Tc, double e)
{
    double x = 1.0 - T / Tc;
    x = x < 1.0 ? (x > 0.0 ? x : 0.0) : 1.0;
    return pow(x, e);
}

// find_temp_scaling of one quantity (0 Ms, 1 alpha, 2 Hk) at temperature T: the power law
// for Ms and Hk when its exponent is >= 0, else the layer's table
__device__ __forceinline__ bool scale_follows_t(const LayerParams& L, int q)
{
    if (!L.tscale) return false;
    if (q == 0) return !L.use_mag && (L.msexp >= 0.0 || L.tab_n[0] > 1);
    if (q == 1) return L.tab_n[1] > 1;
    return L.hkexp >= 0.0 || L.tab_n[2] > 1;
}
__device__ __forceinline__ double scale_at_t(const LayerParams& L, const double* tabs, int q, double T, double Tc)
{
    const double e = q == 0 ? L.msexp : L.hkexp;
    if (q != 1 && e >= 0.0) return temp_power(T, Tc, e);
    const double* tt = tabs + L.tab_off[q];
    return table_eval(tt, tt + L.tab_n[q], L.tab_n[q], T / Tc);
}

// tout (may be null): the grain temperature at the evaluation and at the next evaluation
// from the layer's temperature source (field.f90:545-556); unchanged without one.
// Loads are arranged for memory-level parallelism: the offsets and constants first, then
// per round (each lane two cells and two exchange entries) all index loads, then all value
// loads, then the sums.  Out-of-range positions load a valid entry (the arrays carry one
// spare entry at the end) and add nothing.
// Per-grain offsets and field constants, loaded by the caller with the grain's state.
template <typename T> struct GrainPre {
    int b0, b1, r0, r1, x0, x1;   // cell range (grain2gridindex: its count), run range, exchange CSR range
    T hk_s, hk1, hk2, u[3], happ[3];
};

template <typename T>
__device__ __forceinline__ void grain_pre(const GrainData<T>& gd, int g, GrainPre<T>& P)
{
    const int G = gd.G;
    P.b0 = gd.gc_off[g]; P.b1 = gd.gc_off[g + 1];
    P.r0 = gd.run_off[g]; P.r1 = gd.run_off[g + 1];
    P.x0 = gd.x_off[g]; P.x1 = gd.x_off[g + 1];
    P.hk_s = gd.hk_scale[g]; P.hk1 = gd.hk1[g]; P.hk2 = gd.hk2[g];
    P.u[0] = gd.uni[g]; P.u[1] = gd.uni[G + g]; P.u[2] = gd.uni[2 * G + g];
    P.happ[0] = P.happ[1] = P.happ[2] = 0;
    if (gd.happ) {
        P.happ[0] = gd.happ[g]; P.happ[1] = gd.happ[G + g]; P.happ[2] = gd.happ[2 * G + g];
    }
}

template <typename T, int LPG, typename TC>
__device__ __forceinline__ void field_at(const GrainData<T>& gd, const LayerParams& L, int k, int g, int lane,
                                         const T* __restrict__ mf, const double* __restrict__ Hr,
                                         double t, const T* m, T* h, double* tout, GrainPre<T>& P)
{
    const int G = gd.G;
    const bool fsrc = L.src[0].kind == MRMG_SRC_FIELD_TABLE, tsrc = L.src[1].kind != MRMG_SRC_NONE;
    const int b0 = P.b0, b1 = P.b1, x0 = P.x0, x1 = P.x1;
    const T hk_s = P.hk_s, hk1 = P.hk1, hk2 = P.hk2;
    const T u[3] = {P.u[0], P.u[1], P.u[2]};
    T happ[3] = {P.happ[0], P.happ[1], P.happ[2]};
    const T* __restrict__ msc = mf + (size_t)ST_MSC * G;
    // get_magnetostatic gather (field.f90:463-468), summed over the grain's runs (each run
    // the sum of its cells, written by the row c2r), and the exchange CSR
    // (field.f90:326-363: intralayer, then the layers above and below, merged per grain)
    const int r0 = P.r0, r1 = P.r1;
    const size_t NR = (size_t)gd.nruns;
    T hd[3] = {0, 0, 0}, hx[3] = {0, 0, 0};
    double hs[3] = {0, 0, 0}, ts0 = 0, ts1 = 0;
    if (!fsrc && !tsrc) {
        T s0 = 0, s1 = 0, s2 = 0, e0 = 0, e1 = 0, e2 = 0;
        const int nrun = r1 - r0, ne = x1 - x0;
        const int ncr = (nrun + 2 * LPG - 1) / (2 * LPG), ner = (ne + 2 * LPG - 1) / (2 * LPG);
        const int nr = ncr > ner ? ncr : ner;
        for (int r = 0; r < nr; r++) {
            const int qa = r0 + lane + 2 * r * LPG, qb = qa + LPG;
            const int ja = x0 + lane + 2 * r * LPG, jb = ja + LPG;
            const bool hqa = qa < r1, hqb = qb < r1, hja = ja < x1, hjb = jb < x1;
            const int na = gd.x_nb[hja ? ja : x0], nbb = gd.x_nb[hjb ? jb : x0];
            const T wa = gd.x_w[hja ? ja : x0], wb = gd.x_w[hjb ? jb : x0];
            T c0 = 0, c1 = 0, c2 = 0, d0 = 0, d1 = 0, d2 = 0;
            if (Hr) {
                const int ra = hqa ? qa : r0, rb = hqb ? qb : r0;
                c0 = (T)Hr[ra]; c1 = (T)Hr[NR + ra]; c2 = (T)Hr[2 * NR + ra];
                d0 = (T)Hr[rb]; d1 = (T)Hr[NR + rb]; d2 = (T)Hr[2 * NR + rb];
            }
            const T pa = msc[na], pb = msc[nbb];
            const T ax = mf[na], ay = mf[G + na], az = mf[2 * G + na];
            const T bx = mf[nbb], by = mf[G + nbb], bz = mf[2 * G + nbb];
            if (hqa) { s0 += c0; s1 += c1; s2 += c2; }
            if (hqb) { s0 += d0; s1 += d1; s2 += d2; }
            if (hja) {
                const T w = wa * pa;
                e0 += w * ax; e1 += w * ay; e2 += w * az;
            }
            if (hjb) {
                const T w = wb * pb;
                e0 += w * bx; e1 += w * by; e2 += w * bz;
            }
        }
        if (Hr) {
            const T cnt = (T)(b1 - b0);
            hd[0] = lane_sum<LPG>(s0) / cnt;
            hd[1] = lane_sum<LPG>(s1) / cnt;
            hd[2] = lane_sum<LPG>(s2) / cnt;
        }
        hx[0] = lane_sum<LPG>(e0);
        hx[1] = lane_sum<LPG>(e1);
        hx[2] = lane_sum<LPG>(e2);
    } else {
        // moving sources: a table is sampled at every cell of the grain's runs; the
        // separable super-Gaussian temperature is summed per run from its running sums
        T s0 = 0, s1 = 0, s2 = 0;
        const bool tsg = tsrc && L.src[1].kind == MRMG_SRC_TEMP_SG;
        const bool per_cell = fsrc || (tsrc && !tsg);
        for (int q = r0 + lane; q < r1; q += LPG) {
            if (Hr) {
                s0 += (T)Hr[q];
                s1 += (T)Hr[NR + q];
                s2 += (T)Hr[2 * NR + q];
            }
            const int geo = gd.run_geo[q], j = gd.run_row[q];
            const int i0 = geo & 0xffff, len = geo >> 16;
            if (tsg) {
                ts0 += sg_run(gd.ax, k, 0, i0, len, j);
                ts1 += sg_run(gd.ax, k, 1, i0, len, j);
            }
            if (per_cell)
                for (int i = i0; i < i0 + len; i++) {
                    if (fsrc)
#pragma unroll
                        for (int c = 0; c < 3; c++) hs[c] += src_cell(gd.ax, L.src[0], k, 0, 0, i, j, c, 3);
                    if (tsrc && !tsg) {
                        ts0 += src_cell(gd.ax, L.src[1], k, 1, 0, i, j, 0, 1);
                        ts1 += src_cell(gd.ax, L.src[1], k, 1, 1, i, j, 0, 1);
                    }
                }
        }
        const T cnt = (T)(b1 - b0);
        if (Hr) {
            hd[0] = lane_sum<LPG>(s0) / cnt;
            hd[1] = lane_sum<LPG>(s1) / cnt;
            hd[2] = lane_sum<LPG>(s2) / cnt;
        }
        const double dc = (double)(b1 - b0);
        if (fsrc)
#pragma unroll
            for (int c = 0; c < 3; c++) hs[c] = lane_sum<LPG>(hs[c]) / dc;
        if (tsrc && tout) {
            const double tst = gd.temp[g];
            tout[0] = lane_sum<LPG>(ts0) / dc + tst;
            tout[1] = lane_sum<LPG>(ts1) / dc + tst;
        }
        csr_sum<T, LPG>(x0, x1, gd.x_nb, gd.x_w, mf, G, lane, hx);
    }
    // get_applied: static per-grain field + source average + uniform ramped field
    // (applied.f90:871-883)
    if (fsrc)
#pragma unroll
        for (int i = 0; i < 3; i++) happ[i] = happ[i] + (T)hs[i];
    if (L.has_uf) {
        const double sc = tscale_eval(L.uf_fn, L.uf_t0, L.uf_t1, L.uf_s0, L.uf_s1, L.uf_scale, t);
#pragma unroll
        for (int i = 0; i < 3; i++) happ[i] = happ[i] + (T)(sc * L.uf_dir[i]);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) h[i] = happ[i] + hd[i];
    // find_temp_scaling of Hk at the evaluation's temperature (temp_scaling.f90:392-395)
    T hks = hk_s;
    if (tsrc && tout && scale_follows_t(L, 2)) hks = (T)scale_at_t(L, gd.tabs, 2, tout[0], gd.tc[g]);
    // get_cubic_hk (field.f90:377-395)
    if (L.has_cub && hks > (T)0) {
        const T* A = gd.cub;
        T d1[3], d2[3], d3[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            d1[c] = A[(3 * c + 0) * G + g] * m[0] + A[(3 * c + 1) * G + g] * m[1] + A[(3 * c + 2) * G + g] * m[2];
#pragma unroll
        for (int c = 0; c < 3; c++) d2[c] = d1[c] * d1[c];
        d3[0] = d1[0] * (hk1 * (d2[1] + d2[2]) + hk2 * d2[1] * d2[2]);
        d3[1] = d1[1] * (hk1 * (d2[0] + d2[2]) + hk2 * d2[0] * d2[2]);
        d3[2] = d1[2] * (hk1 * (d2[0] + d2[1]) + hk2 * d2[0] * d2[1]);
#pragma unroll
        for (int r = 0; r < 3; r++)
            h[r] = h[r] - (A[(r + 0) * G + g] * d3[0] + A[(r + 3) * G + g] * d3[1] + A[(r + 6) * G + g] * d3[2]) * hks;
    }
    // get_uni_hk (field.f90:366-375)
    if (L.has_uni && hks > (T)0) {
        const T p = dot3(m, u), yy = (T)1 - p * p;
        const T s = p * (hk1 + hk2 * yy) * hks;
#pragma unroll
        for (int i = 0; i < 3; i++) h[i] = h[i] + s * u[i];
    }
    // exchange, summed over the merged CSR above
#pragma unroll
    for (int i = 0; i < 3; i++) h[i] = h[i] + hx[i];
}

// Grain temperature of a layer's temperature source at the two times of the stage's axes
// (lanes split the grain's cells; every lane gets the sums): tout = mean + static temp.
template <typename T, int LPG>
__device__ __forceinline__ void src_temp(const GrainData<T>& gd, const LayerParams& L, int g, int lane, double* tout)
{
    const int k = gd.layer_of[g];
    const int b0 = gd.gc_off[g], b1 = gd.gc_off[g + 1], r0 = gd.run_off[g], r1 = gd.run_off[g + 1];
    double ts0 = 0, ts1 = 0;
    const bool tsg = L.src[1].kind == MRMG_SRC_TEMP_SG;
    for (int q = r0 + lane; q < r1; q += LPG) {
        const int geo = gd.run_geo[q], j = gd.run_row[q];
        const int i0 = geo & 0xffff, len = geo >> 16;
        if (tsg) {
            ts0 += sg_run(gd.ax, k, 0, i0, len, j);
            ts1 += sg_run(gd.ax, k, 1, i0, len, j);
            continue;
        }
        for (int i = i0; i < i0 + len; i++) {
            ts0 += src_cell(gd.ax, L.src[1], k, 1, 0, i, j, 0, 1);
            ts1 += src_cell(gd.ax, L.src[1], k, 1, 1, i, j, 0, 1);
        }
    }
    const double dc = (double)(b1 - b0), tst = gd.temp[g];
    tout[0] = lane_sum<LPG>(ts0) / dc + tst;
    tout[1] = lane_sum<LPG>(ts1) / dc + tst;
}

// get_thermal's standard deviation (field.f90:407-423) for a grain whose temperature moves
__device__ __forceinline__ double thermal_sigma_dev(double tsig, double floor_ms, double as, double T, double ms,
                                                    double vol, bool no_ms, double mag, double msc)
{
    double x;
    if (no_ms) x = as * T / (ms * vol);
    else if (mag != 1.0) x = as * T / (ms * vol * fmax(floor_ms, fabs(mag)));
    else if (msc > 0.0) x = as * T / (ms * vol * fmax(floor_ms, msc));
    else x = 0.0;
    return sqrt(x) * tsig;
}

__device__ __forceinline__ double ipow_dev(double x, int n)
{
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r = r * x;
        x = x * x;
        n >>= 1;
    }
    return r;
}

// Axis tables of every moving source at the evaluation time t0 and the next evaluation
// time t1: set() (applied.f90:1757-1776, node search by bisection, which finds the index
// its monotone scan finds for increasing nodes) or set_sg() (1778-1790).
__global__ void k_src_axes(const LayerParams* __restrict__ lp, int nl, int nx, int ny, double dx, double dy,
                           double t0, double t1, int* __restrict__ xi, double* __restrict__ xa, int* __restrict__ yi,
                           double* __restrict__ ya, double* __restrict__ yb, const double* __restrict__ pos)
{
    const int n = nx + ny;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nl * 4 * n) return;
    int r = tid;
    const int e = r % n; r /= n;
    const int ts = r % 2; r /= 2;
    const int slot = r % 2; r /= 2;
    const int k = r;
    const SrcParams& sp = lp[k].src[slot];
    if (sp.kind == MRMG_SRC_NONE) return;
    const double t = ts ? t1 : t0;
    const double sc = tscale_eval(sp.fn, sp.t0, sp.t1, sp.s0, sp.s1, sp.scale, t);
    const bool isx = e < nx;
    const int i = isx ? e : e - nx;   // 0-based cell index; the reference's i is i+1
    const size_t base = (size_t)(k * 2 + slot) * 2 + ts;
    if (sp.kind == MRMG_SRC_TEMP_SG) {
        if (isx) {
            const double del = sp.ref_x - sp.vel_x * (t - sp.t_origin) - sp.offset_x;
            xa[base * nx + i] = exp(-ipow_dev((dx * ((i + 1) - 0.5) + del) * sp.isigx, sp.expn));
        } else {
            const double del = sp.ref_y - sp.vel_y * (t - sp.t_origin) - sp.offset_y;
            ya[base * ny + i] = sp.dtemp * sc * exp(-ipow_dev((dy * ((i + 1) - 0.5) + del) * sp.isigy, sp.expn));
        }
        return;
    }
    const double off = isx ? sp.offset_x + sp.vel_x * (t - sp.t_origin) - sp.ref_x
                           : sp.offset_y + sp.vel_y * (t - sp.t_origin) - sp.ref_y;
    const double dd = isx ? dx : dy;
    const int npos = isx ? sp.nxp : sp.nyp;
    const double* P = pos + (isx ? sp.poff_x : sp.poff_y);
    const double x = ((i + 1) - .5) * dd - off;
    int lo = 2, hi = npos;   // first i0 in [2, npos] with P(i0) > x, else npos (1-based)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (P[mid - 1] > x) hi = mid;
        else lo = mid + 1;
    }
    const int i0 = lo;
    const int ind = npos - 1 < i0 - 1 ? npos - 1 : i0 - 1;
    double fr = (P[i0 - 1] - x) / (P[i0 - 1] - P[i0 - 2]);
    fr = 1.0 < fr ? 1.0 : fr;
    fr = 0.0 > fr ? 0.0 : fr;
    if (isx) {
        xi[base * nx + i] = ind;
        xa[base * nx + i] = fr;
    } else {
        yi[base * ny + i] = ind;
        ya[base * ny + i] = sc * fr;
        yb[base * ny + i] = sc * (1.0 - fr);
    }
}

// Running sums of the super-Gaussian x factors: one workgroup per (layer, slot, time)
// table; xc[0] = 0, xc[i+1] = xc[i] + xa[i].  Each thread sums a contiguous chunk, the
// chunk totals are scanned in LDS in a fixed order, then each chunk is written out.
__global__ void __launch_bounds__(256) k_src_cum(const LayerParams* __restrict__ lp, int nx,
                                                 const double* __restrict__ xa, double* __restrict__ xc)
{
    const int b = blockIdx.x;                 // (k * 2 + slot) * 2 + ts
    const int k = b >> 2, slot = (b >> 1) & 1;
    if (lp[k].src[slot].kind != MRMG_SRC_TEMP_SG) return;
    __shared__ double part[257];
    const int nt = (int)blockDim.x, t = (int)threadIdx.x;
    const int chunk = (nx + nt - 1) / nt;
    const int a0 = min(t * chunk, nx), a1 = min(a0 + chunk, nx);
    const double* x = xa + (size_t)b * nx;
    double* c = xc + (size_t)b * (nx + 1);
    double s = 0.0;
    for (int i = a0; i < a1; i++) s += x[i];
    part[t + 1] = s;
    __syncthreads();
    if (t == 0) {
        part[0] = 0.0;
        for (int i = 1; i <= nt; i++) part[i] += part[i - 1];
    }
    __syncthreads();
    s = part[t];
    if (t == 0) c[0] = 0.0;
    for (int i = a0; i < a1; i++) {
        s += x[i];
        c[i + 1] = s;
    }
}

// get_mag (data.f90:344-353): the value Ms*ms_scale*m that every cell of grain g takes,
// kept per grain in Mg[c][g] (in the demag's type); the row r2c spreads it over the grain's
// runs, and non-magnetic cells stay 0.
template <typename T, typename TC>
__device__ __forceinline__ void store_mg(TC* __restrict__ Mg, int G, int g, T f, const T* m)
{
    Mg[g] = (TC)(f * m[0]);
    Mg[G + g] = (TC)(f * m[1]);
    Mg[2 * G + g] = (TC)(f * m[2]);
}

// recompute: layers whose Ms follows a moving temperature take it at the axes' first time
template <typename T, int LPG, typename TC>
__global__ void __launch_bounds__(256) k_scatter(GrainData<T> gd, T* __restrict__ m, TC* __restrict__ Mg,
                                                 int recompute)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = tid / LPG, lane = tid % LPG;
    if (g >= gd.G) return;
    const LayerParams& L = gd.lp[gd.layer_of[g]];
    const T mm[3] = {m[g], m[gd.G + g], m[2 * gd.G + g]};
    T f = L.use_mag ? gd.ms[g] * m[(size_t)ST_MSC * gd.G + g] : gd.msf[g];
    if (recompute && L.src[1].kind != MRMG_SRC_NONE && scale_follows_t(L, 0)) {
        double tt[2];
        src_temp<T, LPG>(gd, L, g, lane, tt);
        const T msc = (T)scale_at_t(L, gd.tabs, 0, tt[0], gd.tc[g]);
        if (lane == 0) m[(size_t)ST_MSC * gd.G + g] = msc;
        f = gd.ms[g] * msc;
    }
    if (lane == 0) store_mg<T, TC>(Mg, gd.G, g, f, mm);
}

// ------------------------------------------------------------------------- LLB push
// Langevin function and its derivative (llg.f90:704-710, 726-733)
__device__ __forceinline__ double llb_lv(double x) { return fabs(x) > 1.e-2 ? 1.0 / tanh(x) - 1.0 / x : x / 3.0; }
__device__ __forceinline__ double llb_dlv(double x)
{
    if (fabs(x) > 4.e-3) {
        const double b = sinh(x);
        return 1.0 / (x * x) - 1.0 / (b * b);
    }
    return 1.0 / 3.0;
}
__device__ __forceinline__ double dot3d(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cross3d(const double* a, const double* b, double* t)
{
    t[0] = a[1] * b[2] - a[2] * b[1];
    t[1] = a[2] * b[0] - a[0] * b[2];
    t[2] = a[0] * b[1] - a[1] * b[0];
}
#define GMUB_KB 1.343427768e-4   // g muB / kB (llg.f90:925)

struct LlbGrain {
    double T, Tc, lam, spin, alpha, as, ms;
};

// push_llb (llg.f90:916-1065) for one grain, in FP64.  stage2: the corrector (Heun) or
// second half (mid-point) that starts from the predicted state mp/magp.  mi/magi: state
// at the start of the step; ph: field without the thermal part; th: the thermal draw.
// Heun's dm is read (stage 2) or written (stage 1) through dmv.  Returns the new
// direction and |m|, and the errmx/errmx1 contributions.
__device__ void llb_push(const LlbGrain& q, int type, bool stage2, double scale1, double dt0,
                         const double* mi, double magi, const double* mp, double magp, const double* ph,
                         const double* th, double* dmv, double* mo, double* mago, double* e, double* e1)
{
    const double T = q.T, Tc = q.Tc, lam = q.lam, alpha = q.alpha, spin = q.spin;
    double m[3], h[3], t[3], d[3], o[3], htot[3], mag, mhtot, meq, aperp, apar, y, scale11, scale2;
    *e = 0.0;
    if (magi > 0.0) {
        if (type > -3 && type < 3) {
            scale11 = scale1;
            if (!stage2) {   // first or only stage: relax |m| (llg.f90:937-947)
                for (int c = 0; c < 3; c++) m[c] = mi[c];
                mag = magi;
                for (int c = 0; c < 3; c++) htot[c] = ph[c] + lam * mag * m[c];
                mhtot = sqrt(dot3d(htot, htot));
                if (mhtot != 0.0) for (int c = 0; c < 3; c++) htot[c] = htot[c] / mhtot;
                meq = GMUB_KB * spin * mhtot / T;
                mag = llb_lv(meq) * dot3d(m, htot);
                mag = mag + (magi - mag) * exp(scale1 * 2.0 * alpha * lam * lam * T * T * llb_lv(meq) /
                                               (3.0 * Tc * Tc * llb_dlv(meq) * (mhtot + 1.e-20)));
            } else {
                for (int c = 0; c < 3; c++) m[c] = mp[c];
                mag = magp;
            }
            for (int c = 0; c < 3; c++) htot[c] = ph[c] + lam * mag * m[c];
            mhtot = sqrt(dot3d(htot, htot));
            if (mhtot != 0.0) for (int c = 0; c < 3; c++) htot[c] = htot[c] / mhtot;
            meq = GMUB_KB * spin * mhtot / T;
            aperp = alpha * (1.0 - lam * llb_lv(meq) * T / (3.0 * mhtot * Tc));
            apar = alpha * 2.0 * T / (3.0 * Tc);
            scale2 = -aperp / (fabs(mag) + 1.e-20);
            const double sth = sqrt(fabs(scale2 / (alpha * q.as * dt0 * (fabs(mag) + 1.e-20))));
            for (int c = 0; c < 3; c++) h[c] = ph[c] + th[c] * sth;
            for (int c = 0; c < 3; c++) m[c] = mi[c];
            cross3d(m, h, t);
            *e = dot3d(t, t) / (q.ms * q.ms);
            if (type != 2 && type != -2)
                for (int c = 0; c < 3; c++) o[c] = (1.0 + scale2 * scale2) * scale11 * h[c] + scale2 * m[c];
            else
                for (int c = 0; c < 3; c++) o[c] = scale11 * h[c] + scale2 * m[c];
            double mo3[3], to[3], mh[3];
            cross3d(m, o, mo3);
            for (int c = 0; c < 3; c++) t[c] = m[c] + mo3[c];
            y = 2.0 / (1.0 + dot3d(o, o));
            cross3d(t, o, to);
            for (int c = 0; c < 3; c++) d[c] = y * to[c];
            for (int c = 0; c < 3; c++) mo[c] = m[c] + d[c];
            for (int c = 0; c < 3; c++) mh[c] = m[c] + 0.5 * d[c];
            for (int c = 0; c < 3; c++) htot[c] = ph[c] + lam * mag * mh[c];
            mhtot = sqrt(dot3d(htot, htot));
            if (mhtot != 0.0) for (int c = 0; c < 3; c++) htot[c] = htot[c] / mhtot;
            meq = GMUB_KB * spin * mhtot / T;
            const double sg = signbit(mag) ? -1.0 : 1.0;
            double smh[3];
            for (int c = 0; c < 3; c++) smh[c] = sg * mh[c];
            mag = llb_lv(meq) * dot3d(smh, htot);
            meq = mag + (magi - mag) * exp(scale1 * 4.0 * alpha * lam * lam * T * T * llb_lv(meq) /
                                           (3.0 * Tc * Tc * llb_dlv(meq) * (mhtot + 1.e-20)));
            const double spar = sqrt(fabs(apar / (alpha * q.as * dt0)));
            double tp[3];
            for (int c = 0; c < 3; c++) tp[c] = th[c] * spar;
            meq = meq - 2.0 * scale11 * dot3d(tp, mh);
            *mago = fabs(meq);
            if (meq < 0.0) {
                for (int c = 0; c < 3; c++) o[c] = -mh[c];
                const double r = sqrt(dot3d(o, o));
                for (int c = 0; c < 3; c++) mo[c] = o[c] / r;
                for (int c = 0; c < 3; c++) d[c] = mo[c] - m[c];
            }
            y = 0.5 * (*mago + magi);
            const double dmg = *mago - magi;
            double ev[3];
            for (int c = 0; c < 3; c++) ev[c] = d[c] * y + m[c] * dmg;
            *e1 = dot3d(ev, ev);
        } else {   // Heun (llg.f90:992-1041)
            if (stage2) {
                for (int c = 0; c < 3; c++) m[c] = mp[c];
                mag = magp;
                if (mag <= 0.0) {
                    mag = 0.0;
                    y = sqrt(dot3d(ph, ph));
                    if (y != 0.0) for (int c = 0; c < 3; c++) m[c] = ph[c] / y;
                }
            } else {
                for (int c = 0; c < 3; c++) m[c] = mi[c];
                mag = magi;
            }
            for (int c = 0; c < 3; c++) htot[c] = ph[c] + lam * mag * m[c];
            mhtot = sqrt(dot3d(htot, htot));
            if (mhtot != 0.0) for (int c = 0; c < 3; c++) htot[c] = htot[c] / mhtot;
            meq = GMUB_KB * spin * mhtot / T;
            aperp = alpha * (1.0 - lam * llb_lv(meq) * T / (3.0 * mhtot * Tc));
            apar = alpha * 2.0 * T / (3.0 * Tc);
            const double sq = sqrt(alpha * q.as * dt0);
            double hpar[3], hperp[3];
            for (int c = 0; c < 3; c++) h[c] = th[c] / sq;
            const double hm = dot3d(h, m);
            for (int c = 0; c < 3; c++) { hpar[c] = hm * m[c]; hperp[c] = h[c] - hpar[c]; }
            const double sp = sqrt(apar), sr = sqrt((aperp - apar > 0.0 ? aperp - apar : 0.0) / (aperp * aperp));
            for (int c = 0; c < 3; c++) { hpar[c] = hpar[c] * sp; hperp[c] = hperp[c] * sr; }
            scale11 = 2.0 * scale1;
            meq = apar * (lam / mhtot) * lam * 2.0 * scale1 * T * llb_lv(meq) / (3.0 * llb_dlv(meq) * Tc) *
                  (mag - llb_lv(meq) * dot3d(m, htot));
            double hp[3], mxh[3], mxt[3];
            for (int c = 0; c < 3; c++) hp[c] = ph[c] + hperp[c];
            cross3d(m, hp, t);
            cross3d(m, ph, mxh);
            cross3d(m, t, mxt);
            const double am = fabs(meq) < 0.5 ? fabs(meq) : 0.5, rl = signbit(meq) ? -am : am;
            for (int c = 0; c < 3; c++) d[c] = scale11 * (mag * mxh[c] + aperp * mxt[c] + hpar[c]) + rl * m[c];
            if (stage2) for (int c = 0; c < 3; c++) d[c] = 0.5 * (d[c] + dmv[c]);
            else for (int c = 0; c < 3; c++) dmv[c] = d[c];
            *e = dot3d(t, t) / (q.ms * q.ms);
            *e1 = dot3d(d, d);
            for (int c = 0; c < 3; c++) o[c] = mi[c] * magi + d[c];
            y = sqrt(dot3d(o, o));
            if (y > 0.0) {
                for (int c = 0; c < 3; c++) mo[c] = o[c] / y;
            } else {
                y = 0.0;
                for (int c = 0; c < 3; c++) mo[c] = -mi[c];
            }
            *mago = y;
        }
    } else {   // |m| = 0 (llg.f90:1042-1056)
        y = sqrt(dot3d(ph, ph));
        if (y > 0.0) {
            for (int c = 0; c < 3; c++) mo[c] = ph[c] / y;
            scale11 = 2.0 * scale1;
            meq = GMUB_KB * spin * y / T;
            scale2 = lam * T / (3.0 * llb_dlv(meq) * Tc);
            *mago = (y / scale2 + llb_lv(meq)) * (1.0 - exp(scale11 * scale2 * 2.0 * T * alpha / (3.0 * Tc)));
        } else {
            for (int c = 0; c < 3; c++) mo[c] = mi[c];
            *mago = magi;
        }
        const double dmg = *mago - magi;
        *e1 = dmg * dmg;
    }
}

// Rotational pushes of push_llg (llg.f90:1086-1121): Boris (|type|==2) or analytic
// rotation with quaternions (|type|==1).  Returns d.
template <typename T>
__device__ __forceinline__ void rot_push(int type, T scale1, T scale2, const T* m, const T* H, T* d)
{
    if (type == 2 || type == -2) {
        T o[3], tt[3], c3[3];
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = scale1 * H[c] - scale2 * m[c];
        cross3(m, o, tt);
#pragma unroll
        for (int c = 0; c < 3; c++) tt[c] = m[c] + tt[c];
        const T y = (T)2 / ((T)1 + dot3(o, o));
        cross3(tt, o, c3);
#pragma unroll
        for (int c = 0; c < 3; c++) d[c] = y * c3[c];
    } else {
        T y = dot3(H, H);
        if (y > (T)0) {
            y = sqrt(y);
            const T scale11 = (T)2 * scale1 * y / ((T)1 + scale2 * scale2);
            T o[3], u0[3];
#pragma unroll
            for (int c = 0; c < 3; c++) o[c] = H[c] / y;
            T dp = dot3(o, m);
            dp = dp < (T)1 ? dp : (T)1;
            dp = dp > (T)-1 ? dp : (T)-1;
            const T th0 = acos(dp);
            cross3(m, o, u0);
            // rotate_2(theta0, u0 (normalised), theta1, u1 = o, m)  (transform.f90:97-115)
            T theta0 = th0 - (T)2 * atan(exp(scale11 * scale2) * tan((T)0.5 * th0));
            T theta1 = fabs(scale11);
            T st1 = (T)0.5 * theta1, st0 = (T)0.5 * theta0;
            T t0[3] = {u0[0], u0[1], u0[2]};
            T ct0 = sqrt(dot3(t0, t0));
            if (ct0 != (T)0) {
                t0[0] = t0[0] / ct0; t0[1] = t0[1] / ct0; t0[2] = t0[2] / ct0;
            } else {
                t0[0] = 0; t0[1] = 0; t0[2] = 1; st0 = 0;
            }
            if (st0 == (T)0 && st0 == st1) {
                d[0] = d[1] = d[2] = 0;
                return;
            }
            const T ct1 = cos(st1);
            st1 = sin(st1);
            ct0 = cos(st0);
            st0 = sin(st0);
            const T* t1 = o;
            T w[4];
            w[0] = ct0 * ct1 - dot3(t0, t1) * st0 * st1;
            w[1] = t0[0] * ct1 * st0 + (t1[0] * ct0 + (t1[1] * t0[2] - t0[1] * t1[2]) * st0) * st1;
            w[2] = t0[1] * ct1 * st0 + (t1[1] * ct0 + (t1[2] * t0[0] - t0[2] * t1[0]) * st0) * st1;
            w[3] = t0[2] * ct1 * st0 + (t1[2] * ct0 + (t1[0] * t0[1] - t0[0] * t1[1]) * st0) * st1;
            const T w1 = w[0] * w[0], w2 = w[1] * w[1], w3 = w[2] * w[2], w4 = w[3] * w[3];
            T r[3];
            r[0] = (T)2 * (m[2] * (w[0] * w[2] + w[1] * w[3]) + m[1] * (w[1] * w[2] - w[0] * w[3])) + m[0] * (w1 + w2 - w3 - w4);
            r[1] = (T)2 * (m[0] * (w[0] * w[3] + w[1] * w[2]) + m[2] * (w[2] * w[3] - w[0] * w[1])) + m[1] * (w1 + w3 - w2 - w4);
            r[2] = (T)2 * (m[1] * (w[0] * w[1] + w[2] * w[3]) + m[0] * (w[1] * w[3] - w[0] * w[2])) + m[2] * (w1 + w4 - w3 - w2);
#pragma unroll
            for (int c = 0; c < 3; c++) d[c] = r[c] - m[c];
        } else {
            d[0] = d[1] = d[2] = 0;
        }
    }
}


template <typename T, int MODE, int LPG, typename TC>
__global__ void __launch_bounds__(256) k_grain(GrainData<T> gd, T* __restrict__ ma, T* __restrict__ mb,
                                               T* __restrict__ dm, T* __restrict__ hthm, T* __restrict__ hout,
                                               const double* __restrict__ Hr, TC* __restrict__ Mg,
                                               StepArgs a, double* __restrict__ diag_part, T* __restrict__ mc)
{
    // grid padded to a multiple of 8: each XCD takes a contiguous range of grains, so the
    // cell lines neighbouring grains share stay in one L2
    const int nblk = (gd.G * LPG + 255) / 256;
    const int bl = xcd_remap(blockIdx.x, gridDim.x);
    if (bl >= nblk) return;
    const int tid = bl * blockDim.x + threadIdx.x;
    const int G = gd.G;
    // threads past the last grain repeat the last grain's work and store nothing, so the
    // loads below need no guard; every lane of a grain computes its results, lane 0 stores
    const bool act = tid / LPG < G;
    const int g = act ? tid / LPG : G - 1, lane = tid % LPG;
    const bool w0 = act && lane == 0;
    __shared__ double red[DIAG_STRIDE];
    __shared__ LayerParams slp[MRMG_MAX_LAYERS];
    // the grain's state and constants, issued first; the layer parameters' copy into LDS
    // then waits for all of them at once
    const int k = gd.layer_of[g];
    GrainPre<T> P;
    grain_pre(gd, g, P);
    const T* mf = (MODE == GM_HEUN2 || MODE == GM_ROT2) ? mb : ma;
    T mfield[3] = {mf[g], mf[G + g], mf[2 * G + g]};
    const T mi[3] = {ma[g], ma[G + g], ma[2 * G + g]};
    const T msc = mf[(size_t)ST_MSC * G + g];
    const T g_ms = gd.ms[g], g_msf = gd.msf[g], g_vol = gd.vol[g], g_sig = gd.sigma[g];
    T g_as = gd.alpha_scale[g];
    const double g_temp = gd.temp[g];
    T dmv[3] = {0, 0, 0}, thv[3] = {0, 0, 0};
    if (MODE == GM_HEUN2)
#pragma unroll
        for (int c = 0; c < 3; c++) dmv[c] = dm[(size_t)c * G + g];
    if (!(MODE == GM_HEUN1 || MODE == GM_ROT1 || MODE == GM_FIELD1))
#pragma unroll
        for (int c = 0; c < 3; c++) thv[c] = hthm[(size_t)c * G + g];
    {
        typedef unsigned long long u64;
        const int nw = gd.nl * (int)(sizeof(LayerParams) / sizeof(u64));
        const u64* src = reinterpret_cast<const u64*>(gd.lp);
        u64* dst = reinterpret_cast<u64*>(slp);
        for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
    }
    if (a.diag)
        for (int i = threadIdx.x; i < DIAG_STRIDE; i += blockDim.x) red[i] = 0.0;
    __syncthreads();
    const LayerParams& L = slp[k];
    T h[3] = {0, 0, 0};
    double tev[2] = {0, 0};   // grain temperature now and at the next evaluation (moving sources)
    if (MODE != GM_PUSH1) {
        field_at<T, LPG, TC>(gd, L, k, g, lane, mf, Hr, a.clk ? a.clk[0] + a.t : a.t, mfield, h, tev, P);
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) h[c] = hout[(size_t)c * G + g];
        if (L.src[1].kind != MRMG_SRC_NONE) src_temp<T, LPG>(gd, L, g, lane, tev);
    }
    T mo[3] = {0, 0, 0};
    T fsc = 0;   // Ms * ms_scale of the new state: the cell magnetisation factor
    {
        // the grain temperature of this evaluation and the next one
        const bool tsrc = L.src[1].kind != MRMG_SRC_NONE;
        double Tev = g_temp, Tnx = Tev;
        if (tsrc) {
            Tnx = tev[1];
            if (MODE == GM_PUSH1) {
                Tev = a.tlast[g];
            } else {
                Tev = tev[0];
                if (w0 && a.tlast && (MODE == GM_FIELD1 || MODE == GM_HEUN2)) a.tlast[g] = Tev;
            }
        }
        // the Ms scale of the new state at the next evaluation follows the temperature
        const bool dyn_ms = tsrc && scale_follows_t(L, 0);
        // alpha_scale at this evaluation's temperature (push and thermal variance)
        if (tsrc && scale_follows_t(L, 1)) g_as = (T)scale_at_t(L, gd.tabs, 1, Tev, gd.tc[g]);
        // thermal draw (get_thermal) happens with the first field of a step
        T th[3];
        if (MODE == GM_HEUN1 || MODE == GM_ROT1 || MODE == GM_FIELD1) {
            th[0] = th[1] = th[2] = 0;
            if (a.noise != 0 && L.thermal) {
                double n[3];
                if (a.noise == 1) {
                    n[0] = a.normals[3 * (size_t)g]; n[1] = a.normals[3 * (size_t)g + 1]; n[2] = a.normals[3 * (size_t)g + 2];
                } else {
                    philox_normals3(a.seed, a.clk ? (unsigned long long)a.clk[1] + a.step : a.step, g, n);
                }
                T sg = g_sig;
                if (tsrc)
                    sg = (T)thermal_sigma_dev(L.tsigma, L.floor_ms, (double)g_as, Tev, (double)g_ms, (double)g_vol,
                                              a.llb != 0, (double)ma[(size_t)ST_MAG * G + g],
                                              (double)ma[(size_t)ST_MSC * G + g]);
#pragma unroll
                for (int c = 0; c < 3; c++) th[c] = sg * (T)n[c];
            }
            if (w0) { hthm[g] = th[0]; hthm[G + g] = th[1]; hthm[2 * G + g] = th[2]; }
        } else {
            th[0] = thv[0]; th[1] = thv[1]; th[2] = thv[2];
        }
        if (w0 && MODE != GM_HEUN1 && MODE != GM_PUSH1) { hout[g] = h[0]; hout[G + g] = h[1]; hout[2 * G + g] = h[2]; }
        if (MODE == GM_FIELD1) {
            // get_fields, llg.f90:373-381: the step is chosen on the host from these maxima
            if (a.diag && w0) {
                const T di = (T)a.dti;
                T v[3];
#pragma unroll
                for (int c = 0; c < 3; c++) v[c] = h[c] + th[c] * di;
                const double q = (double)dot3(v, v);
                atomicMax(reinterpret_cast<unsigned long long*>(&red[D_HMAG + k]),
                          (unsigned long long)__double_as_longlong(q));
            }
        } else if (a.llb) {
            // push_llb in FP64 (llg.f90:916-1065)
            const bool stage2 = MODE == GM_HEUN2 || MODE == GM_ROT2;
            const bool heun = MODE == GM_HEUN1 || MODE == GM_HEUN2 || MODE == GM_PUSH1;
            LlbGrain q;
            q.T = Tev; q.Tc = gd.tc[g]; q.lam = gd.lam[g]; q.spin = L.spin; q.alpha = L.alpha;
            q.as = (double)g_as; q.ms = (double)g_ms;
            double mid[3], mpd[3], phd[3], thd[3], dmd[3] = {0, 0, 0}, mod[3], mago, e, e1;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                mid[c] = (double)ma[(size_t)c * G + g];
                mpd[c] = (double)mb[(size_t)c * G + g];
                phd[c] = (double)h[c];
                thd[c] = (double)th[c];
            }
            const double magi = (double)ma[(size_t)ST_MAG * G + g];
            const double magp = stage2 ? (double)mb[(size_t)ST_MAG * G + g] : 0.0;
            if (heun)
#pragma unroll
                for (int c = 0; c < 3; c++) dmd[c] = (double)dm[(size_t)c * G + g];
            llb_push(q, a.type, stage2, -0.5 * L.gamma * a.dts, a.dt, mid, magi, mpd, magp, phd, thd, dmd, mod, &mago,
                     &e, &e1);
            if (w0 && heun && !stage2)
#pragma unroll
                for (int c = 0; c < 3; c++) dm[(size_t)c * G + g] = (T)dmd[c];
#pragma unroll
            for (int c = 0; c < 3; c++) mo[c] = (T)mod[c];
            const T mscn = L.use_mag ? (T)(mago < 1.0 ? mago : 1.0)
                                     : dyn_ms ? (T)scale_at_t(L, gd.tabs, 0, Tnx, gd.tc[g]) : gd.ms_scale[g];
            fsc = (L.use_mag || dyn_ms) ? g_ms * mscn : g_msf;
            T* dst = (MODE == GM_HEUN1 || MODE == GM_ROT1 || MODE == GM_PUSH1) ? mb : (mc ? mc : ma);
            if (w0) {
                dst[g] = mo[0]; dst[G + g] = mo[1]; dst[2 * G + g] = mo[2];
                dst[(size_t)ST_MAG * G + g] = (T)mago;
                dst[(size_t)ST_MSC * G + g] = mscn;
            }
            if (a.diag && w0) {
                const double vol = (double)g_vol;
                atomicMax(reinterpret_cast<unsigned long long*>(&red[D_ERRMX + k]), (unsigned long long)__double_as_longlong(e));
                atomicMax(reinterpret_cast<unsigned long long*>(&red[D_ERRMX1 + k]), (unsigned long long)__double_as_longlong(e1));
                atomicAdd(&red[D_MMAG + k], mago * vol);
            }
        } else {
        const T sq = (T)a.sqrt_dt;
        T H[3];
#pragma unroll
        for (int c = 0; c < 3; c++) H[c] = th[c] / sq + h[c];
        const T mscn = dyn_ms ? (T)scale_at_t(L, gd.tabs, 0, Tnx, gd.tc[g]) : msc;
        fsc = (L.use_mag || dyn_ms) ? g_ms * mscn : g_msf;
        T d[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
        const T scale1 = (T)(-0.5 * L.gamma) * (T)a.dts;
        const T scale2 = g_as * (T)L.alpha;
        if (msc > (T)0) {
            if (MODE == GM_HEUN1 || MODE == GM_HEUN2 || MODE == GM_PUSH1) {
                const T scale11 = (T)2 * scale1 / ((T)1 + scale2 * scale2);
                const T* m = mfield;   // stage 1: m_in, stage 2: predicted m
                T mt[3];
                cross3(m, H, tq);
                cross3(m, tq, mt);
#pragma unroll
                for (int c = 0; c < 3; c++) d[c] = scale11 * (tq[c] + scale2 * mt[c]);
                if (MODE == GM_HEUN2) {
#pragma unroll
                    for (int c = 0; c < 3; c++) d[c] = (T)0.5 * (d[c] + dmv[c]);
                } else if (w0) {
#pragma unroll
                    for (int c = 0; c < 3; c++) dm[(size_t)c * G + g] = d[c];
                }
#pragma unroll
                for (int c = 0; c < 3; c++) mo[c] = mi[c] + d[c];
                const T y = dot3(mo, mo);
                if (y > (T)0) {
                    const T r = sqrt(y);
#pragma unroll
                    for (int c = 0; c < 3; c++) mo[c] = mo[c] / r;
                }
            } else {
                // rotational types push from m_in (llg.f90:1085)
                cross3(mi, H, tq);
                rot_push<T>(a.type, scale1, scale2, mi, H, d);
#pragma unroll
                for (int c = 0; c < 3; c++) mo[c] = mi[c] + d[c];
            }
        } else {
            // ms_scale = 0: m along the (non-thermal) field (llg.f90:1140-1147)
            const T y = sqrt(dot3(h, h));
            if (y > (T)0) {
#pragma unroll
                for (int c = 0; c < 3; c++) mo[c] = h[c] / y;
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) mo[c] = mi[c];
            }
        }
        T* dst = (MODE == GM_HEUN1 || MODE == GM_ROT1 || MODE == GM_PUSH1) ? mb : (mc ? mc : ma);
        if (w0) {
            dst[g] = mo[0]; dst[G + g] = mo[1]; dst[2 * G + g] = mo[2];
            if (dyn_ms) dst[(size_t)ST_MSC * G + g] = mscn;
        }
        if (a.diag && w0) {
            const T msv = g_ms;
            const double e = msc > (T)0 ? (double)dot3(tq, tq) / ((double)msv * (double)msv) : 0.0;
            const double e1 = msc > (T)0 ? (double)dot3(d, d) : 0.0;
            const double vol = (double)g_vol;
            atomicMax(reinterpret_cast<unsigned long long*>(&red[D_ERRMX + k]), (unsigned long long)__double_as_longlong(e));
            atomicMax(reinterpret_cast<unsigned long long*>(&red[D_ERRMX1 + k]), (unsigned long long)__double_as_longlong(e1));
            atomicAdd(&red[D_MMAG + k], sqrt((double)dot3(mo, mo)) * (double)msc * vol);
        }
        }
    }
    // the new magnetisation of the grain's cells, for the next field evaluation
    if (MODE != GM_FIELD1 && Mg && w0) store_mg<T, TC>(Mg, G, g, fsc, mo);
    if (a.diag) {
        __syncthreads();
        for (int i = threadIdx.x; i < DIAG_STRIDE; i += blockDim.x) diag_part[(size_t)bl * DIAG_STRIDE + i] = red[i];
    }
}

// llg()'s loop-top statistics (llg.f90:217-237) of layer blockIdx.x: the volume-weighted
// mean and variance of m and mean of h over the layer's grains, in two passes (mean, then
// squared deviations) with a fixed reduction order.  out[k*10 + ...] = W, <m>, var, <h>.
#define STATS_THREADS 1024
__device__ __forceinline__ double block_sum_fixed(double v, double* sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
    return r;
}
__global__ void __launch_bounds__(STATS_THREADS) k_layer_stats(const double* __restrict__ m,
                                                               const double* __restrict__ h,
                                                               const double* __restrict__ vol,
                                                               const LayerParams* __restrict__ lp, int G,
                                                               double* __restrict__ out)
{
    __shared__ double sh[STATS_THREADS / 64];
    const int k = blockIdx.x;
    const int g0 = lp[k].off, g1 = g0 + lp[k].ng;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int g = g0 + (int)threadIdx.x; g < g1; g += blockDim.x) {
        const double w = vol[g];
        s[0] += w;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            s[1 + c] += w * m[(size_t)c * G + g];
            s[4 + c] += w * h[(size_t)c * G + g];
        }
    }
    double r[7];
#pragma unroll
    for (int i = 0; i < 7; i++) r[i] = block_sum_fixed(s[i], sh);
    const double W = r[0];
    const double mean[3] = {r[1] / W, r[2] / W, r[3] / W};
    double v[3] = {0, 0, 0};
    for (int g = g0 + (int)threadIdx.x; g < g1; g += blockDim.x) {
        const double w = vol[g];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double d = m[(size_t)c * G + g] - mean[c];
            v[c] += w * d * d;
        }
    }
    double vr[3];
#pragma unroll
    for (int c = 0; c < 3; c++) vr[c] = block_sum_fixed(v[c], sh);
    if (threadIdx.x == 0) {
        double* o = out + (size_t)k * 10;
        o[0] = W;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o[1 + c] = mean[c];
            o[4 + c] = vr[c] / W;
            o[7 + c] = r[4 + c] / W;
        }
    }
}

// ------------------------------------------------------------------------- readback
// Reciprocity readback with a Gaussian sensitivity.  k_rb_plane folds the layers and the
// components of the cell magnetisation into one plane P(x,y) = sum_k wk[k] (dir . M); then
// one workgroup per head position sums P against the two Gaussians, whose values over the
// cell centres it first puts in LDS (periodic minimum-image distances).
struct RbWeights {
    double w[MRMG_MAX_LAYERS];
};
template <typename TC>
__global__ void __launch_bounds__(256) k_rb_plane(const TC* __restrict__ Mg, int G, const int* __restrict__ c2g,
                                                  int nxy, int nl, double d0, double d1, double d2, RbWeights wk,
                                                  double* __restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    double s = 0.0;
    for (int k = 0; k < nl; k++) {
        const int g = c2g[(size_t)k * nxy + i];
        if (g >= 0) s += wk.w[k] * (d0 * (double)Mg[g] + d1 * (double)Mg[G + g] + d2 * (double)Mg[2 * G + g]);
    }
    P[i] = s;
}
// periodic: minimum-image distances (the periodic media); else plain ones (open boundaries)
__global__ void __launch_bounds__(256) k_rb_scan(const double* __restrict__ P, int nx, int ny, double dx, double dy,
                                                 double sx, double sy, const double* __restrict__ pos, int n_pos,
                                                 double* __restrict__ out, int periodic)
{
    extern __shared__ double gsh[];   // nx + ny
    __shared__ double red[4];
    const int p = blockIdx.x;
    if (p >= n_pos) return;
    const double x0 = pos[p], y0 = pos[n_pos + p];
    const double Lx = nx * dx, Ly = ny * dy;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        double d = (i + 0.5) * dx - x0;
        if (periodic) d -= Lx * rint(d / Lx);
        gsh[i] = exp(-d * d / (2.0 * sx * sx));
    }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        double d = (j + 0.5) * dy - y0;
        if (periodic) d -= Ly * rint(d / Ly);
        gsh[nx + j] = exp(-d * d / (2.0 * sy * sy));
    }
    __syncthreads();
    double s = 0.0;
    for (int idx = threadIdx.x; idx < nx * ny; idx += blockDim.x) {
        const int j = idx / nx, i = idx - j * nx;
        s += P[idx] * gsh[i] * gsh[nx + j];
    }
    const double r = block_sum_fixed(s, red);
    if (threadIdx.x == 0) out[p] = r;
}


// ------------------------------------------------------------------------- switching record
// After a completed step: a grain not yet marked whose m satisfies dir . m > thr gets the
// step's index, its end time and the applied field there (the grain's static field plus
// the uniform ramped field at that time, as get_applied sums them; a moving table source
// is not included).
template <typename S>
__global__ void __launch_bounds__(256) k_switch(GrainData<S> gd, const S* __restrict__ m, double d0, double d1,
                                                double d2, double thr, long long step, double t,
                                                long long* __restrict__ rs, double* __restrict__ rt,
                                                double* __restrict__ rh)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int G = gd.G;
    if (g >= G || rs[g] >= 0) return;
    const double v = d0 * (double)m[g] + d1 * (double)m[(size_t)G + g] + d2 * (double)m[2 * (size_t)G + g];
    if (v > thr) {
        const LayerParams& L = gd.lp[gd.layer_of[g]];
        double h[3] = {0, 0, 0};
        if (gd.happ)
            for (int i = 0; i < 3; i++) h[i] = (double)gd.happ[(size_t)i * G + g];
        if (L.has_uf) {
            const double sc = tscale_eval(L.uf_fn, L.uf_t0, L.uf_t1, L.uf_s0, L.uf_s1, L.uf_scale, t);
            for (int i = 0; i < 3; i++) h[i] += sc * L.uf_dir[i];
        }
        rs[g] = step;
        rt[g] = t;
        for (int i = 0; i < 3; i++) rh[(size_t)i * G + g] = h[i];
    }
}
__global__ void k_fill_i64(long long* __restrict__ a, size_t n, long long v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// streaming copy (16-byte vectors), the known-size reference of the traffic counters
__global__ void k_copy16(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// the axis tables of the moving sources at the evaluation time t0 and the next t1
template <typename T>
int Engine<T>::src_axes(double t0, double t1)
{
    if (!any_src) return MRMG_OK;
    const int n = c->nl * 4 * (c->nx + c->ny);
    hipLaunchKernelGGL(k_src_axes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, lp.p, c->nl, c->nx,
                       c->ny, c->dx, c->dy, t0, t1, ax_xi.p, ax_xa.p, ax_yi.p, ax_ya.p, ax_yb.p, src_pos.p);
    HIPCHK(hipGetLastError());
    if (any_sg) {
        hipLaunchKernelGGL(k_src_cum, dim3((unsigned)(c->nl * 4)), dim3(256), 0, c->stream, lp.p, c->nx, ax_xa.p,
                           ax_xc.p);
        HIPCHK(hipGetLastError());
    }
    return MRMG_OK;
}

// get_mag for the state m (used where no grain kernel has written the grains' cell values
// yet); with t0 given, Ms scales that follow a moving temperature are taken at t0
template <typename T>
int Engine<T>::scatter(S* m, const double* t0 )
{
    const size_t threads = (size_t)G * 8;
    if (t0) RCHK(src_axes(*t0, *t0));
    hipLaunchKernelGGL((k_scatter<S, 8, T>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream,
                       gdata(), m, Mg.p, t0 && any_src ? 1 : 0);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

template <typename T>
template <int MODE, int LPG> void Engine<T>::launch_grain(const StepArgs& a, S* mout)
{
    const size_t threads = (size_t)G * LPG;
    const unsigned nb8 = (unsigned)(((threads + 255) / 256 + 7) / 8 * 8);
    hipLaunchKernelGGL((k_grain<S, MODE, LPG, T>), dim3(nb8), dim3(256), 0, c->stream,
                       gdata(), ma.p, mb.p, dm.p, hthm.p, h.p, demag_on() ? Hr.p : nullptr,
                       demag_on() ? Mg.p : nullptr, a, diag_part.p, mout);
}

// mout: destination of the corrector (null: ma); t_next: time of the next field
// evaluation, for the moving sources
template <typename T>
template <int MODE> int Engine<T>::grain(const StepArgs& a, S* mout, double t_next )
{
    RCHK(src_axes(a.t, t_next));
    RCHK(ev_begin((MODE == GM_HEUN1 || MODE == GM_ROT1 || MODE == GM_FIELD1 || MODE == GM_PUSH1) ? K_GRAIN1
                                                                                             : K_GRAIN2));
    if (lpg == 4) launch_grain<MODE, 4>(a, mout);
    else if (lpg == 2) launch_grain<MODE, 2>(a, mout);
    else if (lpg == 1) launch_grain<MODE, 1>(a, mout);
    else if (lpg == 16) launch_grain<MODE, 16>(a, mout);
    else launch_grain<MODE, 8>(a, mout);
    HIPCHK(hipGetLastError());
    RCHK(ev_end());
    return MRMG_OK;
}

// llg()'s loop-top statistics of the current state and h, into stats
template <typename T>
int Engine<T>::layer_stats()
{
    hipLaunchKernelGGL(k_layer_stats, dim3(c->nl), dim3(STATS_THREADS), 0, c->stream, ma.p, h.p, vol.p, lp.p, G,
                       stats.p);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

// the readback with caller-provided device buffers: plane (nx*ny), pos (2*n_pos: x
// then y), dout (n_pos); every copy is ordered on the context's stream
template <typename T>
int Engine<T>::readback_into(const mrmg_sensor* sn, int n_pos, const double* xpos, const double* ypos, double* out, double* plane, double* pos, double* dout)
{
    const int nx = c->nx, ny = c->ny, nl = c->nl, nxy = nx * ny;
    RbWeights wk{};
    for (int k = 0; k < nl; k++) wk.w[k] = (sn->w ? sn->w[k] : 1.0) * c->dz[k] * c->dx * c->dy;
    HIPCHK(hipMemcpyAsync(pos, xpos, sizeof(double) * n_pos, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(pos + n_pos, ypos, sizeof(double) * n_pos, hipMemcpyHostToDevice, c->stream));
    RCHK(scatter(ma.p));   // the grains' cell values of the current state
    hipLaunchKernelGGL((k_rb_plane<T>), dim3((nxy + 255) / 256), dim3(256), 0, c->stream, Mg.p, G, cell2grain.p, nxy,
                       nl, sn->dir[0], sn->dir[1], sn->dir[2], wk, plane);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_rb_scan, dim3(n_pos), dim3(256), sizeof(double) * (nx + ny), c->stream, plane, nx, ny,
                       c->dx, c->dy, sn->sigma_x, sn->sigma_y, pos, n_pos, dout, c->open ? 0 : 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * n_pos, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MRMG_OK;
}

template <typename T>
int Engine<T>::switch_arm(int enable, const double* dir, double thr)
{
    if (!enable) {
        sw_on = false;
        return MRMG_OK;
    }
    if (sw_step.n < (size_t)G) {
        RCHK(sw_step.alloc(G));
        RCHK(sw_t.alloc(G));
        RCHK(sw_h.alloc(3 * (size_t)G));
    }
    hipLaunchKernelGGL(k_fill_i64, dim3((G + 255) / 256), dim3(256), 0, c->stream, sw_step.p, (size_t)G, -1LL);
    HIPCHK(hipGetLastError());
    for (int i = 0; i < 3; i++) sw_dir[i] = dir[i];
    sw_thr = thr;
    sw_on = true;
    return MRMG_OK;
}

// after a completed step: first-time marks of the grains past the threshold
template <typename T>
int Engine<T>::switch_mark(long long step, double t)
{
    hipLaunchKernelGGL((k_switch<S>), dim3((G + 255) / 256), dim3(256), 0, c->stream, gdata(), ma.p, sw_dir[0],
                       sw_dir[1], sw_dir[2], sw_thr, step, t, sw_step.p, sw_t.p, sw_h.p);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}


int launch_philox_dump(hipStream_t s, unsigned long long seed, unsigned long long step0, int nsteps, int G,
                       double* out)
{
    const size_t cnt = (size_t)nsteps * G;
    hipLaunchKernelGGL(k_philox_dump, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, seed, step0, nsteps, G,
                       out);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

// the device clocks [t, step] of n batch members after a completed step: t = t + dt (the
// host loop's own addition), step + 1
__global__ void k_clock_advance(double* __restrict__ clk, int n, double dt)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    clk[2 * b] = clk[2 * b] + dt;
    clk[2 * b + 1] = clk[2 * b + 1] + 1.0;
}

int launch_clock_advance(hipStream_t s, double* clk, int n, double dt)
{
    hipLaunchKernelGGL(k_clock_advance, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, clk, n, dt);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

int launch_copy16(hipStream_t s, unsigned grid, const uint4* a, uint4* b, size_t n)
{
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, a, b, n);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

#define MRMG_INST(T)                                                                                       \
    template int Engine<T>::src_axes(double, double);                                                      \
    template int Engine<T>::scatter(double*, const double*);                                               \
    template int Engine<T>::layer_stats();                                                                 \
    template int Engine<T>::readback_into(const mrmg_sensor*, int, const double*, const double*, double*,   \
                                          double*, double*, double*);                                       \
    template int Engine<T>::switch_arm(int, const double*, double);                                        \
    template int Engine<T>::switch_mark(long long, double);                                                \
    template int Engine<T>::grain<GM_HEUN1>(const StepArgs&, double*, double);                             \
    template int Engine<T>::grain<GM_HEUN2>(const StepArgs&, double*, double);                             \
    template int Engine<T>::grain<GM_ROT1>(const StepArgs&, double*, double);                              \
    template int Engine<T>::grain<GM_ROT2>(const StepArgs&, double*, double);                              \
    template int Engine<T>::grain<GM_FIELD1>(const StepArgs&, double*, double);                            \
    template int Engine<T>::grain<GM_PUSH1>(const StepArgs&, double*, double);
MRMG_INST(float)
MRMG_INST(double)
