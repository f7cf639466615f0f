// demag.hip - the demag FFT kernels (get_mag scatter into the row r2c, column FFTs, the
// k-space product of demag.f90:117-121, row c2r) and the Engine members that launch them.
#include "engine.h"
#include "device_common.h"

#ifndef MRMG_FORM
#error "compile with -DMRMG_FORM=float or -DMRMG_FORM=double"
#endif

// ------------------------------------------------------------------------- runs <-> LDS rows
// A thread takes the runs q0 + tid, q0 + tid + nt, ... of its block, RU at a time: the
// descriptors (and grain values) of RU runs are loaded before any is used, so the
// dependent loads of a run (descriptor, then grain value) are waited for once per RU runs
// instead of once per run.  RU is 8 for the one-component row kernels on long rows (E >= 16
// FFT elements per thread, where a thread has several runs) and 4 otherwise: the c2r holds
// its prefetched descriptors across the FFT, and where they are not used (short rows, or
// the three rows of the fused kernels) they cost occupancy.
#define MRMG_RU(NC, E) (((NC) == 1 && (E) >= 16) ? 8 : 4)
#define MRMG_UF 4

// The cell magnetisation of rows y0 .. y0+nrows-1 of layer k into LDS, component c0+ci of
// NC at row pair p = (y - y0)/2: re holds the even rows, im the odd ones, at element
// (ci*np + p)*nx + x (get_mag, data.f90:344-353: Ms ms_scale m on a grain's cells, 0 on
// non-magnetic cells).  Mg[c][g] is the grain's cell value.  Also stages the row
// twiddles.  roff (shared, nrows+1 entries) keeps the block's run offsets for the store.
// The first RU runs' loads are issued before LDS is cleared and overlap it.  im must be
// re + LPN(np * NC * nx), as every caller lays them out.
template <typename T, int NC, int RU>
__device__ __forceinline__ void runs_to_lds(const RowRuns& R, const T* __restrict__ Mg, int G, int c0, int k, int ny,
                                            int y0, int nrows, int np, int nx, T* re, T* im, int* roff, C2<T>* tw,
                                            const C2<T>* __restrict__ twx)
{
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    const int* ro = R.row_off + k * ny + y0;          // y0 + nrows <= ny
    const int q0 = ro[0], q1 = ro[nrows];
    int geo[RU];
    T v[RU][NC];
    auto fetch = [&](int qb) {
        int g[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int q = qb + u * nt;
            if (q < q1) { geo[u] = R.geo[q]; g[u] = R.grain[q]; }
        }
#pragma unroll
        for (int u = 0; u < RU; u++)
            if (qb + u * nt < q1)
#pragma unroll
                for (int ci = 0; ci < NC; ci++) v[u][ci] = Mg[(size_t)(c0 + ci) * G + g[u]];
    };
    fetch(q0 + tid);
    if (tid <= nrows) roff[tid] = ro[tid];
    if (tw) for (int i = tid; i < nx; i += nt) tw[i] = twx[i];   // null: no LDS copy wanted
    const int n = NC * np * nx;
    if ((reinterpret_cast<size_t>(re) & 15) == 0) {
        // re and im (= re + LPN(n)) are one span of 2 LPN(n) elements: cleared whole,
        // padding slots included, with 16-byte stores and a scalar tail
        const int tot = 2 * LPN(n), per = 16 / (int)sizeof(T), n16 = tot / per;
        uint4* z = reinterpret_cast<uint4*>(re);
        for (int i = tid; i < n16; i += nt) z[i] = uint4{0u, 0u, 0u, 0u};
        for (int i = n16 * per + tid; i < tot; i += nt) re[i] = (T)0;
    } else {
        for (int i = tid; i < n; i += nt) {
            re[LP(i)] = (T)0;
            im[LP(i)] = (T)0;
        }
    }
    __syncthreads();
    // a thread's runs come in increasing q, so its row index yy only moves forward
    int yy = 0;
    for (int qb = q0 + tid; qb < q1; qb += RU * nt) {
        if (qb != q0 + tid) fetch(qb);
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int q = qb + u * nt;
            if (q >= q1) break;
            while (yy + 1 < nrows && q >= roff[yy + 1]) yy++;
            const int i0 = geo[u] & 0xffff, len = geo[u] >> 16;
            T* dst = (yy & 1) ? im : re;
#pragma unroll
            for (int ci = 0; ci < NC; ci++) {
                const int b = (ci * np + (yy >> 1)) * nx + i0;
                // MRMG_UF cells per iteration; past the run's end the last cell is written
                // again (the same value), so no write is predicated
                for (int t = 0; t < len; t += MRMG_UF)
#pragma unroll
                    for (int j = 0; j < MRMG_UF; j++) dst[LP(b + min(t + j, len - 1))] = v[u][ci];
            }
        }
    }
}

// The run descriptors a thread sums in lds_to_runs, its first RU of them loaded ahead
// (at kernel start, before the spectra and the FFT) by prefetch.
template <int RU>
struct RunsAhead {
    int q0, q1;
    int geo[RU], gm[RU];
    __device__ __forceinline__ void fetch(const RowRuns& R, int qb, int nt)
    {
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int q = qb + u * nt;
            if (q < q1) { geo[u] = R.geo[q]; gm[u] = R.gm[q]; }
        }
    }
    __device__ __forceinline__ void prefetch(const RowRuns& R, int k, int ny, int y0, int nrows)
    {
        const int* ro = R.row_off + k * ny + y0;
        q0 = ro[0];
        q1 = ro[nrows];
        fetch(R, q0 + (int)threadIdx.x, (int)blockDim.x);
    }
};

// The cell field of the same rows, summed over each run (in x order, in FP64) into
// Hr[c][run] (c = c0+ci), with the layout of runs_to_lds.
template <typename T, int NC, int RU>
__device__ __forceinline__ void lds_to_runs(const RowRuns& R, RunsAhead<RU>& A, double* __restrict__ Hr, int c0,
                                            int nrows, int np, int nx, const T* re, const T* im, const int* roff)
{
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    int yy = 0;   // increasing q: the row index only moves forward
    for (int qb = A.q0 + tid; qb < A.q1; qb += RU * nt) {
        if (qb != A.q0 + tid) A.fetch(R, qb, nt);
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int q = qb + u * nt;
            if (q >= A.q1) break;
            while (yy + 1 < nrows && q >= roff[yy + 1]) yy++;
            const int i0 = A.geo[u] & 0xffff, len = A.geo[u] >> 16;
            const T* src = (yy & 1) ? im : re;
#pragma unroll
            for (int ci = 0; ci < NC; ci++) {
                const int b = (ci * np + (yy >> 1)) * nx + i0;
                // in x order, MRMG_UF reads issued per iteration; past the run's end the
                // last cell is read again and 0 added (exact: s is never -0)
                double s = 0.0;
                for (int t = 0; t < len; t += MRMG_UF) {
                    T a[MRMG_UF];
#pragma unroll
                    for (int j = 0; j < MRMG_UF; j++) a[j] = src[LP(b + min(t + j, len - 1))];
#pragma unroll
                    for (int j = 0; j < MRMG_UF; j++) s += (double)(t + j < len ? a[j] : (T)0);
                }
                Hr[(size_t)(c0 + ci) * R.nruns + A.gm[u]] = s;
            }
        }
    }
}

// ------------------------------------------------------------------------- demag kernels
// Row r2c: one workgroup per (row pair, layer).  Two rows of each component of the cell
// magnetisation (written by the grain kernels, get_mag data.f90:344-353) are packed as
// z = M(y0) + i M(y1), transformed with one complex FFT and separated.
// nx: the FFT row length; ny: the media rows of a layer; sy: the spectra's rows (the FFT
// column length, 2 ny with open boundaries, whose padding rows the kernel leaves alone).
// BAT: one launch over a batch of contexts, blockIdx.z the member (operands from tab);
// else the context's own operands (one).  The same holds for every demag kernel below.
template <typename T, int E, bool BAT = false>
__global__ void __launch_bounds__(256) k_row_fwd(DemagOps<T> one, const DemagOps<T>* __restrict__ tab,
                                                 int nx, int ny, int sy, int nl, int nxh, FFTPlan px,
                                                 const C2<T>* __restrict__ twx)
{
    const DemagOps<T> o = BAT ? tab[blockIdx.z] : one;
    const T* __restrict__ Mg = o.Mg;
    const int G = o.G;
    const RowRuns R = o.R;
    C2<T>* __restrict__ S = o.S1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int roff[3];
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    T* re = reinterpret_cast<T*>(tw + nx);
    T* im = re + LPN(3 * nx);
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1, k = blockIdx.y;
    const bool has1 = y1 < ny;
    runs_to_lds<T, 3, MRMG_RU(3, E)>(R, Mg, G, 0, k, ny, y0, has1 ? 2 : 1, 1, nx, re, im, roff, tw, twx);
    __syncthreads();
    lds_fft<T, E, false, true>(re, im, 3, px, tw);
    for (int idx = threadIdx.x; idx < 3 * nxh; idx += blockDim.x) {
        const int c = idx / nxh, kx = idx - c * nxh;
        const int kn = kx == 0 ? 0 : nx - kx;
        const T zr = re[LP(c * nx + kx)], zi = im[LP(c * nx + kx)], nr = re[LP(c * nx + kn)], ni = im[LP(c * nx + kn)];
        C2<T> A = {(zr + nr) * (T)0.5, (zi - ni) * (T)0.5};
        C2<T> B = {(zi + ni) * (T)0.5, (nr - zr) * (T)0.5};
        C2<T>* row = S + (((size_t)k * 3 + c) * sy + y0) * nxh;
        row[kx] = A;
        if (has1) row[nxh + kx] = B;
    }
}


// Row c2r: FFTW's 1D c2r semantics (imaginary parts of the DC and Nyquist bins ignored).
// CELLS: the cell field into Hcell [c][k][y][x] (mrmg_demag_field; mx: the media row
// length), else its run sums into Hr.
template <typename T, int E, bool CELLS = false, bool BAT = false>
__global__ void __launch_bounds__(256) k_row_inv(DemagOps<T> one, const DemagOps<T>* __restrict__ tab,
                                                 int nx, int ny, int sy, int mx, int nl, int nxh, FFTPlan px,
                                                 const C2<T>* __restrict__ twx)
{
    const DemagOps<T> o = BAT ? tab[blockIdx.z] : one;
    const C2<T>* __restrict__ S2 = o.S2;
    T* __restrict__ Hcell = o.Hcell;
    double* __restrict__ Hr = o.Hr;
    const RowRuns R = o.R;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int roff[3];
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    T* re = reinterpret_cast<T*>(tw + nx);
    T* im = re + LPN(3 * nx);
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1, k = blockIdx.y;
    const bool has1 = y1 < ny;
    RunsAhead<MRMG_RU(3, E)> ra;
    if (!CELLS) ra.prefetch(R, k, ny, y0, has1 ? 2 : 1);
    staged_load<(E < 4 ? E : 4)>(
        3 * nxh,
        [&](int idx) {
            const int c = idx / nxh, kx = idx - c * nxh;
            const C2<T>* row = S2 + (((size_t)k * 3 + c) * sy + y0) * nxh;
            return C4<T>{row[kx], row[(has1 ? nxh : 0) + kx]};
        },
        [&](int idx, C4<T> v) {
            const int c = idx / nxh, kx = idx - c * nxh;
            C2<T> A = v.a, B = has1 ? v.b : C2<T>{0, 0};
            if (kx == 0 || 2 * kx == nx) { A.y = 0; B.y = 0; }
            re[LP(c * nx + kx)] = A.x - B.y;
            im[LP(c * nx + kx)] = A.y + B.x;
            if (kx > 0 && nx - kx != kx) {
                re[LP(c * nx + nx - kx)] = A.x + B.y;
                im[LP(c * nx + nx - kx)] = B.x - A.y;
            }
        },
        tw, twx, nx);
    __syncthreads();
    if (!CELLS && threadIdx.x <= (has1 ? 2 : 1)) roff[threadIdx.x] = R.row_off[k * ny + min(y0 + (int)threadIdx.x, ny)];
    lds_fft<T, E, true, true>(re, im, 3, px, tw);
    if (!CELLS) {
        lds_to_runs<T, 3, MRMG_RU(3, E)>(R, ra, Hr, 0, has1 ? 2 : 1, 1, nx, re, im, roff);
        return;
    }
    const size_t plane = (size_t)mx * ny;
    for (int idx = threadIdx.x; idx < 3 * mx; idx += blockDim.x) {
        const int c = idx / mx, x = idx - c * mx;
        T* hc = Hcell + ((size_t)c * nl + k) * plane;
        hc[(size_t)y0 * mx + x] = re[LP(c * nx + x)];
        if (has1) hc[(size_t)y1 * mx + x] = im[LP(c * nx + x)];
    }
}

// ------------------------------------------------------------------------- split column path
// For grids whose 3*n_lay columns of one kx do not fit one workgroup's LDS, the demag
// runs as five kernels over column-contiguous spectra ST[k][c][kx][y]:
//   k_rowT_fwd  r2c of TY rows of one (component, layer), written as TY-long y runs
//   k_colT      y-FFT of whole columns (CB columns per workgroup), in place
//   k_prod      the k-space product (demag.f90:117-121), one thread per (kx, ky)
//   k_colT      inverse y-FFT, in place
//   k_rowT_inv  c2r of TY rows, read back from TY-long y runs
template <typename T, int E, bool WV = false, bool BAT = false>
__global__ void __launch_bounds__(256) k_rowT_fwd(DemagOps<T> one, const DemagOps<T>* __restrict__ tab,
                                                  int nx, int ny, int sy, int nl, int nxh, int TY, FFTPlan px,
                                                  const C2<T>* __restrict__ twx)
{
    const DemagOps<T> o = BAT ? tab[blockIdx.z] : one;
    const T* __restrict__ Mg = o.Mg;
    const int G = o.G;
    const RowRuns R = o.R;
    C2<T>* __restrict__ S = o.S1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int roff[9];
    // the twiddles are read from the plan's table in global memory (L2): its per-pass
    // layout makes the reads coalesced, and the LDS it would take (8 nx bytes) goes to
    // more workgroups per CU (measured at 2048 points: 253 against 307 us)
    const C2<T>* __restrict__ tw = twx;
    const int B = TY / 2;
    T* re = reinterpret_cast<T*>(smem_raw);
    T* im = re + LPN(B * nx);
    // 1D grid padded to a multiple of 8, XCD-contiguous: the row blocks that share the
    // cache lines of one kx run (TY*sizeof(C2<T>) < 128 bytes) are read/written in one L2
    const int nyb = (ny + TY - 1) / TY;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if (L >= nyb * 3 * nl) return;
    const int y0 = (L % nyb) * TY, kc = L / nyb;        // kc = k*3 + c
    const int k = kc / 3, c = kc - 3 * k;
    runs_to_lds<T, 1, MRMG_RU(1, E)>(R, Mg, G, c, k, ny, y0, min(TY, ny - y0), B, nx, re, im, roff, (C2<T>*)nullptr, twx);
    __syncthreads();
    if (WV) {
        lds_fft_wave<T, E, false>(re, im, B, px, tw);
        __syncthreads();
    } else {
        lds_fft<T, E, false, true>(re, im, B, px, tw);
    }
    // one thread per (kx, row pair): both rows of the pair come from the same four LDS
    // values, and go out as one 2-element store where the pair is aligned (sy even).
    // B = TY/2 is 1, 2 or 4: kx and p are a shift and a mask of idx
    const int lb = B == 4 ? 2 : B == 2 ? 1 : 0;
    C2<T>* dst = S + (size_t)kc * nxh * sy;
    for (int idx = threadIdx.x; idx < nxh * B; idx += blockDim.x) {
        const int kx = idx >> lb, p = idx & (B - 1);
        const int y = y0 + 2 * p;
        if (y >= ny) continue;
        const int kn = kx == 0 ? 0 : nx - kx;
        const T zr = re[LP(p * nx + kx)], zi = im[LP(p * nx + kx)], nr = re[LP(p * nx + kn)], ni = im[LP(p * nx + kn)];
        const C2<T> a = {(zr + nr) * (T)0.5, (zi - ni) * (T)0.5};
        const C2<T> b = {(zi + ni) * (T)0.5, (nr - zr) * (T)0.5};
        C2<T>* d = dst + (size_t)kx * sy + y;
        if (y + 1 < ny && (sy & 1) == 0) {
            *reinterpret_cast<C4<T>*>(d) = C4<T>{a, b};
        } else {
            d[0] = a;
            if (y + 1 < ny) d[1] = b;
        }
    }
}

// y-FFT of ncols contiguous columns of length ny, CB columns per workgroup, in place.
// Twiddles are read from the global table (cached), not staged.  Rows y >= nyv (the
// padding of open boundaries) are read as zero by the forward transform and not stored by
// the inverse one (no row kernel reads them).
template <typename T, int E, bool INV, bool BAT = false>
__global__ void __launch_bounds__(256) k_colT(DemagOps<T> one, const DemagOps<T>* __restrict__ tab, int ny, int nyv,
                                              int ncols, int CB, FFTPlan py, const C2<T>* __restrict__ twy)
{
    C2<T>* __restrict__ S = INV ? (BAT ? tab[blockIdx.z].S2 : one.S2) : (BAT ? tab[blockIdx.z].S1 : one.S1);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* re = reinterpret_cast<T*>(smem_raw);
    const int q0 = blockIdx.x * CB;
    const int nb = min(CB, ncols - q0);
    T* im = re + LPN(CB * ny);
    C2<T>* base = S + (size_t)q0 * ny;
    if (INV || nyv == ny)
        staged_load<(E < 8 ? E : 8), T>(
            nb * ny, [&](int idx) { return base[idx]; }, [&](int idx, C2<T> v) { re[LP(idx)] = v.x; im[LP(idx)] = v.y; });
    else
        staged_load<(E < 8 ? E : 8), T>(
            nb * ny, [&](int idx) { return idx % ny < nyv ? base[idx] : C2<T>{0, 0}; },
            [&](int idx, C2<T> v) { re[LP(idx)] = v.x; im[LP(idx)] = v.y; });
    __syncthreads();
    lds_fft<T, E, INV, true>(re, im, nb, py, twy);
    for (int idx = threadIdx.x; idx < nb * ny; idx += blockDim.x)
        if (!INV || nyv == ny || idx % ny < nyv) base[idx] = C2<T>{re[LP(idx)], im[LP(idx)]};
}

// The k-space product over column-contiguous spectra: thread per (kx, ky), observation
// layers in chunks of OC, each output summed over (ks, b) in the reference's order.
// NT: the tensor and the field spectra as non-temporal streams (see prod_streams).
template <typename T, bool SYM, int OC, bool BAT = false, bool NT = false>
__global__ void __launch_bounds__(256) k_prod(DemagOps<T> one, const DemagOps<T>* __restrict__ tab, int ny, int nl,
                                              int nxh, int nslot, int toep)
{
    const C2<T>* __restrict__ S1 = BAT ? tab[blockIdx.z].S1 : one.S1;
    C2<T>* __restrict__ S2 = BAT ? tab[blockIdx.z].S2 : one.S2;
    const C2<T>* __restrict__ Tspec = BAT ? tab[blockIdx.z].Tspec : one.Tspec;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t Kc = (size_t)nxh * ny;
    if (idx >= Kc) return;
    const int kx = (int)(idx / ny), ky = (int)(idx - (size_t)kx * ny);
    for (int kc0 = 0; kc0 < nl; kc0 += OC) {
        C2<T> h[OC][3];
#pragma unroll
        for (int j = 0; j < OC; j++)
#pragma unroll
            for (int a = 0; a < 3; a++) h[j][a] = {0, 0};
        for (int ks = 0; ks < nl; ks++) {
            const C2<T> mx = S1[((size_t)(3 * ks + 0)) * Kc + idx];
            const C2<T> my = S1[((size_t)(3 * ks + 1)) * Kc + idx];
            const C2<T> mz = S1[((size_t)(3 * ks + 2)) * Kc + idx];
#pragma unroll
            for (int j = 0; j < OC; j++) {
                const int ko = kc0 + j;
                if (ko >= nl) break;
                const C2<T>* t = Tspec + tensor_block(kx, ko, ks, nl, toep) * nslot * ny + ky;
                C2<T> a;
                if (SYM) {
                    const C2<T> txx = ld_s<NT>(t), tyy = ld_s<NT>(t + ny), tzz = ld_s<NT>(t + 2 * ny),
                                txy = ld_s<NT>(t + 3 * ny), txz = ld_s<NT>(t + 4 * ny),
                                tyz = ld_s<NT>(t + 5 * ny);
                    a = cmul(txx, mx); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(txy, my); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(txz, mz); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(txy, mx); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(tyy, my); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(tyz, mz); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(txz, mx); h[j][2].x += a.x; h[j][2].y += a.y;
                    a = cmul(tyz, my); h[j][2].x += a.x; h[j][2].y += a.y;
                    a = cmul(tzz, mz); h[j][2].x += a.x; h[j][2].y += a.y;
                } else {
                    a = cmul(ld_s<NT>(t + 0 * ny), mx); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(ld_s<NT>(t + 1 * ny), my); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(ld_s<NT>(t + 2 * ny), mz); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(ld_s<NT>(t + 3 * ny), mx); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(ld_s<NT>(t + 4 * ny), my); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(ld_s<NT>(t + 5 * ny), mz); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(ld_s<NT>(t + 6 * ny), mx); h[j][2].x += a.x; h[j][2].y += a.y;
                    a = cmul(ld_s<NT>(t + 7 * ny), my); h[j][2].x += a.x; h[j][2].y += a.y;
                    a = cmul(ld_s<NT>(t + 8 * ny), mz); h[j][2].x += a.x; h[j][2].y += a.y;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < OC; j++) {
            const int ko = kc0 + j;
            if (ko >= nl) break;
#pragma unroll
            for (int a = 0; a < 3; a++) st_s<NT>(S2 + ((size_t)(3 * ko + a)) * Kc + idx, h[j][a]);
        }
    }
}

// The product with the z-offset tensor (tensor_block's 2 nl - 1 blocks): thread per
// (kx, ky), the NL layers' spectra and field sums all in registers.  The offsets run from
// NL-1 down to -(NL-1), so every field sum H_ko still adds its sources in ascending ks, as
// k_prod and the reference do (demag.f90:117-121), while each spectrum and each stored
// tensor value is read from memory once.
template <typename T, bool SYM, int NL, bool BAT = false, bool NT = false>
__global__ void __launch_bounds__(256) k_prod_offsets(DemagOps<T> one, const DemagOps<T>* __restrict__ tab, int ny,
                                                      int nxh, int nslot)
{
    const C2<T>* __restrict__ S1 = BAT ? tab[blockIdx.z].S1 : one.S1;
    C2<T>* __restrict__ S2 = BAT ? tab[blockIdx.z].S2 : one.S2;
    const C2<T>* __restrict__ Tspec = BAT ? tab[blockIdx.z].Tspec : one.Tspec;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t Kc = (size_t)nxh * ny;
    if (idx >= Kc) return;
    const int kx = (int)(idx / ny), ky = (int)(idx - (size_t)kx * ny);
    C2<T> m[NL][3], h[NL][3];
#pragma unroll
    for (int k = 0; k < NL; k++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            m[k][b] = ld_s<NT>(S1 + ((size_t)(3 * k + b)) * Kc + idx);
            h[k][b] = {0, 0};
        }
#pragma unroll
    for (int d = NL - 1; d >= -(NL - 1); d--) {
        const C2<T>* t = Tspec + tensor_block(kx, d > 0 ? d : 0, d > 0 ? 0 : -d, NL, 1) * nslot * ny + ky;
        C2<T> tv[SYM ? 6 : 9];
#pragma unroll
        for (int sl = 0; sl < (SYM ? 6 : 9); sl++) tv[sl] = ld_s<NT>(t + (size_t)sl * ny);
#pragma unroll
        for (int ko = 0; ko < NL; ko++) {
            const int ks = ko - d;
            if (ks < 0 || ks >= NL) continue;
            const C2<T> mx = m[ks][0], my = m[ks][1], mz = m[ks][2];
            C2<T> a;
            if (SYM) {
                a = cmul(tv[0], mx); h[ko][0].x += a.x; h[ko][0].y += a.y;
                a = cmul(tv[3], my); h[ko][0].x += a.x; h[ko][0].y += a.y;
                a = cmul(tv[4], mz); h[ko][0].x += a.x; h[ko][0].y += a.y;
                a = cmul(tv[3], mx); h[ko][1].x += a.x; h[ko][1].y += a.y;
                a = cmul(tv[1], my); h[ko][1].x += a.x; h[ko][1].y += a.y;
                a = cmul(tv[5], mz); h[ko][1].x += a.x; h[ko][1].y += a.y;
                a = cmul(tv[4], mx); h[ko][2].x += a.x; h[ko][2].y += a.y;
                a = cmul(tv[5], my); h[ko][2].x += a.x; h[ko][2].y += a.y;
                a = cmul(tv[2], mz); h[ko][2].x += a.x; h[ko][2].y += a.y;
            } else {
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    a = cmul(tv[3 * r + 0], mx); h[ko][r].x += a.x; h[ko][r].y += a.y;
                    a = cmul(tv[3 * r + 1], my); h[ko][r].x += a.x; h[ko][r].y += a.y;
                    a = cmul(tv[3 * r + 2], mz); h[ko][r].x += a.x; h[ko][r].y += a.y;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NL; k++)
#pragma unroll
        for (int b = 0; b < 3; b++) st_s<NT>(S2 + ((size_t)(3 * k + b)) * Kc + idx, h[k][b]);
}

// c2r of TY rows of one (component, layer) from column-contiguous spectra (FFTW's c2r:
// imaginary parts of the DC and Nyquist bins ignored).
template <typename T, int E, bool WV = false, bool CELLS = false, bool BAT = false>
__global__ void __launch_bounds__(256) k_rowT_inv(DemagOps<T> one, const DemagOps<T>* __restrict__ tab,
                                                  int nx, int ny, int sy, int mx, int nl, int nxh, int TY, FFTPlan px,
                                                  const C2<T>* __restrict__ twx)
{
    const DemagOps<T> o = BAT ? tab[blockIdx.z] : one;
    const C2<T>* __restrict__ S = o.S2;
    T* __restrict__ Hcell = o.Hcell;
    double* __restrict__ Hr = o.Hr;
    const RowRuns R = o.R;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int roff[9];
    const C2<T>* __restrict__ tw = twx;   // from L2, as in k_rowT_fwd
    const int B = TY / 2;
    T* re = reinterpret_cast<T*>(smem_raw);
    T* im = re + LPN(B * nx);
    const int nyb = (ny + TY - 1) / TY;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if (L >= nyb * 3 * nl) return;
    const int y0 = (L % nyb) * TY, kc = L / nyb;
    const int k = kc / 3, c = kc - 3 * k;
    const C2<T>* src = S + (size_t)kc * nxh * sy;
    const int nrows = min(TY, ny - y0);
    RunsAhead<MRMG_RU(1, E)> ra;
    if (!CELLS) ra.prefetch(R, k, ny, y0, nrows);
    // each thread takes one (kx, row pair) and forms z = A + iB at kx and its mirror.  B =
    // TY/2 is 1, 2 or 4: kx and p are a shift and a mask of idx
    const int lb = B == 4 ? 2 : B == 2 ? 1 : 0;
    staged_load<(E < 4 ? E : 4), T>(
        nxh * B,
        [&](int idx) {
            const int kx = idx >> lb, p = idx & (B - 1);
            const int ya = y0 + 2 * p, yb = ya + 1;
            return C4<T>{src[(size_t)kx * sy + (ya < ny ? ya : ny - 1)], src[(size_t)kx * sy + (yb < ny ? yb : ny - 1)]};
        },
        [&](int idx, C4<T> v) {
            const int kx = idx >> lb, p = idx & (B - 1);
            const int ya = y0 + 2 * p;
            C2<T> A = ya < ny ? v.a : C2<T>{0, 0}, Bv = ya + 1 < ny ? v.b : C2<T>{0, 0};
            if (kx == 0 || 2 * kx == nx) { A.y = 0; Bv.y = 0; }
            re[LP(p * nx + kx)] = A.x - Bv.y;
            im[LP(p * nx + kx)] = A.y + Bv.x;
            if (kx > 0 && nx - kx != kx) {
                re[LP(p * nx + nx - kx)] = A.x + Bv.y;
                im[LP(p * nx + nx - kx)] = Bv.x - A.y;
            }
        });
    __syncthreads();
    if (!CELLS && (int)threadIdx.x <= nrows) roff[threadIdx.x] = R.row_off[k * ny + min(y0 + (int)threadIdx.x, ny)];
    if (WV) {
        lds_fft_wave<T, E, true>(re, im, B, px, tw);
        __syncthreads();
    } else {
        lds_fft<T, E, true, true>(re, im, B, px, tw);
    }
    if (!CELLS) {
        lds_to_runs<T, 1, MRMG_RU(1, E)>(R, ra, Hr, c, nrows, B, nx, re, im, roff);
        return;
    }
    T* hc = Hcell + ((size_t)c * nl + k) * (size_t)mx * ny;
    for (int idx = threadIdx.x; idx < B * mx; idx += blockDim.x) {
        const int p = idx / mx, x = idx - p * mx;
        const int ya = y0 + 2 * p, yb = ya + 1;
        if (ya < ny) hc[(size_t)ya * mx + x] = re[LP(p * nx + x)];
        if (yb < ny) hc[(size_t)yb * mx + x] = im[LP(p * nx + x)];
    }
}

template <typename T>
template <int MB> int Engine<T>::lds_attr_mb()
{
    const int r = (int)lds_row;
    HIPCHK(hipFuncSetAttribute((const void*)k_row_fwd<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_row_inv<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_row_inv<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_row_fwd<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_row_inv<T, MB, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               r));
    RCHK(lds_attr_col<MB>());
    return MRMG_OK;
}

template <typename T>
template <int MB> int Engine<T>::lds_attr_split()
{
    const int r = (int)lds_rowT, cl = (int)lds_colT;
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_fwd<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_fwd<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               r));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               r));
    HIPCHK(hipFuncSetAttribute((const void*)k_colT<T, MB, false>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_colT<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_fwd<T, MB, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               r));
    HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, r));
    HIPCHK(hipFuncSetAttribute((const void*)k_colT<T, MB, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_colT<T, MB, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    return MRMG_OK;
}

// dynamic LDS above 64 KiB must be requested per kernel
template <typename T>
int Engine<T>::lds_attr()
{
    if (fused_ok && (lds_row > 64 * 1024 || lds_col > 64 * 1024)) {
        RCHK(lds_attr_mb<4>()); RCHK(lds_attr_mb<8>()); RCHK(lds_attr_mb<16>());
        RCHK(lds_attr_mb<32>()); RCHK(lds_attr_mb<64>());
    }
    if (split_ok && (lds_rowT > 64 * 1024 || lds_colT > 64 * 1024)) {
        RCHK(lds_attr_split<4>()); RCHK(lds_attr_split<8>()); RCHK(lds_attr_split<16>());
        RCHK(lds_attr_split<32>()); RCHK(lds_attr_split<64>());
    }
    return MRMG_OK;
}

// one launch of a demag kernel class (K_ROW_FWD, K_COL, K_ROW_INV on the fused paths;
// K_ROW_FWD, K_COLF, K_PROD, K_COLI, K_ROW_INV on the split path); with a batch table set
// (btab, bn members), one launch for all members, blockIdx.z the member
// k_prod_offsets for NL layers: six or nine tensor slots, one context or a batch, cached or
// non-temporal streams
template <typename T, int NL>
static void launch_prod_offsets(bool sym, bool bat, bool nt, dim3 g, hipStream_t s, const DemagOps<T>& o,
                                const DemagOps<T>* tab, int fy, int nxh, int nslot)
{
#define KPO(S, B, N)                                                                                                 \
    if (sym == S && bat == B && nt == N)                                                                             \
        hipLaunchKernelGGL((k_prod_offsets<T, S, NL, B, N>), g, dim3(256), 0, s, o, tab, fy, nxh, nslot);
    KPO(true, true, true) else KPO(true, true, false) else KPO(true, false, true) else KPO(true, false, false)
    else KPO(false, true, true) else KPO(false, true, false) else KPO(false, false, true) else KPO(false, false, false)
#undef KPO
}

template <typename T>
int Engine<T>::launch_class(int cls)
{
    // row kernels: the media rows; column and product kernels: the FFT grid
    const int ny = c->ny, fy = c->fy, nl = c->nl, nxh = c->nxh;
    const unsigned z = btab ? (unsigned)bn : 1u;
    const dim3 grow((ny + 1) / 2, nl, z), growT(rowT_blocks(), 1, z), blk(row_threads), b256(256);
    const int ncols = 3 * nl * nxh;
    const dim3 gcol((ncols + CBc - 1) / CBc, 1, z);
    if (btab && (wv || cells_out))
        return set_err(MRMG_ERR_STATE, "batched demag launches run without wave-local FFTs or cell output");
    switch (cls) {
    case K_ROW_FWD:
        if (tl || split) LAUNCH_E(maxb_rowT, launch_rowT_fwd, growT, blk);
        else LAUNCH_E(maxb_row, launch_row_fwd, grow, blk);
        break;
    case K_COL: {
        if (split) return set_err(MRMG_ERR_STATE, "the split path has no fused column kernel");
        const int nb8 = (nxh * kg + 7) / 8 * 8;
        if (nslot == 6) LAUNCH_E(maxb_col, launch_col_sym, nb8);
        else if (col_threads > 512) return set_err(MRMG_ERR_STATE, "col_threads 1024 needs the 6-slot tensor");
        else LAUNCH_E(maxb_col, launch_col_gen, nb8);
        break;
    }
    case K_COLF:
    case K_COLI:
        if (!split && !(pre && cls == K_COLF))
            return set_err(MRMG_ERR_STATE, "the fused paths have no separate column FFTs");
        if (cls == K_COLF) LAUNCH_E(maxb_colT, launch_colT_fwd, gcol, b256);
        else LAUNCH_E(maxb_colT, launch_colT_inv, gcol, b256);
        break;
    case K_PROD: {
        if (!split) return set_err(MRMG_ERR_STATE, "the fused paths have no separate product kernel");
        const size_t Kc = (size_t)nxh * fy;
        const dim3 gp((unsigned)((Kc + 255) / 256), 1, z);
        const DemagOps<T> o = ops();
        const bool sym = nslot == 6, bat = btab != nullptr, nt = prod_streams();
        if (toep && nl <= 8) {
            switch (nl) {
            case 1: launch_prod_offsets<T, 1>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            case 2: launch_prod_offsets<T, 2>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            case 3: launch_prod_offsets<T, 3>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            case 4: launch_prod_offsets<T, 4>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            case 5: launch_prod_offsets<T, 5>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            case 6: launch_prod_offsets<T, 6>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            case 7: launch_prod_offsets<T, 7>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            default: launch_prod_offsets<T, 8>(sym, bat, nt, gp, c->stream, o, btab, fy, nxh, nslot); break;
            }
        } else {
#define KPR(S, B, N)                                                                                                 \
    if (sym == S && bat == B && nt == N)                                                                             \
        hipLaunchKernelGGL((k_prod<T, S, 4, B, N>), gp, b256, 0, c->stream, o, btab, fy, nl, nxh, nslot, (int)toep);
            KPR(true, true, true) else KPR(true, true, false) else KPR(true, false, true) else KPR(true, false, false)
            else KPR(false, true, true) else KPR(false, true, false) else KPR(false, false, true)
            else KPR(false, false, false)
#undef KPR
        }
        break;
    }
    case K_ROW_INV:
        if (tl || split) LAUNCH_E(maxb_rowT, launch_rowT_inv, growT, blk);
        else LAUNCH_E(maxb_row, launch_row_inv, grow, blk);
        break;
    default:
        return set_err(MRMG_ERR_ARG, "not a demag kernel class");
    }
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

template <typename T>
template <int MB> void Engine<T>::launch_rowT_fwd(dim3 g, dim3 b)
{
    const DemagOps<T> o = ops();
    if (btab)
        hipLaunchKernelGGL((k_rowT_fwd<T, MB, false, true>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy,
                           c->nl, c->nxh, TYr, prow, twx.p);
    else if (wv)
        hipLaunchKernelGGL((k_rowT_fwd<T, MB, true>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy, c->nl,
                           c->nxh, TYr, prow, twx.p);
    else
        hipLaunchKernelGGL((k_rowT_fwd<T, MB>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy, c->nl, c->nxh,
                           TYr, prow, twx.p);
}

template <typename T>
template <int MB> void Engine<T>::launch_colT_fwd(dim3 g, dim3 b)
{
    const DemagOps<T> o = ops();
    if (btab)
        hipLaunchKernelGGL((k_colT<T, MB, false, true>), g, b, lds_colT, c->stream, o, btab, c->fy, c->ny,
                           3 * c->nl * c->nxh, CBc, pcol, twy.p);
    else
        hipLaunchKernelGGL((k_colT<T, MB, false>), g, b, lds_colT, c->stream, o, btab, c->fy, c->ny,
                           3 * c->nl * c->nxh, CBc, pcol, twy.p);
}

template <typename T>
template <int MB> void Engine<T>::launch_colT_inv(dim3 g, dim3 b)
{
    const DemagOps<T> o = ops();
    if (btab)
        hipLaunchKernelGGL((k_colT<T, MB, true, true>), g, b, lds_colT, c->stream, o, btab, c->fy, c->ny,
                           3 * c->nl * c->nxh, CBc, pcol, twy.p);
    else
        hipLaunchKernelGGL((k_colT<T, MB, true>), g, b, lds_colT, c->stream, o, btab, c->fy, c->ny,
                           3 * c->nl * c->nxh, CBc, pcol, twy.p);
}

template <typename T>
template <int MB> void Engine<T>::launch_rowT_inv(dim3 g, dim3 b)
{
    const DemagOps<T> o = ops();
    if (btab)
        hipLaunchKernelGGL((k_rowT_inv<T, MB, false, false, true>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny,
                           c->fy, c->nx, c->nl, c->nxh, TYr, prow, twx.p);
    else if (wv && cells_out)
        hipLaunchKernelGGL((k_rowT_inv<T, MB, true, true>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy,
                           c->nx, c->nl, c->nxh, TYr, prow, twx.p);
    else if (wv)
        hipLaunchKernelGGL((k_rowT_inv<T, MB, true>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy, c->nx,
                           c->nl, c->nxh, TYr, prow, twx.p);
    else if (cells_out)
        hipLaunchKernelGGL((k_rowT_inv<T, MB, false, true>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy,
                           c->nx, c->nl, c->nxh, TYr, prow, twx.p);
    else
        hipLaunchKernelGGL((k_rowT_inv<T, MB>), g, b, lds_rowT, c->stream, o, btab, c->fx, c->ny, c->fy, c->nx, c->nl,
                           c->nxh, TYr, prow, twx.p);
}

template <typename T>
template <int MB> void Engine<T>::launch_row_fwd(dim3 g, dim3 b)
{
    const DemagOps<T> o = ops();
    if (btab)
        hipLaunchKernelGGL((k_row_fwd<T, MB, true>), g, b, lds_row, c->stream, o, btab, c->fx, c->ny, c->fy, c->nl,
                           c->nxh, prow, twx.p);
    else
        hipLaunchKernelGGL((k_row_fwd<T, MB>), g, b, lds_row, c->stream, o, btab, c->fx, c->ny, c->fy, c->nl, c->nxh,
                           prow, twx.p);
}

template <typename T>
template <int MB> void Engine<T>::launch_row_inv(dim3 g, dim3 b)
{
    const DemagOps<T> o = ops();
    if (btab)
        hipLaunchKernelGGL((k_row_inv<T, MB, false, true>), g, b, lds_row, c->stream, o, btab, c->fx, c->ny, c->fy,
                           c->nx, c->nl, c->nxh, prow, twx.p);
    else if (cells_out)
        hipLaunchKernelGGL((k_row_inv<T, MB, true>), g, b, lds_row, c->stream, o, btab, c->fx, c->ny, c->fy, c->nx,
                           c->nl, c->nxh, prow, twx.p);
    else
        hipLaunchKernelGGL((k_row_inv<T, MB>), g, b, lds_row, c->stream, o, btab, c->fx, c->ny, c->fy, c->nx, c->nl,
                           c->nxh, prow, twx.p);
}

// each precision form is its own object (the unit is compiled with MRMG_FORM=float and
// with MRMG_FORM=double)
template int Engine<MRMG_FORM>::lds_attr();
template int Engine<MRMG_FORM>::launch_class(int);
