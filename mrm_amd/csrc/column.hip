// column.hip - the fused column kernel of the demag: y-FFT of the 3 n_lay source columns of
// one kx in LDS, the k-space product of demag.f90:117-121 and the inverse y-FFT, with the
// Engine members that launch it.  Compiled once per precision form (MRMG_FORM).
#include "engine.h"
#include "device_common.h"

#ifndef MRMG_FORM
#error "compile with -DMRMG_FORM=float or -DMRMG_FORM=double"
#endif

// Column pass with the k-space product fused (demag.f90:117-121): one workgroup per
// (kx, group of observation layers).  The 3*n_lay source columns of this kx are
// transformed along y in LDS, the field columns of the group's layers are formed in place
// (all products read into registers, barrier, then written over the group's own source
// columns), and those columns are transformed back.  kg groups per kx keep all CUs busy.
#define COL_MAX_THREADS 512
// TL: spectra in the column-contiguous layout [k][c][kx][y] of the split path (whole
// columns are contiguous and the loads coalesce), else [k][c][y][kx].
// WV: each wave transforms its own run of columns (no workgroup barrier inside the FFTs);
// E then bounds one lane's elements of its wave's run.
// BIG: 1024-thread workgroups (twice the lanes per FFT pass, at most 128 VGPRs), two
// product items per thread.
// BAT: one launch over a batch of contexts (blockIdx.z the member, operands from tab).
// PRE: the columns in S1 are already transformed along y (by k_colT), so the kernel only
// forms the products and the inverse transforms of its observation layers.
// ny: the column length (the FFT grid); nyv: the rows the row kernels write and read
// (the media rows; below ny for open boundaries, whose padding rows are zero in S1 and
// not needed in S2).
template <typename T, int E, bool SYM, bool TL = false, bool WV = false, bool BIG = false, bool BAT = false,
          bool PRE = false>
__global__ void __launch_bounds__(BIG ? 1024 : COL_MAX_THREADS) k_col(DemagOps<T> one, const DemagOps<T>* __restrict__ tab,
                                                     int ny, int nyv, int nl, int nxh,
                                                     int nslot, int toep, int kg, int nb8, FFTPlan py,
                                                     const C2<T>* __restrict__ twy)
{
    const C2<T>* __restrict__ S1 = BAT ? tab[blockIdx.z].S1 : one.S1;
    C2<T>* __restrict__ S2 = BAT ? tab[blockIdx.z].S2 : one.S2;
    const C2<T>* __restrict__ Tspec = BAT ? tab[blockIdx.z].Tspec : one.Tspec;
    const int L = xcd_remap(blockIdx.x, nb8);
    const int kx = L / kg, grp = L - kx * kg;
    if (kx >= nxh) return;
    const int per = (nl + kg - 1) / kg;
    const int ko0 = grp * per, ko1 = ko0 + per < nl ? ko0 + per : nl;
    if (ko0 >= ko1) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int ncol = 3 * nl;
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    T* re = reinterpret_cast<T*>(tw + ny);
    T* im = re + LPN(ncol * ny);
    staged_load<(E < 8 ? E : 8)>(
        ncol * ny,
        [&](int idx) {
            const int col = idx / ny, y = idx - col * ny;
            if (!PRE && y >= nyv) return C2<T>{0, 0};
            return TL ? S1[((size_t)col * nxh + kx) * ny + y] : S1[((size_t)col * ny + y) * nxh + kx];
        },
        [&](int idx, C2<T> v) { re[LP(idx)] = v.x; im[LP(idx)] = v.y; }, tw, twy, ny);
    __syncthreads();
    if (PRE) {
    } else if (WV) {
        lds_fft_wave<T, E, false>(re, im, ncol, py, tw);
        __syncthreads();
    } else {
        lds_fft<T, E, false, true>(re, im, ncol, py, tw);
    }
    // product: work item (ko, ky); up to 4 items per thread held in registers.  The tensor
    // values of KSB source layers are loaded before any of them is used, so a thread waits
    // for memory once per KSB layers instead of once per layer.
    const int items = (ko1 - ko0) * ny;
    constexpr int MAXI = BIG ? 2 : 4, KSB = 2, NS = SYM ? 6 : 9;
    C2<T> hv[MAXI][3];
#pragma unroll
    for (int q = 0; q < MAXI; q++) {
        const int it = threadIdx.x + q * blockDim.x;
        if (it < items) {
            const int ko = ko0 + it / ny, ky = it % ny;
            const C2<T>* Tk = Tspec + ky;
            C2<T> h0 = {0, 0}, h1 = {0, 0}, h2 = {0, 0};
            for (int kb = 0; kb < nl; kb += KSB) {
                C2<T> tv[KSB][NS];
#pragma unroll
                for (int j = 0; j < KSB; j++) {
                    if (kb + j < nl) {
                        const C2<T>* t = Tk + tensor_block(kx, ko, kb + j, nl, toep) * nslot * ny;
#pragma unroll
                        for (int sl = 0; sl < NS; sl++) tv[j][sl] = t[sl * ny];
                    }
                }
#pragma unroll
                for (int j = 0; j < KSB; j++) {
                    const int ks = kb + j;
                    if (ks >= nl) break;
                    const C2<T> mx = {re[LP((3 * ks + 0) * ny + ky)], im[LP((3 * ks + 0) * ny + ky)]};
                    const C2<T> my = {re[LP((3 * ks + 1) * ny + ky)], im[LP((3 * ks + 1) * ny + ky)]};
                    const C2<T> mz = {re[LP((3 * ks + 2) * ny + ky)], im[LP((3 * ks + 2) * ny + ky)]};
                    C2<T> a;
                    if (SYM) {  // slots xx yy zz xy xz yz
                        const C2<T> txx = tv[j][0], tyy = tv[j][1], tzz = tv[j][2], txy = tv[j][3], txz = tv[j][4],
                                    tyz = tv[j][5];
                        a = cmul(txx, mx); h0.x += a.x; h0.y += a.y;
                        a = cmul(txy, my); h0.x += a.x; h0.y += a.y;
                        a = cmul(txz, mz); h0.x += a.x; h0.y += a.y;
                        a = cmul(txy, mx); h1.x += a.x; h1.y += a.y;
                        a = cmul(tyy, my); h1.x += a.x; h1.y += a.y;
                        a = cmul(tyz, mz); h1.x += a.x; h1.y += a.y;
                        a = cmul(txz, mx); h2.x += a.x; h2.y += a.y;
                        a = cmul(tyz, my); h2.x += a.x; h2.y += a.y;
                        a = cmul(tzz, mz); h2.x += a.x; h2.y += a.y;
                    } else {    // slot 3a+b: field component a from source component b
                        a = cmul(tv[j][0], mx); h0.x += a.x; h0.y += a.y;
                        a = cmul(tv[j][1], my); h0.x += a.x; h0.y += a.y;
                        a = cmul(tv[j][2], mz); h0.x += a.x; h0.y += a.y;
                        a = cmul(tv[j][3], mx); h1.x += a.x; h1.y += a.y;
                        a = cmul(tv[j][4], my); h1.x += a.x; h1.y += a.y;
                        a = cmul(tv[j][5], mz); h1.x += a.x; h1.y += a.y;
                        a = cmul(tv[j][6], mx); h2.x += a.x; h2.y += a.y;
                        a = cmul(tv[j][7], my); h2.x += a.x; h2.y += a.y;
                        a = cmul(tv[j][8], mz); h2.x += a.x; h2.y += a.y;
                    }
                }
            }
            hv[q][0] = h0; hv[q][1] = h1; hv[q][2] = h2;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXI; q++) {
        const int it = threadIdx.x + q * blockDim.x;
        if (it < items) {
            const int ko = ko0 + it / ny, ky = it % ny;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                re[LP((3 * ko + a) * ny + ky)] = hv[q][a].x;
                im[LP((3 * ko + a) * ny + ky)] = hv[q][a].y;
            }
        }
    }
    __syncthreads();
    const int e0 = 3 * ko0 * ny, ne = 3 * (ko1 - ko0) * ny;
    if (WV) {
        lds_fft_wave<T, E, true>(re, im, 3 * (ko1 - ko0), py, tw, e0);
        __syncthreads();
    } else {
        lds_fft<T, E, true, true>(re, im, 3 * (ko1 - ko0), py, tw, e0);
    }
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const int idx = e0 + i;
        const int col = idx / ny, y = idx - col * ny;
        if (y >= nyv) continue;
        const C2<T> v = {re[LP(idx)], im[LP(idx)]};
        if (TL) S2[((size_t)col * nxh + kx) * ny + y] = v;
        else S2[((size_t)col * ny + y) * nxh + kx] = v;
    }
}

#define COL_ARGS o, btab, c->fy, c->ny, c->nl, c->nxh, nslot, (int)toep, kg, nb8, pcol, twy.p
template <typename T>
template <int MB> void Engine<T>::launch_col_sym(int nb8)
{
    const DemagOps<T> o = ops();
    const dim3 g(nb8, 1, btab ? (unsigned)bn : 1u), b(col_threads);
    if (pre) {
        if (col_threads > 512)
            hipLaunchKernelGGL((k_col<T, MB, true, true, false, true, false, true>), g, b, lds_col, c->stream,
                               COL_ARGS);
        else
            hipLaunchKernelGGL((k_col<T, MB, true, true, false, false, false, true>), g, b, lds_col, c->stream,
                               COL_ARGS);
    } else if (btab) {
        if (tl)
            hipLaunchKernelGGL((k_col<T, MB, true, true, false, false, true>), g, b, lds_col, c->stream, COL_ARGS);
        else
            hipLaunchKernelGGL((k_col<T, MB, true, false, false, false, true>), g, b, lds_col, c->stream, COL_ARGS);
    } else if (col_threads > 512) {
        if (tl)
            hipLaunchKernelGGL((k_col<T, MB, true, true, false, true>), g, b, lds_col, c->stream, COL_ARGS);
        else
            hipLaunchKernelGGL((k_col<T, MB, true, false, false, true>), g, b, lds_col, c->stream, COL_ARGS);
    } else if (tl && wv)
        hipLaunchKernelGGL((k_col<T, MB, true, true, true>), g, b, lds_col, c->stream, COL_ARGS);
    else if (tl)
        hipLaunchKernelGGL((k_col<T, MB, true, true>), g, b, lds_col, c->stream, COL_ARGS);
    else
        hipLaunchKernelGGL((k_col<T, MB, true>), g, b, lds_col, c->stream, COL_ARGS);
}

template <typename T>
template <int MB> void Engine<T>::launch_col_gen(int nb8)
{
    const DemagOps<T> o = ops();
    const dim3 g(nb8, 1, btab ? (unsigned)bn : 1u), b(col_threads);
    if (pre)
        hipLaunchKernelGGL((k_col<T, MB, false, true, false, false, false, true>), g, b, lds_col, c->stream,
                           COL_ARGS);
    else if (btab) {
        if (tl)
            hipLaunchKernelGGL((k_col<T, MB, false, true, false, false, true>), g, b, lds_col, c->stream, COL_ARGS);
        else
            hipLaunchKernelGGL((k_col<T, MB, false, false, false, false, true>), g, b, lds_col, c->stream, COL_ARGS);
    } else if (tl && wv)
        hipLaunchKernelGGL((k_col<T, MB, false, true, true>), g, b, lds_col, c->stream, COL_ARGS);
    else if (tl)
        hipLaunchKernelGGL((k_col<T, MB, false, true>), g, b, lds_col, c->stream, COL_ARGS);
    else
        hipLaunchKernelGGL((k_col<T, MB, false>), g, b, lds_col, c->stream, COL_ARGS);
}
#undef COL_ARGS

template <typename T>
template <int MB> int Engine<T>::lds_attr_col()
{
    const int cl = (int)lds_col;
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, false>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, false, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, true, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, true, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, false, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, false, true, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, false, false, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, true, false, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, true, true, false, true, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    HIPCHK(hipFuncSetAttribute((const void*)k_col<T, MB, false, true, false, false, false, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, cl));
    return MRMG_OK;
}

// The instantiations per elements-per-lane bound; MRMG_COL_PART 1 or 2 compiles one half
// of them (the two halves are separate objects, built in parallel), unset both.
#if !defined(MRMG_COL_PART) || MRMG_COL_PART == 1
template void Engine<MRMG_FORM>::launch_col_sym<4>(int);
template void Engine<MRMG_FORM>::launch_col_gen<4>(int);
template int Engine<MRMG_FORM>::lds_attr_col<4>();
template void Engine<MRMG_FORM>::launch_col_sym<8>(int);
template void Engine<MRMG_FORM>::launch_col_gen<8>(int);
template int Engine<MRMG_FORM>::lds_attr_col<8>();
template void Engine<MRMG_FORM>::launch_col_sym<16>(int);
template void Engine<MRMG_FORM>::launch_col_gen<16>(int);
template int Engine<MRMG_FORM>::lds_attr_col<16>();
#endif
#if !defined(MRMG_COL_PART) || MRMG_COL_PART == 2
template void Engine<MRMG_FORM>::launch_col_sym<32>(int);
template void Engine<MRMG_FORM>::launch_col_gen<32>(int);
template int Engine<MRMG_FORM>::lds_attr_col<32>();
template void Engine<MRMG_FORM>::launch_col_sym<64>(int);
template void Engine<MRMG_FORM>::launch_col_gen<64>(int);
template int Engine<MRMG_FORM>::lds_attr_col<64>();
#endif
