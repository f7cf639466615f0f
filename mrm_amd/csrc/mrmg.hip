// mrmg.hip - MI355X (gfx950) implementation of mrm's granular-media time step behind the
// C ABI of include/mrmg.h.  Reference behaviour (paths relative to the reference's src/):
//   data.f90:322-390   get_mag      -> k_row_fwd   (scatter fused into the row r2c FFT)
//   demag.f90:33-135   get_demag    -> k_row_fwd, k_col (column FFT + tensor product +
//                                       inverse column FFT in one LDS pass), k_row_inv
//   field.f90:160-273  get_field    -> k_grain     (gather, anisotropy, exchange, applied)
//   field.f90:397-436  get_thermal  -> k_grain     (injected normals or Philox4x32-10)
//   llg.f90:1066-1155  push_llg     -> k_grain     (Heun / Boris / quaternion)
//   llg.f90:1195-1271  push_m       -> mrmg_step   (stage sequencing)
//   tensor_element.f90:13-102, 219-261 -> k_plane_to_plane (FP64 tensor build)
//   demag.f90:259-267  init_demag   -> k_tensor_rows / k_tensor_cols / k_tensor_pack
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mrmg.h"
#include "kernels.h"

// ------------------------------------------------------------------------- errors
static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess)                                                                 \
            return set_err(MRMG_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,             \
                           hipGetErrorString(e_));                                            \
    } while (0)

#define RCHK(x)                        \
    do {                               \
        int r_ = (x);                  \
        if (r_ != MRMG_OK) return r_;  \
    } while (0)

// ------------------------------------------------------------------------- device helpers
template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* t)
{
    t[0] = a[1] * b[2] - a[2] * b[1];
    t[1] = a[2] * b[0] - a[0] * b[2];
    t[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename T>
__device__ __forceinline__ T dot3(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter (grain, step lo, step hi, 0),
// key = seed.  Box-Muller on the first two pairs gives the grain's three normals.
__device__ __forceinline__ void philox4x32_10(unsigned int c[4], unsigned int k0, unsigned int k1)
{
    const unsigned int M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        unsigned int hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        unsigned int hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        unsigned int n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

__device__ __forceinline__ void philox_normals3(unsigned long long seed, unsigned long long step, int g, double* n)
{
    unsigned int c[4] = {(unsigned int)g, (unsigned int)(step & 0xffffffffull), (unsigned int)(step >> 32), 0u};
    philox4x32_10(c, (unsigned int)(seed & 0xffffffffull), (unsigned int)(seed >> 32));
    // u in (0,1]: (x + 1) * 2^-32 ; angle from the second word
    const double inv32 = 2.3283064365386963e-10;
    double u1 = ((double)c[0] + 1.0) * inv32, a1 = (double)c[1] * inv32 * 6.283185307179586;
    double u2 = ((double)c[2] + 1.0) * inv32, a2 = (double)c[3] * inv32 * 6.283185307179586;
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u2));
    double s1, c1, s2, c2;
    sincos(a1, &s1, &c1);
    sincos(a2, &s2, &c2);
    n[0] = r1 * c1;
    n[1] = r1 * s1;
    n[2] = r2 * c2;
}

__global__ void k_philox_dump(unsigned long long seed, unsigned long long step0, int nsteps, int G, double* out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nsteps * G) return;
    int it = (int)(i / G), g = (int)(i - (long long)it * G);
    double n[3];
    philox_normals3(seed, step0 + it, g, n);
    out[3 * i] = n[0]; out[3 * i + 1] = n[1]; out[3 * i + 2] = n[2];
}

// XCD-aware remap: hardware places block b on XCD b % 8; give the blocks of one XCD a
// contiguous range of logical ids so neighbouring columns share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb8)
{
    const int per = nb8 >> 3;
    return (b & 7) * per + (b >> 3);
}

// ------------------------------------------------------------------------- demag kernels
// Row r2c: one workgroup per (row pair, layer).  Two rows of each component of the cell
// magnetisation (written by the grain kernels, get_mag data.f90:344-353) are packed as
// z = M(y0) + i M(y1), transformed with one complex FFT and separated.
template <typename T, int E>
__global__ void __launch_bounds__(256) k_row_fwd(const T* __restrict__ Mcell, C2<T>* __restrict__ S,
                                                 int nx, int ny, int nl, int nxh, FFTPlan px,
                                                 const C2<T>* __restrict__ twx)
{
    extern __shared__ unsigned char smem_raw[];
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    T* re = reinterpret_cast<T*>(tw + nx);
    T* im = re + LPN(3 * nx);
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1, k = blockIdx.y;
    const bool has1 = y1 < ny;
    const size_t CP = (size_t)nl * ny * nx;
    const T* r0 = Mcell + ((size_t)k * ny + y0) * nx;
    staged_load<(E < 8 ? E : 8)>(
        3 * nx,
        [&](int idx) {
            const int c = idx / nx, x = idx - c * nx;
            return C2<T>{r0[c * CP + x], r0[c * CP + (has1 ? nx : 0) + x]};
        },
        [&](int idx, C2<T> v) { re[LP(idx)] = v.x; im[LP(idx)] = has1 ? v.y : (T)0; }, tw, twx, nx);
    __syncthreads();
    lds_fft<T, E, false>(re, im, 3, px, tw);
    for (int idx = threadIdx.x; idx < 3 * nxh; idx += blockDim.x) {
        const int c = idx / nxh, kx = idx - c * nxh;
        const int kn = kx == 0 ? 0 : nx - kx;
        const T zr = re[LP(c * nx + kx)], zi = im[LP(c * nx + kx)], nr = re[LP(c * nx + kn)], ni = im[LP(c * nx + kn)];
        C2<T> A = {(zr + nr) * (T)0.5, (zi - ni) * (T)0.5};
        C2<T> B = {(zi + ni) * (T)0.5, (nr - zr) * (T)0.5};
        C2<T>* row = S + (((size_t)k * 3 + c) * ny + y0) * nxh;
        row[kx] = A;
        if (has1) row[nxh + kx] = B;
    }
}

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
template <typename T, int E, bool SYM, bool TL = false, bool WV = false, bool BIG = false>
__global__ void __launch_bounds__(BIG ? 1024 : COL_MAX_THREADS) k_col(const C2<T>* __restrict__ S1, C2<T>* __restrict__ S2,
                                                     const C2<T>* __restrict__ Tspec, int ny, int nl, int nxh,
                                                     int nslot, int kg, int nb8, FFTPlan py,
                                                     const C2<T>* __restrict__ twy)
{
    const int L = xcd_remap(blockIdx.x, nb8);
    const int kx = L / kg, grp = L - kx * kg;
    if (kx >= nxh) return;
    const int per = (nl + kg - 1) / kg;
    const int ko0 = grp * per, ko1 = ko0 + per < nl ? ko0 + per : nl;
    if (ko0 >= ko1) return;
    extern __shared__ unsigned char smem_raw[];
    const int ncol = 3 * nl;
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    T* re = reinterpret_cast<T*>(tw + ny);
    T* im = re + LPN(ncol * ny);
    staged_load<(E < 8 ? E : 8)>(
        ncol * ny,
        [&](int idx) {
            const int col = idx / ny, y = idx - col * ny;
            return TL ? S1[((size_t)col * nxh + kx) * ny + y] : S1[((size_t)col * ny + y) * nxh + kx];
        },
        [&](int idx, C2<T> v) { re[LP(idx)] = v.x; im[LP(idx)] = v.y; }, tw, twy, ny);
    __syncthreads();
    if (WV) {
        lds_fft_wave<T, E, false>(re, im, ncol, py, tw);
        __syncthreads();
    } else {
        lds_fft<T, E, false>(re, im, ncol, py, tw);
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
            const C2<T>* Tk = Tspec + (((size_t)kx * nl + ko) * nl) * nslot * ny + ky;
            C2<T> h0 = {0, 0}, h1 = {0, 0}, h2 = {0, 0};
            for (int kb = 0; kb < nl; kb += KSB) {
                C2<T> tv[KSB][NS];
#pragma unroll
                for (int j = 0; j < KSB; j++) {
                    if (kb + j < nl) {
                        const C2<T>* t = Tk + (size_t)(kb + j) * nslot * ny;
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
        lds_fft<T, E, true>(re, im, 3 * (ko1 - ko0), py, tw, e0);
    }
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const int idx = e0 + i;
        const int col = idx / ny, y = idx - col * ny;
        const C2<T> v = {re[LP(idx)], im[LP(idx)]};
        if (TL) S2[((size_t)col * nxh + kx) * ny + y] = v;
        else S2[((size_t)col * ny + y) * nxh + kx] = v;
    }
}

// Row c2r: FFTW's 1D c2r semantics (imaginary parts of the DC and Nyquist bins ignored).
template <typename T, int E>
__global__ void __launch_bounds__(256) k_row_inv(const C2<T>* __restrict__ S2, T* __restrict__ Hcell,
                                                 int nx, int ny, int nl, int nxh, FFTPlan px,
                                                 const C2<T>* __restrict__ twx)
{
    extern __shared__ unsigned char smem_raw[];
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    T* re = reinterpret_cast<T*>(tw + nx);
    T* im = re + LPN(3 * nx);
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1, k = blockIdx.y;
    const bool has1 = y1 < ny;
    staged_load<(E < 4 ? E : 4)>(
        3 * nxh,
        [&](int idx) {
            const int c = idx / nxh, kx = idx - c * nxh;
            const C2<T>* row = S2 + (((size_t)k * 3 + c) * ny + y0) * nxh;
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
    lds_fft<T, E, true>(re, im, 3, px, tw);
    const size_t plane = (size_t)nx * ny;
    for (int idx = threadIdx.x; idx < 3 * nx; idx += blockDim.x) {
        const int c = idx / nx, x = idx - c * nx;
        T* hc = Hcell + ((size_t)c * nl + k) * plane;
        hc[(size_t)y0 * nx + x] = re[LP(idx)];
        if (has1) hc[(size_t)y1 * nx + x] = im[LP(idx)];
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
template <typename T, int E, bool WV = false>
__global__ void __launch_bounds__(256) k_rowT_fwd(const T* __restrict__ Mcell, C2<T>* __restrict__ S,
                                                  int nx, int ny, int nl, int nxh, int TY, FFTPlan px,
                                                  const C2<T>* __restrict__ twx)
{
    extern __shared__ unsigned char smem_raw[];
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    const int B = TY / 2;
    T* re = reinterpret_cast<T*>(tw + nx);
    T* im = re + LPN(B * nx);
    // 1D grid padded to a multiple of 8, XCD-contiguous: the row blocks that share the
    // cache lines of one kx run (TY*sizeof(C2<T>) < 128 bytes) are read/written in one L2
    const int nyb = (ny + TY - 1) / TY;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if (L >= nyb * 3 * nl) return;
    const int y0 = (L % nyb) * TY, kc = L / nyb;        // kc = k*3 + c
    const int k = kc / 3, c = kc - 3 * k;
    const size_t CP = (size_t)nl * ny * nx;
    const T* src = Mcell + c * CP + (size_t)k * ny * nx;
    staged_load<(E < 8 ? E : 8)>(
        B * nx,
        [&](int idx) {
            const int p = idx / nx, x = idx - p * nx;
            const int ya = y0 + 2 * p, yb = ya + 1;
            return C2<T>{src[(size_t)(ya < ny ? ya : ny - 1) * nx + x], src[(size_t)(yb < ny ? yb : ny - 1) * nx + x]};
        },
        [&](int idx, C2<T> v) {
            const int p = idx / nx, ya = y0 + 2 * p;
            re[LP(idx)] = ya < ny ? v.x : (T)0;
            im[LP(idx)] = ya + 1 < ny ? v.y : (T)0;
        },
        tw, twx, nx);
    __syncthreads();
    if (WV) {
        lds_fft_wave<T, E, false>(re, im, B, px, tw);
        __syncthreads();
    } else {
        lds_fft<T, E, false>(re, im, B, px, tw);
    }
    C2<T>* dst = S + (size_t)kc * nxh * ny;
    for (int idx = threadIdx.x; idx < nxh * TY; idx += blockDim.x) {
        const int kx = idx / TY, yy = idx - kx * TY;
        const int y = y0 + yy;
        if (y >= ny) continue;
        const int p = yy >> 1;
        const int kn = kx == 0 ? 0 : nx - kx;
        const T zr = re[LP(p * nx + kx)], zi = im[LP(p * nx + kx)], nr = re[LP(p * nx + kn)], ni = im[LP(p * nx + kn)];
        C2<T> v;
        if ((yy & 1) == 0) v = {(zr + nr) * (T)0.5, (zi - ni) * (T)0.5};
        else v = {(zi + ni) * (T)0.5, (nr - zr) * (T)0.5};
        dst[(size_t)kx * ny + y] = v;
    }
}

// y-FFT of ncols contiguous columns of length ny, CB columns per workgroup, in place.
// Twiddles are read from the global table (cached), not staged.
template <typename T, int E, bool INV>
__global__ void __launch_bounds__(256) k_colT(C2<T>* __restrict__ S, int ny, int ncols, int CB, FFTPlan py,
                                              const C2<T>* __restrict__ twy)
{
    extern __shared__ unsigned char smem_raw[];
    T* re = reinterpret_cast<T*>(smem_raw);
    const int q0 = blockIdx.x * CB;
    const int nb = min(CB, ncols - q0);
    T* im = re + LPN(CB * ny);
    C2<T>* base = S + (size_t)q0 * ny;
    staged_load<(E < 8 ? E : 8), T>(
        nb * ny, [&](int idx) { return base[idx]; }, [&](int idx, C2<T> v) { re[LP(idx)] = v.x; im[LP(idx)] = v.y; });
    __syncthreads();
    lds_fft<T, E, INV>(re, im, nb, py, twy);
    for (int idx = threadIdx.x; idx < nb * ny; idx += blockDim.x) base[idx] = C2<T>{re[LP(idx)], im[LP(idx)]};
}

// The k-space product over column-contiguous spectra: thread per (kx, ky), observation
// layers in chunks of OC, each output summed over (ks, b) in the reference's order.
template <typename T, bool SYM, int OC>
__global__ void __launch_bounds__(256) k_prod(const C2<T>* __restrict__ S1, C2<T>* __restrict__ S2,
                                              const C2<T>* __restrict__ Tspec, int ny, int nl, int nxh, int nslot)
{
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
                const C2<T>* t = Tspec + ((((size_t)kx * nl + ko) * nl + ks) * nslot) * ny + ky;
                C2<T> a;
                if (SYM) {
                    const C2<T> txx = t[0], tyy = t[ny], tzz = t[2 * ny], txy = t[3 * ny], txz = t[4 * ny], tyz = t[5 * ny];
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
                    a = cmul(t[0 * ny], mx); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(t[1 * ny], my); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(t[2 * ny], mz); h[j][0].x += a.x; h[j][0].y += a.y;
                    a = cmul(t[3 * ny], mx); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(t[4 * ny], my); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(t[5 * ny], mz); h[j][1].x += a.x; h[j][1].y += a.y;
                    a = cmul(t[6 * ny], mx); h[j][2].x += a.x; h[j][2].y += a.y;
                    a = cmul(t[7 * ny], my); h[j][2].x += a.x; h[j][2].y += a.y;
                    a = cmul(t[8 * ny], mz); h[j][2].x += a.x; h[j][2].y += a.y;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < OC; j++) {
            const int ko = kc0 + j;
            if (ko >= nl) break;
#pragma unroll
            for (int a = 0; a < 3; a++) S2[((size_t)(3 * ko + a)) * Kc + idx] = h[j][a];
        }
    }
}

// c2r of TY rows of one (component, layer) from column-contiguous spectra (FFTW's c2r:
// imaginary parts of the DC and Nyquist bins ignored).
template <typename T, int E, bool WV = false>
__global__ void __launch_bounds__(256) k_rowT_inv(const C2<T>* __restrict__ S, T* __restrict__ Hcell,
                                                  int nx, int ny, int nl, int nxh, int TY, FFTPlan px,
                                                  const C2<T>* __restrict__ twx)
{
    extern __shared__ unsigned char smem_raw[];
    C2<T>* tw = reinterpret_cast<C2<T>*>(smem_raw);
    const int B = TY / 2;
    T* re = reinterpret_cast<T*>(tw + nx);
    T* im = re + LPN(B * nx);
    const int nyb = (ny + TY - 1) / TY;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if (L >= nyb * 3 * nl) return;
    const int y0 = (L % nyb) * TY, kc = L / nyb;
    const int k = kc / 3, c = kc - 3 * k;
    const C2<T>* src = S + (size_t)kc * nxh * ny;
    // each thread takes one (kx, row pair) and forms z = A + iB at kx and its mirror
    staged_load<(E < 4 ? E : 4)>(
        nxh * B,
        [&](int idx) {
            const int kx = idx / B, p = idx - kx * B;
            const int ya = y0 + 2 * p, yb = ya + 1;
            return C4<T>{src[(size_t)kx * ny + (ya < ny ? ya : ny - 1)], src[(size_t)kx * ny + (yb < ny ? yb : ny - 1)]};
        },
        [&](int idx, C4<T> v) {
            const int kx = idx / B, p = idx - kx * B;
            const int ya = y0 + 2 * p;
            C2<T> A = ya < ny ? v.a : C2<T>{0, 0}, Bv = ya + 1 < ny ? v.b : C2<T>{0, 0};
            if (kx == 0 || 2 * kx == nx) { A.y = 0; Bv.y = 0; }
            re[LP(p * nx + kx)] = A.x - Bv.y;
            im[LP(p * nx + kx)] = A.y + Bv.x;
            if (kx > 0 && nx - kx != kx) {
                re[LP(p * nx + nx - kx)] = A.x + Bv.y;
                im[LP(p * nx + nx - kx)] = Bv.x - A.y;
            }
        },
        tw, twx, nx);
    __syncthreads();
    if (WV) {
        lds_fft_wave<T, E, true>(re, im, B, px, tw);
        __syncthreads();
    } else {
        lds_fft<T, E, true>(re, im, B, px, tw);
    }
    T* hc = Hcell + ((size_t)c * nl + k) * (size_t)nx * ny;
    for (int idx = threadIdx.x; idx < B * nx; idx += blockDim.x) {
        const int p = idx / nx, x = idx - p * nx;
        const int ya = y0 + 2 * p, yb = ya + 1;
        if (ya < ny) hc[(size_t)ya * nx + x] = re[LP(idx)];
        if (yb < ny) hc[(size_t)yb * nx + x] = im[LP(idx)];
    }
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
    int b0, b1, x0, x1;   // cell range (grain2gridindex) and exchange CSR range
    T hk_s, hk1, hk2, u[3], happ[3];
    // the lane's cells of the first gather round, for the scatter (valid when the grain
    // needed a single round)
    int ca, cb;
    bool one_round;
};

template <typename T>
__device__ __forceinline__ void grain_pre(const GrainData<T>& gd, int g, GrainPre<T>& P)
{
    const int G = gd.G;
    P.b0 = gd.gc_off[g]; P.b1 = gd.gc_off[g + 1];
    P.x0 = gd.x_off[g]; P.x1 = gd.x_off[g + 1];
    P.hk_s = gd.hk_scale[g]; P.hk1 = gd.hk1[g]; P.hk2 = gd.hk2[g];
    P.u[0] = gd.uni[g]; P.u[1] = gd.uni[G + g]; P.u[2] = gd.uni[2 * G + g];
    P.happ[0] = P.happ[1] = P.happ[2] = 0;
    if (gd.happ) {
        P.happ[0] = gd.happ[g]; P.happ[1] = gd.happ[G + g]; P.happ[2] = gd.happ[2 * G + g];
    }
    P.ca = P.cb = 0;
    P.one_round = false;
}

template <typename T, int LPG, typename TC>
__device__ __forceinline__ void field_at(const GrainData<T>& gd, const LayerParams& L, int k, int g, int lane,
                                         const T* __restrict__ mf, const TC* __restrict__ Hcell, size_t CP,
                                         double t, const T* m, T* h, double* tout, GrainPre<T>& P)
{
    const int G = gd.G;
    const bool fsrc = L.src[0].kind == MRMG_SRC_FIELD_TABLE, tsrc = L.src[1].kind != MRMG_SRC_NONE;
    const int b0 = P.b0, b1 = P.b1, x0 = P.x0, x1 = P.x1;
    const T hk_s = P.hk_s, hk1 = P.hk1, hk2 = P.hk2;
    const T u[3] = {P.u[0], P.u[1], P.u[2]};
    T happ[3] = {P.happ[0], P.happ[1], P.happ[2]};
    const T* __restrict__ msc = mf + (size_t)ST_MSC * G;
    // get_magnetostatic gather (field.f90:463-468) and the exchange CSR (field.f90:326-363:
    // intralayer, then the layers above and below, merged per grain)
    T hd[3] = {0, 0, 0}, hx[3] = {0, 0, 0};
    double hs[3] = {0, 0, 0}, ts0 = 0, ts1 = 0;
    if (!fsrc && !tsrc) {
        T s0 = 0, s1 = 0, s2 = 0, e0 = 0, e1 = 0, e2 = 0;
        const int nc = b1 - b0, ne = x1 - x0;
        const int ncr = (nc + 2 * LPG - 1) / (2 * LPG), ner = (ne + 2 * LPG - 1) / (2 * LPG);
        const int nr = ncr > ner ? ncr : ner;
        P.one_round = ncr <= 1;
        for (int r = 0; r < nr; r++) {
            const int qa = b0 + lane + 2 * r * LPG, qb = qa + LPG;
            const int ja = x0 + lane + 2 * r * LPG, jb = ja + LPG;
            const bool hqa = qa < b1, hqb = qb < b1, hja = ja < x1, hjb = jb < x1;
            const int ca = gd.gc_idx[hqa ? qa : b0], cb = gd.gc_idx[hqb ? qb : b0];
            if (r == 0) { P.ca = ca; P.cb = cb; }
            const int na = gd.x_nb[hja ? ja : x0], nbb = gd.x_nb[hjb ? jb : x0];
            const T wa = gd.x_w[hja ? ja : x0], wb = gd.x_w[hjb ? jb : x0];
            T c0 = 0, c1 = 0, c2 = 0, d0 = 0, d1 = 0, d2 = 0;
            if (Hcell) {
                c0 = (T)Hcell[ca]; c1 = (T)Hcell[CP + ca]; c2 = (T)Hcell[2 * CP + ca];
                d0 = (T)Hcell[cb]; d1 = (T)Hcell[CP + cb]; d2 = (T)Hcell[2 * CP + cb];
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
        if (Hcell) {
            const T cnt = (T)nc;
            hd[0] = lane_sum<LPG>(s0) / cnt;
            hd[1] = lane_sum<LPG>(s1) / cnt;
            hd[2] = lane_sum<LPG>(s2) / cnt;
        }
        hx[0] = lane_sum<LPG>(e0);
        hx[1] = lane_sum<LPG>(e1);
        hx[2] = lane_sum<LPG>(e2);
    } else {
        T s0 = 0, s1 = 0, s2 = 0;
        for (int q = b0 + lane; q < b1; q += LPG) {
            const int ci = gd.gc_idx[q];
            if (Hcell) {
                s0 += (T)Hcell[ci];
                s1 += (T)Hcell[CP + ci];
                s2 += (T)Hcell[2 * CP + ci];
            }
            const int i = ci % gd.ax.nx, j = (ci / gd.ax.nx) % gd.ax.ny;
            if (fsrc)
#pragma unroll
                for (int c = 0; c < 3; c++) hs[c] += src_cell(gd.ax, L.src[0], k, 0, 0, i, j, c, 3);
            if (tsrc) {
                ts0 += src_cell(gd.ax, L.src[1], k, 1, 0, i, j, 0, 1);
                ts1 += src_cell(gd.ax, L.src[1], k, 1, 1, i, j, 0, 1);
            }
        }
        const T cnt = (T)(b1 - b0);
        if (Hcell) {
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
    const int b0 = gd.gc_off[g], b1 = gd.gc_off[g + 1];
    double ts0 = 0, ts1 = 0;
    for (int q = b0 + lane; q < b1; q += LPG) {
        const int ci = gd.gc_idx[q];
        const int i = ci % gd.ax.nx, j = (ci / gd.ax.nx) % gd.ax.ny;
        ts0 += src_cell(gd.ax, L.src[1], k, 1, 0, i, j, 0, 1);
        ts1 += src_cell(gd.ax, L.src[1], k, 1, 1, i, j, 0, 1);
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

// get_mag (data.f90:344-353): M = Ms*ms_scale*m on every cell of grain g, lanes split the
// grain's cells.  Non-magnetic cells are never written and stay 0.
template <typename T, int LPG, typename TC>
__device__ __forceinline__ void scatter_cells(const GrainData<T>& gd, TC* __restrict__ Mcell, size_t CP, int g,
                                              int lane, T f, const T* m, const GrainPre<T>* P = nullptr)
{
    const TC v0 = (TC)(f * m[0]), v1 = (TC)(f * m[1]), v2 = (TC)(f * m[2]);
    if (P && P->one_round) {
        // the cells this lane gathered: b0+lane and b0+lane+LPG
        const int qa = P->b0 + lane, qb = qa + LPG;
        if (qa < P->b1) { Mcell[P->ca] = v0; Mcell[CP + P->ca] = v1; Mcell[2 * CP + P->ca] = v2; }
        if (qb < P->b1) { Mcell[P->cb] = v0; Mcell[CP + P->cb] = v1; Mcell[2 * CP + P->cb] = v2; }
        return;
    }
    for (int q = gd.gc_off[g] + lane; q < gd.gc_off[g + 1]; q += LPG) {
        const int ci = gd.gc_idx[q];
        Mcell[ci] = v0;
        Mcell[CP + ci] = v1;
        Mcell[2 * CP + ci] = v2;
    }
}

// recompute: layers whose Ms follows a moving temperature take it at the axes' first time
template <typename T, int LPG, typename TC>
__global__ void __launch_bounds__(256) k_scatter(GrainData<T> gd, T* __restrict__ m, TC* __restrict__ Mcell,
                                                 size_t CP, int recompute)
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
    scatter_cells<T, LPG, TC>(gd, Mcell, CP, g, lane, f, mm);
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
                                               const TC* __restrict__ Hcell, TC* __restrict__ Mcell, size_t CP,
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
        field_at<T, LPG, TC>(gd, L, k, g, lane, mf, Hcell, CP, a.t, mfield, h, tev, P);
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
                    philox_normals3(a.seed, a.step, g, n);
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
    // the new magnetisation of the grain onto its cells, for the next field evaluation
    // (every lane of the grain holds mo and fsc)
    if (MODE != GM_FIELD1 && Mcell && act) scatter_cells<T, LPG, TC>(gd, Mcell, CP, g, lane, fsc, mo, &P);
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
__global__ void __launch_bounds__(256) k_rb_plane(const TC* __restrict__ Mcell, size_t CP, int nxy, int nl,
                                                  double d0, double d1, double d2, RbWeights wk,
                                                  double* __restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    double s = 0.0;
    for (int k = 0; k < nl; k++) {
        const size_t ci = (size_t)k * nxy + i;
        s += wk.w[k] * (d0 * (double)Mcell[ci] + d1 * (double)Mcell[CP + ci] + d2 * (double)Mcell[2 * CP + ci]);
    }
    P[i] = s;
}
__global__ void __launch_bounds__(256) k_rb_scan(const double* __restrict__ P, int nx, int ny, double dx, double dy,
                                                 double sx, double sy, const double* __restrict__ pos, int n_pos,
                                                 double* __restrict__ out)
{
    extern __shared__ double gsh[];   // nx + ny
    __shared__ double red[4];
    const int p = blockIdx.x;
    if (p >= n_pos) return;
    const double x0 = pos[p], y0 = pos[n_pos + p];
    const double Lx = nx * dx, Ly = ny * dy;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        double d = (i + 0.5) * dx - x0;
        d -= Lx * rint(d / Lx);
        gsh[i] = exp(-d * d / (2.0 * sx * sx));
    }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        double d = (j + 0.5) * dy - y0;
        d -= Ly * rint(d / Ly);
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

// ------------------------------------------------------------------------- tensor kernels
// Newell's volume-averaged tensor of two blocks (tensor_element.f90:13-102): n -= t.
__device__ __forceinline__ double dsignv(double a, double b) { return signbit(b) ? -fabs(a) : fabs(a); }
__device__ __forceinline__ double dmaxv(double a, double b) { return a > b ? a : b; }

__device__ double tg1(double x, double y, double z, double d)
{
    return x * y * z * atan(y * z / dsignv(dmaxv(1.e-300, fabs(x) * d), x))
         + 0.5 * y * (z * z - x * x) * log(dmaxv(1.e-300, d - y))
         + 0.5 * z * (y * y - x * x) * log(dmaxv(1.e-300, d - z))
         + (y * y + z * z - 2.0 * x * x) * d / 6.0;
}
__device__ double tg2(double z, double x, double y, double d)
{
    return -x * y * z * log(dmaxv(1.e-300, d + z))
         + y * (y * y - 3.0 * z * z) * log(dmaxv(1.e-300, d + x)) / 6.0
         + x * (x * x - 3.0 * z * z) * log(dmaxv(1.e-300, d + y)) / 6.0
         + 0.5 * x * x * z * atan(y * z / dsignv(dmaxv(1.e-300, fabs(x) * d), x))
         + 0.5 * y * y * z * atan(x * z / dsignv(dmaxv(1.e-300, fabs(y) * d), y))
         + z * z * z * atan(x * y / dsignv(dmaxv(1.e-300, fabs(z) * d), z)) / 6.0 + x * y * d / 3.0;
}

__device__ void van(double rx, double ry, double rz, double xs, double ys, double zs, double xo, double zo, double* n)
{
    const double dxo[2] = {-0.5 * xo, 0.5 * xo}, dyo[2] = {-0.5 * xo, 0.5 * xo}, dzo[2] = {-0.5 * zo, 0.5 * zo};
    const double dxs[2] = {-0.5 * xs, 0.5 * xs}, dys[2] = {-0.5 * ys, 0.5 * ys}, dzs[2] = {-0.5 * zs, 0.5 * zs};
    double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i6 = 0; i6 < 2; i6++)
    for (int i5 = 0; i5 < 2; i5++)
    for (int i4 = 0; i4 < 2; i4++)
    for (int i3 = 0; i3 < 2; i3++) {
        const double dx = -rx + dxo[i3] - dxs[i6];
        for (int i2 = 0; i2 < 2; i2++) {
            const double dy = -ry + dyo[i2] - dys[i5];
            for (int i1 = 0; i1 < 2; i1++) {
                const double dz = -rz + dzo[i1] - dzs[i4];
                const double d = sqrt(dx * dx + dy * dy + dz * dz);
                if (d != 0.0) {
                    const double s = ((i1 + i2 + i3 + i4 + i5 + i6 + 5) & 1) ? -1.0 : 1.0;
                    double x;
                    t[0] = t[0] + s * tg1(dx, dy, dz, d);
                    t[4] = t[4] + s * tg1(dy, dz, dx, d);
                    t[8] = t[8] + s * tg1(dz, dx, dy, d);
                    x = s * tg2(dx, dy, dz, d); t[5] = t[5] + x; t[7] = t[7] + x;
                    x = s * tg2(dy, dz, dx, d); t[2] = t[2] + x; t[6] = t[6] + x;
                    x = s * tg2(dz, dx, dy, d); t[1] = t[1] + x; t[3] = t[3] + x;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) n[i] = n[i] - t[i];
}

// plane_to_plane for one layer pair: one thread per output cell; the offsets (i,j) that
// alias onto it (i = +-nx/2, j = +-ny/2 for even sizes) are summed in the reference's
// j-outer, i-inner order (tensor_element.f90:233-259).  Writes the normalised block.
// mode 0: blk = t / norm (a direct term alone); mode 1: blk = t (a direct term that images
// follow); mode 2: blk += f t with f = fxy for the x and y rows (a = c % 3 < 2), fz for z
// (an image term, tensor.f90:163-181).
__global__ void k_plane_to_plane(int nx, int ny, double dx, double dy, double dzs, double dzo, double rz,
                                 int nimg, double norm, double* __restrict__ blk, int mode = 0, double fxy = 0.0,
                                 double fz = 0.0)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nx * ny) return;
    const int oi = p % nx, oj = p / nx;
    int is[2], js[2], ni = 0, nj = 0;
    if (nx % 2 == 0 && oi == nx / 2) { is[0] = -nx / 2; is[1] = nx / 2; ni = 2; }
    else { is[0] = (oi <= nx / 2) ? -oi : nx - oi; ni = 1; }
    if (ny % 2 == 0 && oj == ny / 2) { js[0] = -ny / 2; js[1] = ny / 2; nj = 2; }
    else { js[0] = (oj <= ny / 2) ? -oj : ny - oj; nj = 1; }
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nj; b++) {
        const int j = js[b];
        const int jq = j / ((ny + 1) / 2);
        const int yp0 = -nimg * (1 - (jq > 0 ? jq : 0)), yp1 = nimg * (1 + (jq < 0 ? jq : 0));
        for (int a = 0; a < ni; a++) {
            const int i = is[a];
            const int iq = i / ((nx + 1) / 2);
            const int xp0 = -nimg * (1 - (iq > 0 ? iq : 0)), xp1 = nimg * (1 + (iq < 0 ? iq : 0));
            double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int nyi = yp0; nyi <= yp1; nyi++)
                for (int nxi = xp0; nxi <= xp1; nxi++)
                    van(i * dx + nxi * dx * nx, j * dy + nyi * dy * ny, rz, dx, dy, dzs, dx, dzo, t);
#pragma unroll
            for (int c = 0; c < 9; c++) acc[c] = acc[c] + t[c];
        }
    }
    const size_t plane = (size_t)nx * ny;
    if (mode == 0) {
#pragma unroll
        for (int c = 0; c < 9; c++) blk[p + plane * c] = acc[c] / norm;
    } else if (mode == 1) {
#pragma unroll
        for (int c = 0; c < 9; c++) blk[p + plane * c] = acc[c];
    } else {
#pragma unroll
        for (int c = 0; c < 9; c++) blk[p + plane * c] = blk[p + plane * c] + ((c % 3) < 2 ? fxy : fz) * acc[c];
    }
}

// the normalisation of a layer-pair block with images, n2d / (dx dy dz_ko nx ny)
__global__ void k_tensor_div(double* __restrict__ blk, size_t n, double norm)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) blk[i] = blk[i] / norm;
}

// Tensor FFT in FP64, rows: one workgroup per (row pair, plane) of P real planes.
template <int E>
__global__ void __launch_bounds__(256) k_tensor_rows(const double* __restrict__ R, C2<double>* __restrict__ TC,
                                                     int nx, int ny, int nxh, FFTPlan px,
                                                     const C2<double>* __restrict__ twx)
{
    extern __shared__ unsigned char smem_raw[];
    C2<double>* tw = reinterpret_cast<C2<double>*>(smem_raw);
    double* re = reinterpret_cast<double*>(tw + nx);
    double* im = re + LPN(nx);
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = twx[i];
    const int y0 = 2 * blockIdx.x, y1 = y0 + 1;
    const size_t pl = blockIdx.y;
    const bool has1 = y1 < ny;
    const double* r = R + pl * nx * ny;
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        re[LP(x)] = r[(size_t)y0 * nx + x];
        im[LP(x)] = has1 ? r[(size_t)y1 * nx + x] : 0.0;
    }
    __syncthreads();
    lds_fft<double, E, false>(re, im, 1, px, tw);
    C2<double>* out = TC + pl * nxh * ny;
    for (int kx = threadIdx.x; kx < nxh; kx += blockDim.x) {
        const int kn = kx == 0 ? 0 : nx - kx;
        const double zr = re[LP(kx)], zi = im[LP(kx)], nr = re[LP(kn)], ni = im[LP(kn)];
        out[(size_t)y0 * nxh + kx] = {(zr + nr) * 0.5, (zi - ni) * 0.5};
        if (has1) out[(size_t)y1 * nxh + kx] = {(zi + ni) * 0.5, (nr - zr) * 0.5};
    }
}

// Tensor FFT in FP64, columns in place: one workgroup per (kx, plane).
template <int E>
__global__ void __launch_bounds__(256) k_tensor_cols(C2<double>* __restrict__ TC, int ny, int nxh, FFTPlan py,
                                                     const C2<double>* __restrict__ twy)
{
    extern __shared__ unsigned char smem_raw[];
    C2<double>* tw = reinterpret_cast<C2<double>*>(smem_raw);
    double* re = reinterpret_cast<double*>(tw + ny);
    double* im = re + LPN(ny);
    for (int i = threadIdx.x; i < ny; i += blockDim.x) tw[i] = twy[i];
    const int kx = blockIdx.x;
    C2<double>* col = TC + (size_t)blockIdx.y * nxh * ny + kx;
    for (int y = threadIdx.x; y < ny; y += blockDim.x) {
        const C2<double> v = col[(size_t)y * nxh];
        re[LP(y)] = v.x;
        im[LP(y)] = v.y;
    }
    __syncthreads();
    lds_fft<double, E, false>(re, im, 1, py, tw);
    for (int y = threadIdx.x; y < ny; y += blockDim.x) col[(size_t)y * nxh] = {re[LP(y)], im[LP(y)]};
}

// Regroup the tensor spectra into Tspec[kx][ko][ks][slot][ky] and round to T.
template <typename T>
__global__ void k_tensor_pack(const C2<double>* __restrict__ TC, C2<T>* __restrict__ Tspec, int ny, int nxh,
                              int nl, int nslot, const int* __restrict__ slot2c)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nxh * nl * nl * nslot * ny;
    if (i >= total) return;
    size_t r = i;
    const int ky = r % ny; r /= ny;
    const int s = r % nslot; r /= nslot;
    const int ks = r % nl; r /= nl;
    const int ko = r % nl; r /= nl;
    const int kx = (int)r;
    const int c = slot2c[s];
    // plane (c, ko, ks) of the (nx,ny,9,nl,nl) tensor
    const size_t plane = (size_t)c + 9 * ((size_t)ko + (size_t)nl * ks);
    const C2<double> v = TC[plane * nxh * ny + (size_t)ky * nxh + kx];
    Tspec[i] = C2<T>{(T)v.x, (T)v.y};
}

// ------------------------------------------------------------------------- host side
// Stockham passes of the largest radices up to maxr: fewer, heavier passes for large
// batches, more and lighter ones when a workgroup's batch has few butterflies per pass
static bool factorize(int n, FFTPlan& p, int maxr = 16)
{
    p.n = n;
    p.nfac = 0;
    int m = n;
    while (maxr >= 16 && m % 16 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 16; m /= 16; }
    while (maxr >= 8 && m % 8 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 8; m /= 8; }
    while (m % 4 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 4; m /= 4; }
    while (m % 2 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 3; m /= 3; }
    while (m % 5 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 5; m /= 5; }
    plan_passes(p);
    return m == 1;
}

template <typename T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void swap_with(DBuf& o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count)
    {
        release();
        if (count == 0) return MRMG_OK;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return set_err(MRMG_ERR_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
        }
        n = count;
        return MRMG_OK;
    }
    int upload(const T* h, size_t count, hipStream_t s)
    {
        RCHK(alloc(count));
        if (count) HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return MRMG_OK;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct HostLayer {
    bool set = false;
    int ng = 0, ncell = 0;
    std::vector<int> grid2grain, g2gi, g2g;
    std::vector<double> vol, ms, hk, uni, cub, ms_scale, alpha_scale, hk_scale, temp, happ;
    std::vector<int> hxi, hxg, uhxi, uhxg, lhxi, lhxg;
    std::vector<double> hx, uhx, lhx;
    double alpha = 1, gamma = 1.76e7, thm_alpha = 1, floor_ms = 0.01;
    int thermal = 0;
    bool has_uni = false, has_cub = false, has_happ = false, has_hx = false, has_uhx = false, has_lhx = false;
    mrmg_ufield uf{};
    bool has_uf = false;
    // LLB and temperature scaling: Tc, lambdaMs, S; use_mag = use_mag_p with tc_nom <= 1e10
    std::vector<double> tc, lambda_ms;
    double spin = 0.5, msexp = -1.0, hkexp = -1.0;
    bool use_mag = false, tscale = false;
    std::vector<double> tab_t[3], tab_s[3];   // READ_SCALE tables: 0 Ms, 1 alpha, 2 Hk
    // moving applied sources: slot 0 field table, slot 1 temperature (table or super-Gaussian)
    mrmg_source src[2] = {};
    std::vector<double> src_x[2], src_y[2], src_v[2];
    std::vector<double> m;     // (3,ng) host copy used before the device state exists
    std::vector<double> mag;   // (ng) |m|, likewise
};

// get_thermal's standard deviation per grain (field.f90:407-423): the LLB form (no_ms_p),
// else the |m| form when |m| != 1, else the ms_scale form
static double thermal_sigma(const HostLayer& H, int n, double mag, bool no_ms)
{
    if (!H.thermal) return 0.0;
    const double tsigma = sqrt(2.0 * 1.380662e-16 * H.thm_alpha / H.gamma);
    double x;
    if (no_ms) {
        x = H.alpha_scale[n] * H.temp[n] / (H.ms[n] * H.vol[n]);
    } else if (mag != 1.0) {
        x = H.alpha_scale[n] * H.temp[n] / (H.ms[n] * H.vol[n] * std::max(H.floor_ms, fabs(mag)));
    } else {
        const double msc = H.use_mag ? 1.0 : H.ms_scale[n];
        x = msc > 0.0 ? H.alpha_scale[n] * H.temp[n] / (H.ms[n] * H.vol[n] * std::max(H.floor_ms, msc)) : 0.0;
    }
    return sqrt(x) * tsigma;
}

// the Ms scale the field sees for |m| = mag (find_temp_scaling's use_mag_p form)
static double field_ms_scale(const HostLayer& H, int n, double mag)
{
    return H.use_mag ? std::min(1.0, std::max(0.0, fabs(mag))) : H.ms_scale[n];
}

// set_fld's normalisation of a time function (applied.f90:276-310): duty and moddepth
// clamped to [0, 1]; a function with amp == 0 or gamma == 0 becomes the constant
// amp * f(0) + dc; equal ramp ends fold into the scale
static TFun norm_tfun(int func, double amp, double gamma, double phase, double param, double duty, double moddepth,
                      double dc, double& s0, double& s1, double& scale)
{
    TFun f{};
    f.func = func; f.amp = amp; f.gamma = gamma; f.phase = phase; f.param = param; f.dc = dc;
    f.duty = std::min(1.0, std::max(0.0, duty));
    f.moddepth = std::min(1.0, std::max(0.0, moddepth));
    if (f.func != MRMG_FN_RAMP && (f.amp == 0.0 || f.gamma == 0.0)) {
        s0 = f.amp * funct_eval(f.func, 0.0, f.gamma, f.phase, f.param, f.duty, f.moddepth) + f.dc;
        s1 = s0;
        f.func = MRMG_FN_RAMP;
    }
    if (s0 == s1) { scale = scale * s0; s0 = 1.0; s1 = 1.0; }
    return f;
}

// checks of a time function's arguments
static int check_tfun(int func, double t_start, double t_end, double amp, double gamma, double param)
{
    if (func < MRMG_FN_RAMP || func > MRMG_FN_STE) return set_err(MRMG_ERR_ARG, "time function %d unknown", func);
    if (func == MRMG_FN_RAMP && t_end == t_start) return set_err(MRMG_ERR_ARG, "t_end == t_start");
    if ((func == MRMG_FN_NEW || func == MRMG_FN_STE) && amp != 0.0 && gamma != 0.0 && param == 0.0)
        return set_err(MRMG_ERR_ARG, "NEW and STE need param != 0");
    return MRMG_OK;
}

struct EngineBase {
    virtual ~EngineBase() {}
    virtual int build(std::vector<HostLayer>& L) = 0;
    virtual int set_tensor_planes(const double* d_real) = 0;
    virtual int set_state(int k, const double* m) = 0;
    virtual int get_state(int k, double* m, double* h) = 0;
    virtual int set_mag(int k, const double* mag) = 0;
    virtual int get_mag(int k, double* mag) = 0;
    virtual int update_ufield(std::vector<HostLayer>& L) = 0;
    virtual int step(const mrmg_step_params* p, const double* noise, mrmg_diag* diag) = 0;
    virtual int demag_field(double* h_cell, double* h_grain) = 0;
    virtual size_t footprint() const = 0;
    virtual bool has_tensor() const = 0;
    virtual int profile(int enable) = 0;
    virtual int tune(const char* key, int v) = 0;
    virtual int get_tune(const char* key, int* v) = 0;
    virtual int kernel_stats(mrmg_kstat* out, int max, int* n) = 0;
    virtual int readback(const mrmg_sensor* s, int n_pos, const double* xpos, const double* ypos, double* out) = 0;
    virtual int switch_arm(int enable, const double* dir, double thr) = 0;
    virtual int switch_read(int k, int64_t* step, double* t, double* hap) = 0;
    virtual int time_class(int cls, int reps, double* ms) = 0;
};

struct mrmg_ctx {
    int nx = 0, ny = 0, nl = 0, nxh = 0, precision = 0, demag = 1;
    double dx = 0, dy = 0;
    std::vector<double> dz, il;
    hipStream_t stream = nullptr;
    FFTPlan px{}, py{};
    std::vector<HostLayer> layers;
    bool dirty = true;
    std::unique_ptr<EngineBase> eng;
};

static std::vector<C2<double>> twiddle_table(int n)
{
    std::vector<C2<double>> w(n);
    for (int k = 0; k < n; k++) {
        // exp(-2 pi i k/n), angle reduced by symmetry before the libm call
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        w[k] = {(double)cosl(a), (double)(-sinl(a))};
    }
    return w;
}

// E: complex elements per thread of one pass, B*N <= E*threads, rounded up to 4..64
static int pick_e(int batch, const FFTPlan& p, int threads)
{
    const int need = (batch * p.n + threads - 1) / threads;
    for (int e = 4; e <= 64; e *= 2)
        if (need <= e) return e;
    return -1;
}

#define LAUNCH_E(EV, KERNEL, ...)                                                                     \
    do {                                                                                              \
        switch (EV) {                                                                                 \
        case 4: KERNEL<4>(__VA_ARGS__); break;                                                        \
        case 8: KERNEL<8>(__VA_ARGS__); break;                                                        \
        case 16: KERNEL<16>(__VA_ARGS__); break;                                                      \
        case 32: KERNEL<32>(__VA_ARGS__); break;                                                      \
        default: KERNEL<64>(__VA_ARGS__); break;                                                      \
        }                                                                                             \
    } while (0)

// T: the type of the demag path (cell fields, spectra, tensor, FFT).  The grain state, the
// grain constants and the field assembly are FP64 in both forms, as in the reference's
// build with single-precision FFTs (ffttr = 4, data.f90:326).
template <typename T>
struct Engine : EngineBase {
    using S = double;
    mrmg_ctx* c;
    int G = 0, Ctot = 0, nslot = 0;
    std::vector<int> off;
    // grain data
    DBuf<int> layer_of, gc_off, gc_idx, cell2grain;
    DBuf<int> x_off, x_nb;
    DBuf<S> x_w;
    DBuf<S> ms, msf, ms_scale, alpha_scale, hk_scale, hk1, hk2, uni, cub, happ, sigma, sigma_llb, vol;
    DBuf<double> tc, lam, temp, tabs;
    DBuf<LayerParams> lp;
    // state, ST_ROWS rows of G (mc: corrector output of the adaptive path until the step
    // is accepted)
    DBuf<S> ma, mb, mc, dm, hthm, h;
    // |m| on the device differs from the host copy the LLG thermal variance was built from
    bool mag_dirty = false;
    // moving sources: node positions and values of every table, per-stage axis tables,
    // and the grain temperature of the last field evaluation (adaptive path)
    DBuf<double> src_pos, src_vals, ax_xa, ax_ya, ax_yb, tlast;
    DBuf<int> ax_xi, ax_yi;
    bool any_src = false;
    // demag
    DBuf<C2<T>> S1, S2, Tspec, twx, twy;
    DBuf<T> Hcell, Mcell;
    DBuf<int> slot2c;
    // launch shapes (mrmg_tune): threads of the row and column workgroups, observation-layer
    // groups per column workgroup, lanes per grain in the grain kernels
    int row_threads = 128, col_threads = 512, kg = 1, lpg = 8;
    // split column path (column-contiguous spectra): rows per row workgroup, columns per
    // column workgroup, and whether it is in use (forced by col_path or by the grid size)
    bool split = false, fused_ok = false, split_ok = false;
    bool tl = false;   // fused column kernel over the split path's column-contiguous spectra
    bool wv = false;   // wave-local FFTs in the transposing row kernels and the path-2 column kernel
    int TYr = 8, CBc = 1, maxb_rowT = 4, maxb_colT = 4;
    // FFT plans of the step's row and column transforms and their largest radix
    FFTPlan prow{}, pcol{};
    int row_radix = 16, col_radix = 16;
    size_t lds_rowT = 0, lds_colT = 0;
    DBuf<double> noise_buf, diag_part, stats;
    // switching record and readback buffers (allocated on first use)
    DBuf<long long> sw_step;
    DBuf<double> sw_t, sw_h, rb_plane, rb_pos, rb_out;
    bool sw_on = false;
    double sw_dir[3] = {0, 0, 0}, sw_thr = 0;
    // instrumentation: event pairs of the current mrmg_step call, per-class totals
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_class;
    size_t ev_used = 0;
    double kt_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int kt_n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool tensor_ok = false, has_happ = false;
    bool cur_llb = false;   // the step in flight uses push_llb (selects the thermal variance)
    int maxb_row = 1, maxb_col = 1;
    size_t lds_row = 0, lds_col = 0;
    std::vector<LayerParams> hlp;

    explicit Engine(mrmg_ctx* ctx) : c(ctx) {}
    ~Engine() override
    {
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    }

    // kernel classes of the profile
    enum { K_ROW_FWD = 0, K_COL = 1, K_ROW_INV = 2, K_GRAIN1 = 3, K_GRAIN2 = 4, K_COLF = 5, K_PROD = 6, K_COLI = 7,
           K_NCLASS = 8 };
    int ev_begin(int cls)
    {
        if (!prof) return MRMG_OK;
        if (ev_used + 2 > ev_pool.size()) {
            for (int i = 0; i < 2; i++) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                ev_pool.push_back(e);
            }
            ev_class.resize(ev_pool.size() / 2);
        }
        ev_class[ev_used / 2] = cls;
        HIPCHK(hipEventRecord(ev_pool[ev_used], c->stream));
        return MRMG_OK;
    }
    int ev_end()
    {
        if (!prof) return MRMG_OK;
        HIPCHK(hipEventRecord(ev_pool[ev_used + 1], c->stream));
        ev_used += 2;
        return MRMG_OK;
    }
    // after the stream is synchronised: fold the recorded pairs into the totals
    int ev_collect()
    {
        for (size_t i = 0; i < ev_used; i += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev_pool[i], ev_pool[i + 1]));
            kt_ms[ev_class[i / 2]] += ms;
            kt_n[ev_class[i / 2]] += 1;
        }
        ev_used = 0;
        return MRMG_OK;
    }

    size_t footprint() const override
    {
        size_t b = 0;
        b += layer_of.bytes() + gc_off.bytes() + gc_idx.bytes() + cell2grain.bytes();
        b += x_off.bytes() + x_nb.bytes() + x_w.bytes();
        b += ms.bytes() + msf.bytes() + ms_scale.bytes() + alpha_scale.bytes() + hk_scale.bytes() + hk1.bytes() +
             hk2.bytes() + uni.bytes() + cub.bytes() + happ.bytes() + sigma.bytes() + sigma_llb.bytes() + vol.bytes() +
             lp.bytes() + tc.bytes() + lam.bytes() + temp.bytes();
        b += ma.bytes() + mb.bytes() + mc.bytes() + dm.bytes() + hthm.bytes() + h.bytes();
        b += S1.bytes() + S2.bytes() + Tspec.bytes() + twx.bytes() + twy.bytes() + Hcell.bytes() + Mcell.bytes() +
             slot2c.bytes();
        b += noise_buf.bytes() + diag_part.bytes();
        return b;
    }
    bool has_tensor() const override { return tensor_ok; }

    int init_fft()
    {
        const int nx = c->nx, ny = c->ny, nxh = c->nxh, nl = c->nl;
        auto wx = twiddle_table(nx), wy = twiddle_table(ny);
        std::vector<C2<T>> tx(nx), ty(ny);
        for (int i = 0; i < nx; i++) tx[i] = {(T)wx[i].x, (T)wx[i].y};
        for (int i = 0; i < ny; i++) ty[i] = {(T)wy[i].x, (T)wy[i].y};
        RCHK(twx.upload(tx.data(), nx, c->stream));
        RCHK(twy.upload(ty.data(), ny, c->stream));
        prow = c->px;
        pcol = c->py;
        // fused column path: all 3*n_lay columns of one kx in one workgroup's LDS
        lds_row = sizeof(C2<T>) * nx + sizeof(T) * 2 * LPN(3 * nx);
        lds_col = sizeof(C2<T>) * ny + sizeof(T) * 2 * LPN(3 * nl * ny);
        fused_ok = lds_row <= 160 * 1024 && lds_col <= 160 * 1024 && pick_e(3, c->px, 256) > 0 &&
                   pick_e(3 * nl, c->py, col_threads) > 0 && nl * ny <= 4 * col_threads;
        // split path: TY rows per row workgroup, CB columns per column workgroup
        const size_t lim = 160 * 1024;
        split_ok = false;
        for (int ty = 8; ty >= 2; ty /= 2) {
            lds_rowT = sizeof(C2<T>) * nx + sizeof(T) * 2 * LPN((ty / 2) * nx);
            if (lds_rowT <= lim && pick_e(ty / 2, c->px, 256) > 0 && (lds_rowT <= 64 * 1024 || ty == 2)) {
                TYr = ty;
                split_ok = true;
                break;
            }
        }
        CBc = std::max(1, std::min(16, 2048 / ny));
        CBc = std::min(CBc, 3 * nl * nxh);
        lds_colT = sizeof(T) * 2 * LPN(CBc * ny);
        split_ok = split_ok && lds_colT <= lim && pick_e(CBc, c->py, 256) > 0;
        if (!fused_ok && !split_ok)
            return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d: a row or column FFT does not fit one workgroup", nx, ny,
                           nl);
        split = !fused_ok;
        RCHK(shapes());
        RCHK(lds_attr());
        RCHK(S1.alloc((size_t)nl * 3 * ny * nxh));
        RCHK(S2.alloc((size_t)nl * 3 * ny * nxh));
        RCHK(Hcell.alloc((size_t)3 * nl * nx * ny));
        HIPCHK(hipMemsetAsync(Hcell.p, 0, Hcell.bytes(), c->stream));
        RCHK(Mcell.alloc((size_t)3 * nl * nx * ny));
        HIPCHK(hipMemsetAsync(Mcell.p, 0, Mcell.bytes(), c->stream));
        kg = 1;
        return MRMG_OK;
    }

    // derived launch parameters of the current shapes
    int shapes()
    {
        const int nl = c->nl, ny = c->ny;
        // wave-local transforms: a lane's bound covers its wave's run of ceil(B / waves)
        auto row_e = [&]() {
            const int nw = row_threads / 64, b = TYr / 2;
            return wv ? pick_e((b + nw - 1) / nw, c->px, 64) : pick_e(b, c->px, row_threads);
        };
        if (split || tl) {
            maxb_rowT = row_e();
            if (maxb_rowT < 0)
                return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d: row FFT too large for %d threads", c->nx, ny, nl,
                               row_threads);
        }
        if (split) {
            maxb_colT = pick_e(CBc, c->py, 256);
            if (maxb_rowT < 0 || maxb_colT < 0)
                return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d: split-path FFT too large for %d threads", c->nx, ny,
                               nl, row_threads);
            return MRMG_OK;
        }
        maxb_row = pick_e(3, c->px, row_threads);
        const int ncw = col_threads / 64;
        maxb_col = tl && wv ? pick_e((3 * nl + ncw - 1) / ncw, c->py, 64) : pick_e(3 * nl, c->py, col_threads);
        if (maxb_row < 0 || maxb_col < 0)
            return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d too large for one-workgroup LDS FFTs (threads %d/%d)",
                           c->nx, ny, nl, row_threads, col_threads);
        const int per = (nl + kg - 1) / kg;
        if (col_threads > 512 && wv)
            return set_err(MRMG_ERR_UNSUPPORTED, "col_threads 1024 runs without wave-local FFTs");
        if (col_threads > 512 && maxb_col > 16)   // more elements per lane spill at 128 VGPRs
            return set_err(MRMG_ERR_UNSUPPORTED, "col_threads 1024 is for columns of up to 16 elements per lane");
        if (per * ny > (col_threads > 512 ? 2 : 4) * col_threads)
            return set_err(MRMG_ERR_UNSUPPORTED, "column workgroup of %d threads cannot hold %d products", col_threads,
                           per * ny);
        return MRMG_OK;
    }
    int tune(const char* key, int v) override
    {
        std::string k(key ? key : "");
        int old;
        if (k == "row_threads") {
            if (v != 64 && v != 128 && v != 256) return set_err(MRMG_ERR_ARG, "row_threads must be 64, 128 or 256");
            old = row_threads; row_threads = v;
            if (shapes() != MRMG_OK) { row_threads = old; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "col_threads") {
            if (v != 256 && v != 512 && v != 1024) return set_err(MRMG_ERR_ARG, "col_threads must be 256, 512 or 1024");
            if (v > 512 && nslot != 6)
                return set_err(MRMG_ERR_UNSUPPORTED, "col_threads 1024 needs the symmetric (6-slot) tensor");
            old = col_threads; col_threads = v;
            if (shapes() != MRMG_OK) { col_threads = old; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "col_groups") {
            if (v < 1 || v > c->nl) return set_err(MRMG_ERR_ARG, "col_groups must be in 1..n_lay");
            old = kg; kg = v;
            if (shapes() != MRMG_OK) { kg = old; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "lanes_per_grain") {
            if (v != 4 && v != 8 && v != 16) return set_err(MRMG_ERR_ARG, "lanes_per_grain must be 4, 8 or 16");
            lpg = v;
        } else if (k == "col_path") {
            if (v < 0 || v > 2)
                return set_err(MRMG_ERR_ARG, "col_path must be 0 (fused), 1 (split) or 2 (fused, column-contiguous)");
            if (v != 1 && !fused_ok) return set_err(MRMG_ERR_UNSUPPORTED, "the fused column kernel does not fit this grid");
            if (v != 0 && !split_ok) return set_err(MRMG_ERR_UNSUPPORTED, "the transposing row kernels do not fit this grid");
            const bool os = split, ot = tl;
            split = v == 1;
            tl = v == 2;
            if (shapes() != MRMG_OK) { split = os; tl = ot; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "wave_fft") {
            if (v != 0 && v != 1) return set_err(MRMG_ERR_ARG, "wave_fft must be 0 or 1");
            const bool ow = wv;
            wv = v == 1;
            if (shapes() != MRMG_OK) { wv = ow; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "row_radix" || k == "col_radix") {
            if (v != 4 && v != 8 && v != 16) return set_err(MRMG_ERR_ARG, "%s must be 4, 8 or 16", k.c_str());
            FFTPlan q;
            if (!factorize(k == "row_radix" ? c->nx : c->ny, q, v) || q.nfac > MRMG_MAX_FAC)
                return set_err(MRMG_ERR_UNSUPPORTED, "%s %d: no plan", k.c_str(), v);
            if (k == "row_radix") { prow = q; row_radix = v; }
            else { pcol = q; col_radix = v; }
        } else if (k == "row_block") {
            if (v != 2 && v != 4 && v != 8) return set_err(MRMG_ERR_ARG, "row_block must be 2, 4 or 8");
            const size_t l = sizeof(C2<T>) * c->nx + sizeof(T) * 2 * LPN((v / 2) * c->nx);
            if (l > 160 * 1024 || pick_e(v / 2, c->px, 256) < 0)
                return set_err(MRMG_ERR_UNSUPPORTED, "row_block %d does not fit LDS", v);
            const int ot = TYr;
            const size_t ol = lds_rowT;
            TYr = v;
            lds_rowT = l;
            if (shapes() != MRMG_OK || lds_attr() != MRMG_OK) { TYr = ot; lds_rowT = ol; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else {
            return set_err(MRMG_ERR_ARG, "unknown tunable '%s'", k.c_str());
        }
        return MRMG_OK;
    }
    int get_tune(const char* key, int* v) override
    {
        std::string k(key ? key : "");
        if (k == "row_threads") *v = row_threads;
        else if (k == "col_threads") *v = col_threads;
        else if (k == "col_groups") *v = kg;
        else if (k == "lanes_per_grain") *v = lpg;
        else if (k == "col_path") *v = split ? 1 : tl ? 2 : 0;
        else if (k == "row_block") *v = TYr;
        else if (k == "wave_fft") *v = wv ? 1 : 0;
        else if (k == "row_radix") *v = row_radix;
        else if (k == "col_radix") *v = col_radix;
        else return set_err(MRMG_ERR_ARG, "unknown tunable '%s'", k.c_str());
        return MRMG_OK;
    }

    template <int MB> int lds_attr_mb()
    {
        const int r = (int)lds_row, cl = (int)lds_col;
        HIPCHK(hipFuncSetAttribute((const void*)k_row_fwd<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
        HIPCHK(hipFuncSetAttribute((const void*)k_row_inv<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
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
        return MRMG_OK;
    }
    template <int MB> int lds_attr_split()
    {
        const int r = (int)lds_rowT, cl = (int)lds_colT;
        HIPCHK(hipFuncSetAttribute((const void*)k_rowT_fwd<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
        HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
        HIPCHK(hipFuncSetAttribute((const void*)k_rowT_fwd<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
        HIPCHK(hipFuncSetAttribute((const void*)k_rowT_inv<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, r));
        HIPCHK(hipFuncSetAttribute((const void*)k_colT<T, MB, false>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
        HIPCHK(hipFuncSetAttribute((const void*)k_colT<T, MB, true>, hipFuncAttributeMaxDynamicSharedMemorySize, cl));
        return MRMG_OK;
    }
    // dynamic LDS above 64 KiB must be requested per kernel
    int lds_attr()
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

    int build(std::vector<HostLayer>& L) override
    {
        const int nl = c->nl, nx = c->nx, ny = c->ny;
        const size_t P = (size_t)nx * ny;
        // a switching record refers to the grains of the previous build (and is sized for
        // them): a rebuild ends it
        sw_on = false;
        sw_step.release();
        sw_t.release();
        sw_h.release();
        off.assign(nl + 1, 0);
        for (int k = 0; k < nl; k++) off[k + 1] = off[k] + L[k].ng;
        G = off[nl];
        Ctot = 0;
        for (int k = 0; k < nl; k++) Ctot += L[k].ncell;
        std::vector<int> h_layer(G), h_gco(G + 1), h_gci(Ctot), h_c2g(nl * P, -1);
        std::vector<double> h_ms(G), h_msf(G), h_mss(G), h_as(G), h_hks(G), h_hk1(G), h_hk2(G), h_uni(3 * (size_t)G, 0.0),
            h_cub(9 * (size_t)G, 0.0), h_happ(3 * (size_t)G, 0.0), h_sig(G, 0.0), h_sigl(G, 0.0), h_vol(G);
        std::vector<double> h_tc(G), h_lam(G), h_temp(G), h_tabs;
        std::vector<double> h_m((size_t)ST_ROWS * G);
        has_happ = false;
        hlp.assign(nl, LayerParams{});
        int q = 0;
        h_gco[0] = 0;
        for (int k = 0; k < nl; k++) {
            const HostLayer& H = L[k];
            for (int n = 0; n < H.ng; n++) {
                const int g = off[k] + n;
                h_layer[g] = k;
                for (int e = H.g2gi[n]; e < H.g2gi[n + 1]; e++) {
                    const int i = H.g2g[2 * e] - 1, j = H.g2g[2 * e + 1] - 1;
                    if (i < 0 || i >= nx || j < 0 || j >= ny)
                        return set_err(MRMG_ERR_ARG, "layer %d grain %d: cell (%d,%d) outside grid", k, n + 1, i + 1, j + 1);
                    h_gci[q++] = (int)(k * P + (size_t)j * nx + i);
                }
                h_gco[g + 1] = q;
                if (H.g2gi[n + 1] == H.g2gi[n]) return set_err(MRMG_ERR_ARG, "layer %d grain %d has no cells", k, n + 1);
                h_ms[g] = H.ms[n];
                h_mss[g] = H.ms_scale[n];
                h_msf[g] = H.ms[n] * H.ms_scale[n];
                h_as[g] = H.alpha_scale[n];
                h_hks[g] = H.hk_scale[n];
                h_hk1[g] = H.hk[2 * n];
                h_hk2[g] = H.hk[2 * n + 1];
                h_vol[g] = H.vol[n];
                if (H.has_uni) for (int cc = 0; cc < 3; cc++) h_uni[(size_t)cc * G + g] = H.uni[3 * n + cc];
                if (H.has_cub) for (int cc = 0; cc < 9; cc++) h_cub[(size_t)cc * G + g] = H.cub[9 * n + cc];
                if (H.has_happ) for (int cc = 0; cc < 3; cc++) h_happ[(size_t)cc * G + g] = H.happ[3 * n + cc];
                h_sig[g] = thermal_sigma(H, n, H.mag[n], false);
                h_sigl[g] = thermal_sigma(H, n, H.mag[n], true);
                h_tc[g] = H.tc[n];
                h_lam[g] = H.lambda_ms[n] * (H.spin + 1) / H.spin;
                h_temp[g] = H.temp[n];
                for (int cc = 0; cc < 3; cc++) h_m[(size_t)cc * G + g] = H.m[3 * n + cc];
                h_m[(size_t)ST_MAG * G + g] = H.mag[n];
                h_m[(size_t)ST_MSC * G + g] = field_ms_scale(H, n, H.mag[n]);
            }
            for (size_t p = 0; p < P; p++) {
                const int gr = H.grid2grain[p];
                if (gr < 0 || gr > H.ng) return set_err(MRMG_ERR_ARG, "layer %d grid2grain value %d out of range", k, gr);
                h_c2g[k * P + p] = gr > 0 ? off[k] + gr - 1 : -1;
            }
            has_happ |= H.has_happ;
            LayerParams& lpk = hlp[k];
            lpk.alpha = H.alpha;
            lpk.gamma = H.gamma;
            lpk.spin = H.spin;
            lpk.use_mag = H.use_mag;
            lpk.tsigma = sqrt(2.0 * 1.380662e-16 * H.thm_alpha / H.gamma);
            lpk.floor_ms = H.floor_ms;
            lpk.msexp = H.msexp;
            lpk.hkexp = H.hkexp;
            lpk.tscale = H.tscale;
            for (int i = 0; i < 3; i++) {
                lpk.tab_off[i] = (int)h_tabs.size();
                lpk.tab_n[i] = (int)H.tab_t[i].size();
                h_tabs.insert(h_tabs.end(), H.tab_t[i].begin(), H.tab_t[i].end());
                h_tabs.insert(h_tabs.end(), H.tab_s[i].begin(), H.tab_s[i].end());
            }
            lpk.has_uni = H.has_uni;
            lpk.has_cub = H.has_cub;
            lpk.thermal = H.thermal;
            lpk.off = off[k];
            lpk.ng = H.ng;
        }
        RCHK(fill_sources(L));
        RCHK(fill_ufield(L));
        // exchange CSR in global numbering: per grain the intralayer entries, then those to
        // the layer above, then to the layer below, and one spare entry at the end
        {
            std::vector<int> o(G + 1, 0), nb;
            std::vector<S> w;
            for (int k = 0; k < nl; k++) {
                const HostLayer& H = L[k];
                for (int n = 0; n < H.ng; n++) {
                    for (int which = 0; which < 3; which++) {
                        const std::vector<int>& ix = which == 0 ? H.hxi : which == 1 ? H.uhxi : H.lhxi;
                        const std::vector<int>& gx = which == 0 ? H.hxg : which == 1 ? H.uhxg : H.lhxg;
                        const std::vector<double>& vx = which == 0 ? H.hx : which == 1 ? H.uhx : H.lhx;
                        const bool has = which == 0 ? H.has_hx : which == 1 ? H.has_uhx : H.has_lhx;
                        const int tk = which == 0 ? k : which == 1 ? k - 1 : k + 1;
                        if (!has) continue;
                        if (tk < 0 || tk >= nl) return set_err(MRMG_ERR_ARG, "layer %d: interlayer CSR to missing layer", k);
                        for (int e = ix[n]; e < ix[n + 1]; e++) {
                            const int gg = gx[e];
                            if (gg < 1 || gg > L[tk].ng)
                                return set_err(MRMG_ERR_ARG, "layer %d: exchange neighbour %d out of range", k, gg);
                            nb.push_back(off[tk] + gg - 1);
                            w.push_back(vx[e]);
                        }
                    }
                    o[off[k] + n + 1] = (int)nb.size();
                }
            }
            nb.push_back(0);
            w.push_back(0.0);
            RCHK(x_off.upload(o.data(), o.size(), c->stream));
            RCHK(x_nb.upload(nb.data(), nb.size(), c->stream));
            RCHK(x_w.upload(w.data(), w.size(), c->stream));
        }

        auto upT = [&](DBuf<S>& d, const std::vector<double>& v) -> int { return d.upload(v.data(), v.size(), c->stream); };
        RCHK(layer_of.upload(h_layer.data(), G, c->stream));
        RCHK(gc_off.upload(h_gco.data(), G + 1, c->stream));
        h_gci.push_back(0);   // spare entry
        RCHK(gc_idx.upload(h_gci.data(), h_gci.size(), c->stream));
        RCHK(cell2grain.upload(h_c2g.data(), h_c2g.size(), c->stream));
        RCHK(upT(ms, h_ms)); RCHK(upT(msf, h_msf)); RCHK(upT(ms_scale, h_mss)); RCHK(upT(alpha_scale, h_as));
        RCHK(upT(hk_scale, h_hks)); RCHK(upT(hk1, h_hk1)); RCHK(upT(hk2, h_hk2)); RCHK(upT(uni, h_uni));
        RCHK(upT(cub, h_cub)); RCHK(upT(sigma, h_sig)); RCHK(upT(sigma_llb, h_sigl)); RCHK(upT(vol, h_vol));
        RCHK(tc.upload(h_tc.data(), G, c->stream));
        RCHK(lam.upload(h_lam.data(), G, c->stream));
        RCHK(temp.upload(h_temp.data(), G, c->stream));
        h_tabs.push_back(0.0);   // never empty
        RCHK(tabs.upload(h_tabs.data(), h_tabs.size(), c->stream));
        if (has_happ) RCHK(upT(happ, h_happ)); else happ.release();
        RCHK(upT(ma, h_m));
        mag_dirty = false;
        RCHK(mb.alloc((size_t)ST_ROWS * G)); RCHK(mc.alloc((size_t)ST_ROWS * G)); RCHK(dm.alloc(3 * (size_t)G)); RCHK(hthm.alloc(3 * (size_t)G)); RCHK(h.alloc(3 * (size_t)G));
        HIPCHK(hipMemsetAsync(mb.p, 0, mb.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(dm.p, 0, dm.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(hthm.p, 0, hthm.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(h.p, 0, h.bytes(), c->stream));
        RCHK(diag_part.alloc((size_t)(((size_t)G * 16 + 255) / 256) * DIAG_STRIDE));
        RCHK(stats.alloc((size_t)10 * nl));
        HIPCHK(hipMemsetAsync(Mcell.p, 0, Mcell.bytes(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MRMG_OK;
    }

    // the moving sources of every layer into LayerParams and the device tables
    int fill_sources(std::vector<HostLayer>& L)
    {
        const int nl = c->nl, nx = c->nx, ny = c->ny;
        std::vector<double> pos, vals;
        any_src = false;
        for (int k = 0; k < nl; k++) {
            for (int sl = 0; sl < 2; sl++) {
                SrcParams& sp = hlp[k].src[sl];
                sp = SrcParams{};
                const mrmg_source& q = L[k].src[sl];
                sp.kind = q.kind;
                if (q.kind == MRMG_SRC_NONE) continue;
                any_src = true;
                sp.nxp = q.nxp; sp.nyp = q.nyp; sp.expn = q.expn;
                sp.vel_x = q.vel_x; sp.vel_y = q.vel_y; sp.offset_x = q.offset_x; sp.offset_y = q.offset_y;
                sp.ref_x = q.ref_x; sp.ref_y = q.ref_y; sp.t_origin = q.t_origin; sp.isigx = q.isigx; sp.isigy = q.isigy;
                sp.dtemp = q.dtemp;
                // set_temp/set_fld normalisation of a constant scale (applied.f90:307-310)
                sp.t0 = q.t_start; sp.t1 = q.t_end; sp.s0 = q.s_start; sp.s1 = q.s_end; sp.scale = q.scale;
                sp.fn = norm_tfun(q.func, q.amp, q.gamma, q.phase, q.param, q.duty, q.moddepth, q.dc, sp.s0, sp.s1,
                                  sp.scale);
                if (q.kind != MRMG_SRC_TEMP_SG) {
                    sp.poff_x = (int)pos.size();
                    pos.insert(pos.end(), L[k].src_x[sl].begin(), L[k].src_x[sl].end());
                    sp.poff_y = (int)pos.size();
                    pos.insert(pos.end(), L[k].src_y[sl].begin(), L[k].src_y[sl].end());
                    sp.voff = (int)vals.size();
                    vals.insert(vals.end(), L[k].src_v[sl].begin(), L[k].src_v[sl].end());
                }
            }
        }
        if (!any_src) {
            src_pos.release(); src_vals.release(); ax_xa.release(); ax_ya.release(); ax_yb.release();
            ax_xi.release(); ax_yi.release(); tlast.release();
            return MRMG_OK;
        }
        if (pos.empty()) pos.push_back(0.0);
        if (vals.empty()) vals.push_back(0.0);
        RCHK(src_pos.upload(pos.data(), pos.size(), c->stream));
        RCHK(src_vals.upload(vals.data(), vals.size(), c->stream));
        const size_t nax = (size_t)nl * 4 * nx, nay = (size_t)nl * 4 * ny;
        RCHK(ax_xa.alloc(nax)); RCHK(ax_xi.alloc(nax));
        RCHK(ax_ya.alloc(nay)); RCHK(ax_yb.alloc(nay)); RCHK(ax_yi.alloc(nay));
        HIPCHK(hipMemsetAsync(ax_xi.p, 0, ax_xi.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(ax_yi.p, 0, ax_yi.bytes(), c->stream));
        RCHK(tlast.alloc(G));
        return MRMG_OK;
    }

    // the axis tables of the moving sources at the evaluation time t0 and the next t1
    int src_axes(double t0, double t1)
    {
        if (!any_src) return MRMG_OK;
        const int n = c->nl * 4 * (c->nx + c->ny);
        hipLaunchKernelGGL(k_src_axes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, lp.p, c->nl, c->nx,
                           c->ny, c->dx, c->dy, t0, t1, ax_xi.p, ax_xa.p, ax_yi.p, ax_ya.p, ax_yb.p, src_pos.p);
        HIPCHK(hipGetLastError());
        return MRMG_OK;
    }

    int fill_ufield(std::vector<HostLayer>& L)
    {
        for (int k = 0; k < c->nl; k++) {
            LayerParams& p = hlp[k];
            p.has_uf = L[k].has_uf;
            mrmg_ufield u = L[k].uf;
            p.uf_fn = norm_tfun(u.func, u.amp, u.gamma, u.phase, u.param, u.duty, u.moddepth, u.dc, u.s_start, u.s_end,
                                u.scale);
            for (int i = 0; i < 3; i++) p.uf_dir[i] = u.dir[i];
            p.uf_t0 = u.t_start; p.uf_t1 = u.t_end; p.uf_s0 = u.s_start; p.uf_s1 = u.s_end; p.uf_scale = u.scale;
        }
        RCHK(lp.upload(hlp.data(), hlp.size(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MRMG_OK;
    }

    int update_ufield(std::vector<HostLayer>& L) override { return fill_ufield(L); }

    int set_tensor_planes(const double* d_real) override
    {
        // d_real: device (nx,ny,9,nl,nl) FP64
        const int nx = c->nx, ny = c->ny, nxh = c->nxh, nl = c->nl;
        const int P = 9 * nl * nl;
        DBuf<C2<double>> TC, tdx, tdy;
        RCHK(TC.alloc((size_t)P * nxh * ny));
        auto wx = twiddle_table(nx), wy = twiddle_table(ny);
        RCHK(tdx.upload(wx.data(), nx, c->stream));
        RCHK(tdy.upload(wy.data(), ny, c->stream));
        const int mbr = pick_e(1, c->px, 256), mbc = pick_e(1, c->py, 256);
        if (mbr < 0 || mbc < 0) return set_err(MRMG_ERR_UNSUPPORTED, "tensor FFT too large");
        {
            dim3 grid((ny + 1) / 2, P), blk(256);
            size_t lds = sizeof(double) * (2 * nx + 2 * LPN(nx));
            if (lds > 160 * 1024) return set_err(MRMG_ERR_UNSUPPORTED, "tensor row FFT of %d does not fit LDS", nx);
            if (lds > 64 * 1024) {
                const void* fk[5] = {(const void*)k_tensor_rows<4>, (const void*)k_tensor_rows<8>, (const void*)k_tensor_rows<16>,
                                     (const void*)k_tensor_rows<32>, (const void*)k_tensor_rows<64>};
                for (const void* f : fk) HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            }
#define KTR(MB, ...) hipLaunchKernelGGL(k_tensor_rows<MB>, grid, blk, lds, c->stream, __VA_ARGS__)
            switch (mbr) {
            case 4: KTR(4, d_real, TC.p, nx, ny, nxh, c->px, tdx.p); break;
            case 8: KTR(8, d_real, TC.p, nx, ny, nxh, c->px, tdx.p); break;
            case 16: KTR(16, d_real, TC.p, nx, ny, nxh, c->px, tdx.p); break;
            case 32: KTR(32, d_real, TC.p, nx, ny, nxh, c->px, tdx.p); break;
            default: KTR(64, d_real, TC.p, nx, ny, nxh, c->px, tdx.p); break;
            }
#undef KTR
            HIPCHK(hipGetLastError());
        }
        {
            dim3 grid(nxh, P), blk(256);
            size_t lds = sizeof(double) * (2 * ny + 2 * LPN(ny));
            if (lds > 160 * 1024) return set_err(MRMG_ERR_UNSUPPORTED, "tensor column FFT of %d does not fit LDS", ny);
            if (lds > 64 * 1024) {
                const void* fk[5] = {(const void*)k_tensor_cols<4>, (const void*)k_tensor_cols<8>, (const void*)k_tensor_cols<16>,
                                     (const void*)k_tensor_cols<32>, (const void*)k_tensor_cols<64>};
                for (const void* f : fk) HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            }
#define KTC(MB) hipLaunchKernelGGL(k_tensor_cols<MB>, grid, blk, lds, c->stream, TC.p, ny, nxh, c->py, tdy.p)
            switch (mbc) {
            case 4: KTC(4); break;
            case 8: KTC(8); break;
            case 16: KTC(16); break;
            case 32: KTC(32); break;
            default: KTC(64); break;
            }
#undef KTC
            HIPCHK(hipGetLastError());
        }
        // symmetry decides 6 or 9 slots: checked on the real tensor by the caller
        const size_t total = (size_t)nxh * nl * nl * nslot * ny;
        RCHK(Tspec.alloc(total));
        hipLaunchKernelGGL(k_tensor_pack<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, TC.p,
                           Tspec.p, ny, nxh, nl, nslot, slot2c.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        tensor_ok = true;
        return MRMG_OK;
    }

    int set_slots(bool sym)
    {
        nslot = sym ? 6 : 9;
        if (!sym && col_threads > 512) {   // the 1024-thread column kernel is built for 6 slots only
            col_threads = 512;
            RCHK(shapes());
        }
        const int s6[6] = {0, 4, 8, 1, 2, 5};   // xx yy zz xy xz yz as c = 3a+b
        int s9[9];
        for (int i = 0; i < 9; i++) s9[i] = i;
        return slot2c.upload(sym ? s6 : s9, nslot, c->stream);
    }

    int set_state(int k, const double* m) override
    {
        const int ng = c->layers[k].ng;
        std::vector<S> t(3 * (size_t)ng);
        for (int n = 0; n < ng; n++)
            for (int cc = 0; cc < 3; cc++) t[(size_t)cc * ng + n] = (S)m[3 * n + cc];
        for (int cc = 0; cc < 3; cc++)
            HIPCHK(hipMemcpyAsync(ma.p + (size_t)cc * G + off[k], t.data() + (size_t)cc * ng, sizeof(S) * ng,
                                  hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MRMG_OK;
    }

    int get_state(int k, double* m, double* hh) override
    {
        const int ng = c->layers[k].ng;
        std::vector<S> t(3 * (size_t)ng);
        for (int pass = 0; pass < 2; pass++) {
            double* out = pass == 0 ? m : hh;
            S* src = pass == 0 ? ma.p : h.p;
            if (!out) continue;
            for (int cc = 0; cc < 3; cc++)
                HIPCHK(hipMemcpyAsync(t.data() + (size_t)cc * ng, src + (size_t)cc * G + off[k], sizeof(S) * ng,
                                      hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int n = 0; n < ng; n++)
                for (int cc = 0; cc < 3; cc++) out[3 * n + cc] = (double)t[(size_t)cc * ng + n];
        }
        return MRMG_OK;
    }

    GrainData<S> gdata() const
    {
        GrainData<S> d;
        d.G = G;
        d.layer_of = layer_of.p;
        d.ms = ms.p; d.msf = msf.p; d.ms_scale = ms_scale.p; d.alpha_scale = alpha_scale.p; d.hk_scale = hk_scale.p;
        d.hk1 = hk1.p; d.hk2 = hk2.p; d.uni = uni.p; d.cub = cub.p; d.happ = has_happ ? happ.p : nullptr;
        d.sigma = sigma.p; d.vol = vol.p; d.gc_off = gc_off.p; d.gc_idx = gc_idx.p;
        d.x_off = x_off.p; d.x_nb = x_nb.p; d.x_w = x_w.p;
        d.nl = c->nl;
        d.lp = lp.p;
        d.tc = tc.p; d.lam = lam.p; d.temp = temp.p; d.tabs = tabs.p;
        if (cur_llb) d.sigma = sigma_llb.p;
        d.any_src = any_src;
        d.ax.xi = ax_xi.p; d.ax.xa = ax_xa.p; d.ax.yi = ax_yi.p; d.ax.ya = ax_ya.p; d.ax.yb = ax_yb.p;
        d.ax.pos = src_pos.p; d.ax.vals = src_vals.p; d.ax.nx = c->nx; d.ax.ny = c->ny;
        return d;
    }

    // |m| of the state: rows ST_MAG and ST_MSC of ma are the truth; the other buffers take
    // a copy before a step so that the buffer swaps of the LLG pushes (which leave |m|
    // alone) keep it
    int sync_aux()
    {
        const size_t off3 = (size_t)ST_MAG * G, n = 2 * (size_t)G;
        HIPCHK(hipMemcpyAsync(mb.p + off3, ma.p + off3, n * sizeof(S), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(mc.p + off3, ma.p + off3, n * sizeof(S), hipMemcpyDeviceToDevice, c->stream));
        return MRMG_OK;
    }
    // the LLG thermal variance depends on |m| when it is not 1: rebuild it from the device |m|
    int refresh_sigma()
    {
        std::vector<S> t(G);
        HIPCHK(hipMemcpyAsync(t.data(), ma.p + (size_t)ST_MAG * G, sizeof(S) * G, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<S> sg(G);
        for (int k = 0; k < c->nl; k++) {
            HostLayer& H = c->layers[k];
            H.mag.resize(H.ng);
            for (int n = 0; n < H.ng; n++) {
                H.mag[n] = (double)t[off[k] + n];
                sg[off[k] + n] = (S)thermal_sigma(H, n, H.mag[n], false);
            }
        }
        RCHK(sigma.upload(sg.data(), G, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        mag_dirty = false;
        return MRMG_OK;
    }
    int set_mag(int k, const double* mg) override
    {
        const HostLayer& H = c->layers[k];
        std::vector<S> a(H.ng), b(H.ng);
        for (int n = 0; n < H.ng; n++) {
            a[n] = (S)mg[n];
            b[n] = (S)field_ms_scale(H, n, mg[n]);
        }
        HIPCHK(hipMemcpyAsync(ma.p + (size_t)ST_MAG * G + off[k], a.data(), sizeof(S) * H.ng, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(ma.p + (size_t)ST_MSC * G + off[k], b.data(), sizeof(S) * H.ng, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        mag_dirty = true;
        return MRMG_OK;
    }
    int get_mag(int k, double* mg) override
    {
        const HostLayer& H = c->layers[k];
        std::vector<S> a(H.ng);
        HIPCHK(hipMemcpyAsync(a.data(), ma.p + (size_t)ST_MAG * G + off[k], sizeof(S) * H.ng, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int n = 0; n < H.ng; n++) mg[n] = (double)a[n];
        return MRMG_OK;
    }

    bool demag_on() const { return c->demag && tensor_ok; }

    // get_mag for the state m (used where no grain kernel has written the cells yet); with
    // t0 given, Ms scales that follow a moving temperature are taken at t0
    int scatter(S* m, const double* t0 = nullptr)
    {
        const size_t CP = (size_t)c->nl * c->nx * c->ny;
        const size_t threads = (size_t)G * 8;
        if (t0) RCHK(src_axes(*t0, *t0));
        hipLaunchKernelGGL((k_scatter<S, 8, T>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream,
                           gdata(), m, Mcell.p, CP, t0 && any_src ? 1 : 0);
        HIPCHK(hipGetLastError());
        return MRMG_OK;
    }

    // one launch of a demag kernel class (K_ROW_FWD, K_COL, K_ROW_INV on the fused paths;
    // K_ROW_FWD, K_COLF, K_PROD, K_COLI, K_ROW_INV on the split path)
    int launch_class(int cls)
    {
        const int ny = c->ny, nl = c->nl, nxh = c->nxh;
        const dim3 grow((ny + 1) / 2, nl), growT(rowT_blocks()), blk(row_threads), b256(256);
        const int ncols = 3 * nl * nxh;
        const dim3 gcol((ncols + CBc - 1) / CBc);
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
            if (!split) return set_err(MRMG_ERR_STATE, "the fused paths have no separate column FFTs");
            if (cls == K_COLF) LAUNCH_E(maxb_colT, launch_colT_fwd, gcol, b256);
            else LAUNCH_E(maxb_colT, launch_colT_inv, gcol, b256);
            break;
        case K_PROD: {
            if (!split) return set_err(MRMG_ERR_STATE, "the fused paths have no separate product kernel");
            const size_t Kc = (size_t)nxh * ny;
            const dim3 gp((unsigned)((Kc + 255) / 256));
            if (nslot == 6)
                hipLaunchKernelGGL((k_prod<T, true, 4>), gp, b256, 0, c->stream, S1.p, S2.p, Tspec.p, ny, nl, nxh, nslot);
            else
                hipLaunchKernelGGL((k_prod<T, false, 4>), gp, b256, 0, c->stream, S1.p, S2.p, Tspec.p, ny, nl, nxh,
                                   nslot);
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

    // get_demag: row r2c of the cell magnetisation, column FFT * tensor, row c2r
    int demag()
    {
        static const int fused[3] = {K_ROW_FWD, K_COL, K_ROW_INV};
        static const int splitc[5] = {K_ROW_FWD, K_COLF, K_PROD, K_COLI, K_ROW_INV};
        const int* seq = split ? splitc : fused;
        const int n = split ? 5 : 3;
        for (int i = 0; i < n; i++) {
            RCHK(ev_begin(seq[i]));
            RCHK(launch_class(seq[i]));
            RCHK(ev_end());
        }
        return MRMG_OK;
    }

    // reps back-to-back launches of one demag kernel class between two events on the
    // context's stream: the kernel's average duration without a per-launch event pair.
    // Only the scratch spectra and cell fields are written; the state is untouched.
    int time_class(int cls, int reps, double* ms) override
    {
        if (!demag_on()) return set_err(MRMG_ERR_STATE, "demag is off or has no tensor");
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev() { if (e) (void)hipEventDestroy(e); }
        } a, b;
        HIPCHK(hipEventCreate(&a.e));
        HIPCHK(hipEventCreate(&b.e));
        RCHK(launch_class(cls));   // warm
        HIPCHK(hipEventRecord(a.e, c->stream));
        for (int r = 0; r < reps; r++) RCHK(launch_class(cls));
        HIPCHK(hipEventRecord(b.e, c->stream));
        HIPCHK(hipEventSynchronize(b.e));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, a.e, b.e));
        *ms = (double)t / reps;
        return MRMG_OK;
    }
    // transposing row kernels: (ny/TY) * 3 n_lay blocks on a 1D grid padded to a multiple of 8
    unsigned rowT_blocks() const { return (unsigned)(((c->ny + TYr - 1) / TYr * 3 * c->nl + 7) / 8 * 8); }
    template <int MB> void launch_rowT_fwd(dim3 g, dim3 b)
    {
        if (wv)
            hipLaunchKernelGGL((k_rowT_fwd<T, MB, true>), g, b, lds_rowT, c->stream, Mcell.p, S1.p, c->nx, c->ny, c->nl,
                               c->nxh, TYr, prow, twx.p);
        else
            hipLaunchKernelGGL((k_rowT_fwd<T, MB>), g, b, lds_rowT, c->stream, Mcell.p, S1.p, c->nx, c->ny, c->nl,
                               c->nxh, TYr, prow, twx.p);
    }
    template <int MB> void launch_colT_fwd(dim3 g, dim3 b)
    {
        hipLaunchKernelGGL((k_colT<T, MB, false>), g, b, lds_colT, c->stream, S1.p, c->ny, 3 * c->nl * c->nxh, CBc,
                           pcol, twy.p);
    }
    template <int MB> void launch_colT_inv(dim3 g, dim3 b)
    {
        hipLaunchKernelGGL((k_colT<T, MB, true>), g, b, lds_colT, c->stream, S2.p, c->ny, 3 * c->nl * c->nxh, CBc,
                           pcol, twy.p);
    }
    template <int MB> void launch_rowT_inv(dim3 g, dim3 b)
    {
        if (wv)
            hipLaunchKernelGGL((k_rowT_inv<T, MB, true>), g, b, lds_rowT, c->stream, S2.p, Hcell.p, c->nx, c->ny, c->nl,
                               c->nxh, TYr, prow, twx.p);
        else
            hipLaunchKernelGGL((k_rowT_inv<T, MB>), g, b, lds_rowT, c->stream, S2.p, Hcell.p, c->nx, c->ny, c->nl,
                               c->nxh, TYr, prow, twx.p);
    }
    template <int MB> void launch_row_fwd(dim3 g, dim3 b)
    {
        hipLaunchKernelGGL((k_row_fwd<T, MB>), g, b, lds_row, c->stream, Mcell.p, S1.p, c->nx, c->ny, c->nl, c->nxh,
                           prow, twx.p);
    }
    template <int MB> void launch_col_sym(int nb8)
    {
        if (col_threads > 512) {
            if (tl)
                hipLaunchKernelGGL((k_col<T, MB, true, true, false, true>), dim3(nb8), dim3(col_threads), lds_col,
                                   c->stream, S1.p, S2.p, Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
            else
                hipLaunchKernelGGL((k_col<T, MB, true, false, false, true>), dim3(nb8), dim3(col_threads), lds_col,
                                   c->stream, S1.p, S2.p, Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
        } else if (tl && wv)
            hipLaunchKernelGGL((k_col<T, MB, true, true, true>), dim3(nb8), dim3(col_threads), lds_col, c->stream, S1.p,
                               S2.p, Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
        else if (tl)
            hipLaunchKernelGGL((k_col<T, MB, true, true>), dim3(nb8), dim3(col_threads), lds_col, c->stream, S1.p, S2.p,
                               Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
        else
            hipLaunchKernelGGL((k_col<T, MB, true>), dim3(nb8), dim3(col_threads), lds_col, c->stream, S1.p, S2.p,
                               Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
    }
    template <int MB> void launch_col_gen(int nb8)
    {
        if (tl && wv)
            hipLaunchKernelGGL((k_col<T, MB, false, true, true>), dim3(nb8), dim3(col_threads), lds_col, c->stream,
                               S1.p, S2.p, Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
        else if (tl)
            hipLaunchKernelGGL((k_col<T, MB, false, true>), dim3(nb8), dim3(col_threads), lds_col, c->stream, S1.p,
                               S2.p, Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
        else
            hipLaunchKernelGGL((k_col<T, MB, false>), dim3(nb8), dim3(col_threads), lds_col, c->stream, S1.p, S2.p,
                               Tspec.p, c->ny, c->nl, c->nxh, nslot, kg, nb8, pcol, twy.p);
    }
    template <int MB> void launch_row_inv(dim3 g, dim3 b)
    {
        hipLaunchKernelGGL((k_row_inv<T, MB>), g, b, lds_row, c->stream, S2.p, Hcell.p, c->nx, c->ny, c->nl, c->nxh,
                           prow, twx.p);
    }

    template <int MODE, int LPG> void launch_grain(const StepArgs& a, S* mout)
    {
        const size_t CP = (size_t)c->nl * c->nx * c->ny;
        const T* hc = demag_on() ? Hcell.p : nullptr;
        const size_t threads = (size_t)G * LPG;
        const unsigned nb8 = (unsigned)(((threads + 255) / 256 + 7) / 8 * 8);
        hipLaunchKernelGGL((k_grain<S, MODE, LPG, T>), dim3(nb8), dim3(256), 0, c->stream,
                           gdata(), ma.p, mb.p, dm.p, hthm.p, h.p, hc, demag_on() ? Mcell.p : nullptr, CP, a,
                           diag_part.p, mout);
    }
    // mout: destination of the corrector (null: ma); t_next: time of the next field
    // evaluation, for the moving sources
    template <int MODE> int grain(const StepArgs& a, S* mout = nullptr, double t_next = 0.0)
    {
        RCHK(src_axes(a.t, t_next));
        RCHK(ev_begin((MODE == GM_HEUN1 || MODE == GM_ROT1 || MODE == GM_FIELD1 || MODE == GM_PUSH1) ? K_GRAIN1
                                                                                                 : K_GRAIN2));
        if (lpg == 4) launch_grain<MODE, 4>(a, mout);
        else if (lpg == 16) launch_grain<MODE, 16>(a, mout);
        else launch_grain<MODE, 8>(a, mout);
        HIPCHK(hipGetLastError());
        RCHK(ev_end());
        return MRMG_OK;
    }

    // per-layer reductions of the last diag-enabled grain launch
    struct LayerRed {
        std::vector<double> errmx, errmx1, mmag, hmag;
    };
    int read_partials(LayerRed& r)
    {
        const int nb = (int)(((size_t)G * lpg + 255) / 256), nl = c->nl;
        std::vector<double> part((size_t)nb * DIAG_STRIDE);
        HIPCHK(hipMemcpy(part.data(), diag_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost));
        r.errmx.assign(nl, 0.0); r.errmx1.assign(nl, 0.0); r.mmag.assign(nl, 0.0); r.hmag.assign(nl, 0.0);
        for (int b = 0; b < nb; b++) {
            const double* q = &part[(size_t)b * DIAG_STRIDE];
            for (int k = 0; k < nl; k++) {
                r.errmx[k] = std::max(r.errmx[k], q[D_ERRMX + k]);
                r.errmx1[k] = std::max(r.errmx1[k], q[D_ERRMX1 + k]);
                r.hmag[k] = std::max(r.hmag[k], q[D_HMAG + k]);
                r.mmag[k] += q[D_MMAG + k];
            }
        }
        return MRMG_OK;
    }
    // llg()'s loop-top statistics of the current state and h, into stats
    int layer_stats()
    {
        hipLaunchKernelGGL(k_layer_stats, dim3(c->nl), dim3(STATS_THREADS), 0, c->stream, ma.p, h.p, vol.p, lp.p, G,
                           stats.p);
        HIPCHK(hipGetLastError());
        return MRMG_OK;
    }
    // the statistics into the diagnostics record, the layers combined as llg.f90:231-233 does
    int read_stats(mrmg_diag* d)
    {
        const int nl = c->nl;
        std::vector<double> o((size_t)10 * nl);
        HIPCHK(hipMemcpy(o.data(), stats.p, o.size() * sizeof(double), hipMemcpyDeviceToHost));
        double x = 0.0;
        for (int k = 0; k < nl; k++) {
            const double* q = &o[(size_t)10 * k];
            const double x1 = q[0];
            x = x + x1;
            for (int cc = 0; cc < 3; cc++) {
                d->mave_all[cc] = d->mave_all[cc] + x1 / x * (q[1 + cc] - d->mave_all[cc]);
                d->have_all[cc] = d->have_all[cc] + x1 / x * (q[7 + cc] - d->have_all[cc]);
                d->mvar_all[cc] = d->mvar_all[cc] + x1 / x * (q[4 + cc] - d->mvar_all[cc]);
                if (k < 8) {
                    d->mave[k][cc] = q[1 + cc];
                    d->mvar[k][cc] = q[4 + cc];
                    d->have[k][cc] = q[7 + cc];
                }
            }
        }
        return MRMG_OK;
    }
    int zero_partials()
    {
        const int nb = (int)(((size_t)G * lpg + 255) / 256);
        HIPCHK(hipMemsetAsync(diag_part.p, 0, (size_t)nb * DIAG_STRIDE * sizeof(double), c->stream));
        return MRMG_OK;
    }

    // The llg() loop with get_fields' adaptive dt (llg.f90:320-388) and push_m's rejection
    // of steps that rotate a layer by more than max_dm (llg.f90:1195-1271), Heun only.
    int step_adaptive(const mrmg_step_params* p, const double* noise, mrmg_diag* diag)
    {
        const int nl = c->nl;
        const size_t per = 3 * (size_t)G;
        const double max_dm = p->max_dm * acos(-1.0) / 180.0;
        if (p->noise == MRMG_NOISE_INJECT) RCHK(noise_buf.alloc(per));
        HIPCHK(hipMemsetAsync(hthm.p, 0, hthm.bytes(), c->stream));
        if (demag_on()) RCHK(scatter(ma.p, &p->t0));
        double t = p->t0, dt = p->dt, max_dmdt = 1.0, max_dmdtnext = 1.0;
        int rejected = 0;
        LayerRed r;
        StepArgs a{};
        a.llb = p->llb;
        a.tlast = any_src ? tlast.p : nullptr;
        a.type = p->type;
        a.noise = p->noise;
        a.seed = p->seed;
        for (int it = 0; it < p->n_steps; it++) {
            a.step = p->step0 + it;
            a.normals = nullptr;
            if (p->noise == MRMG_NOISE_INJECT) {
                HIPCHK(hipMemcpyAsync(noise_buf.p, noise + per * it, sizeof(double) * per, hipMemcpyHostToDevice,
                                      c->stream));
                a.normals = noise_buf.p;
            }
            if (diag && it == p->n_steps - 1) RCHK(layer_stats());
            // get_fields(t)
            if (demag_on()) RCHK(demag());
            a.t = t;
            a.dti = 1.0 / sqrt(dt * max_dmdt);
            a.diag = 1;
            RCHK(zero_partials());
            RCHK(grain<GM_FIELD1>(a, nullptr, t));
            HIPCHK(hipStreamSynchronize(c->stream));
            RCHK(read_partials(r));
            double nd = 1.e22;
            for (int k = 0; k < nl; k++) {
                const HostLayer& H = c->layers[k];
                const double hmag = sqrt(r.hmag[k]);
                const double fact1 = (1.0 + H.alpha * H.alpha) / ((1.0 + H.alpha) * H.gamma) * p->dt_frac;
                const double fact0 = fact1 != 0.0 ? 0.0 : 1.0;
                const double cand = fact1 / std::max(1.0, hmag) + fact0 * nd;
                nd = std::min(std::min(nd, p->dt_max), cand);
            }
            dt = nd;
            // push_m(t)
            for (;;) {
                const double dtt = dt * max_dmdt;
                a.dt = dtt; a.dts = dtt; a.sqrt_dt = sqrt(dtt); a.diag = 0;
                RCHK(grain<GM_PUSH1>(a, nullptr, t + dtt));
                if (demag_on()) RCHK(demag());
                a.t = t + dtt; a.diag = 1;
                RCHK(zero_partials());
                RCHK(grain<GM_HEUN2>(a, mc.p, t + dtt));
                HIPCHK(hipStreamSynchronize(c->stream));
                RCHK(read_partials(r));
                bool too_far = false;
                max_dmdtnext = 1.e22;
                for (int k = 0; k < nl; k++) {
                    const double rot = acos(1.0 - 0.5 * r.errmx1[k]) / max_dm;
                    if (rot > 1.0) {
                        max_dmdt = 0.9 * max_dmdt / rot;
                        too_far = true;
                        break;
                    }
                    if (rot < 2.0 / 3.0) max_dmdtnext = std::min(max_dmdtnext, std::min(2.0, max_dmdt / (rot + 1.e-200)));
                    else max_dmdtnext = std::min(max_dmdtnext, max_dmdt);
                }
                if (!too_far) break;
                rejected++;
                if (rejected > 1000000) return set_err(MRMG_ERR_STATE, "push_m: no acceptable step");
            }
            ma.swap_with(mc);
            if (max_dmdtnext > 1.e20) max_dmdtnext = max_dmdt;
            dt = dt * max_dmdt;
            max_dmdt = max_dmdtnext;
            t = t + dt;
            if (sw_on) RCHK(switch_mark((long long)a.step, t));
        }
        if (diag) {
            memset(diag, 0, sizeof(*diag));
            fill_diag(diag, r, t, dt, p->n_steps);
            if (p->n_steps > 0) RCHK(read_stats(diag));
            diag->rejected = rejected;
            diag->max_dmdt = max_dmdt;
        }
        return MRMG_OK;
    }

    void fill_diag(mrmg_diag* d, const LayerRed& r, double t, double dt, int steps)
    {
        d->t = t; d->dt = dt; d->steps = steps; d->n_lay = c->nl;
        if (steps <= 0) return;
        double e = 0, e1 = 0;
        for (int k = 0; k < c->nl; k++) {
            e = std::max(e, sqrt(r.errmx[k]));
            e1 = std::max(e1, r.errmx1[k]);
            double sv = 0;
            for (int n = 0; n < c->layers[k].ng; n++) sv += c->layers[k].vol[n];
            d->mmag += r.mmag[k] / sv;
        }
        d->errmx = e;
        d->errmx1 = e1;
    }

    int step(const mrmg_step_params* p, const double* noise, mrmg_diag* diag) override
    {
        const int type = p->type;
        if (!(type == -3 || type == 2 || type == -2 || type == 1 || type == -1))
            return set_err(MRMG_ERR_UNSUPPORTED, "llg_type %d not implemented", type);
        if (p->dt <= 0.0) return set_err(MRMG_ERR_ARG, "dt must be > 0");
        if (p->noise == MRMG_NOISE_INJECT && !noise) return set_err(MRMG_ERR_ARG, "noise array is NULL");
        if (p->dt_frac != 0.0) {
            if (type != -3) return set_err(MRMG_ERR_UNSUPPORTED, "adaptive dt is implemented for Heun (llg_type -3) only");
            if (p->max_dm <= 0.0 || p->dt_max <= 0.0) return set_err(MRMG_ERR_ARG, "adaptive dt needs max_dm > 0, dt_max > 0");
        }
        if (p->llb) {
            for (int k = 0; k < c->nl; k++)
                if (!(c->layers[k].spin > 0.0))
                    return set_err(MRMG_ERR_ARG, "layer %d: LLB needs spin > 0", k);
        }
        cur_llb = p->llb != 0;
        RCHK(sync_aux());
        if (!cur_llb && mag_dirty) RCHK(refresh_sigma());
        const int rc = p->dt_frac != 0.0 ? step_adaptive(p, noise, diag) : step_fixed(p, noise, diag);
        if (cur_llb) mag_dirty = true;
        cur_llb = false;
        return rc;
    }

    int step_fixed(const mrmg_step_params* p, const double* noise, mrmg_diag* diag)
    {
        const int type = p->type;
        // llg() start: init_applied zeroes hthm (field.f90:703)
        HIPCHK(hipMemsetAsync(hthm.p, 0, hthm.bytes(), c->stream));
        // the cells must hold the current state before the first field evaluation
        if (demag_on()) RCHK(scatter(ma.p, &p->t0));
        // injected normals go up in chunks of steps; Philox normals are drawn a chunk of
        // steps ahead by one grid-wide kernel (the same values the grain kernel would draw)
        // and then read like injected ones, which takes the draw off the grain kernel
        const bool pre = p->noise == MRMG_NOISE_PHILOX && p->n_steps > 1;
        const int chunk = pre ? 16 : 64;
        const size_t per = 3 * (size_t)G;
        if (p->noise == MRMG_NOISE_INJECT || pre) {
            const size_t need = per * std::min(chunk, std::max(p->n_steps, 1));
            if (noise_buf.n < need) RCHK(noise_buf.alloc(need));
        }
        double t = p->t0;
        // get_fields sets llgp%dt = min(dt, dt_max, dt) with a fixed step (fact1 = 0,
        // fact0 = 1, llg.f90:383-386); dt_max <= 0 means no bound here
        const double dt = p->dt_max > 0.0 ? std::min(p->dt, p->dt_max) : p->dt;
        StepArgs a{};
        a.llb = p->llb;
        a.dt = dt;
        a.sqrt_dt = sqrt(dt);
        a.type = type;
        a.noise = pre ? MRMG_NOISE_INJECT : p->noise;
        a.seed = p->seed;
        for (int it = 0; it < p->n_steps; it++) {
            const bool last = it == p->n_steps - 1;
            a.step = p->step0 + it;
            a.normals = nullptr;
            if (p->noise == MRMG_NOISE_INJECT || pre) {
                const int ci = it % chunk;
                if (ci == 0) {
                    const int n = std::min(chunk, p->n_steps - it);
                    if (pre) {
                        const size_t cnt = (size_t)n * G;
                        hipLaunchKernelGGL(k_philox_dump, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream,
                                           (unsigned long long)p->seed, (unsigned long long)(p->step0 + it), n, G,
                                           noise_buf.p);
                        HIPCHK(hipGetLastError());
                    } else {
                        HIPCHK(hipMemcpyAsync(noise_buf.p, noise + per * it, sizeof(double) * per * n,
                                              hipMemcpyHostToDevice, c->stream));
                    }
                }
                a.normals = noise_buf.p + per * ci;
            }
            if (last) RCHK(zero_partials());
            if (last && diag) RCHK(layer_stats());
            if (type == -3) {
                // stage 1 (get_fields at t, push iter=2)
                if (demag_on()) RCHK(demag());
                a.t = t; a.dts = dt; a.diag = 0;
                RCHK(grain<GM_HEUN1>(a, nullptr, t + dt));
                // stage 2 (get_field at t+dt on the predicted m, push iter=1)
                if (demag_on()) RCHK(demag());
                a.t = t + dt; a.diag = last;
                RCHK(grain<GM_HEUN2>(a, nullptr, t + dt));
            } else if (type < 0) {
                if (demag_on()) RCHK(demag());
                a.t = t; a.dts = dt * 0.5; a.diag = 0;
                RCHK(grain<GM_ROT1>(a, nullptr, t + dt * 0.5));
                if (demag_on()) RCHK(demag());
                a.t = t + dt * 0.5; a.dts = dt; a.diag = last;
                RCHK(grain<GM_ROT2>(a, nullptr, t + dt));
            } else {
                if (demag_on()) RCHK(demag());
                a.t = t; a.dts = dt; a.diag = last;
                RCHK(grain<GM_ROT1>(a, nullptr, t + dt));
                ma.swap_with(mb);
            }
            t = t + dt;
            if (sw_on) RCHK(switch_mark((long long)a.step, t));
            if (prof && ev_used >= 512) {
                HIPCHK(hipStreamSynchronize(c->stream));
                RCHK(ev_collect());
            }
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (prof) RCHK(ev_collect());
        if (diag) {
            memset(diag, 0, sizeof(*diag));
            LayerRed r;
            if (p->n_steps > 0) RCHK(read_partials(r));
            else { r.errmx.assign(c->nl, 0); r.errmx1.assign(c->nl, 0); r.mmag.assign(c->nl, 0); }
            fill_diag(diag, r, t, dt, p->n_steps);
            if (p->n_steps > 0) RCHK(read_stats(diag));
            diag->max_dmdt = 1.0;
        }
        return MRMG_OK;
    }

    int profile(int enable) override
    {
        prof = enable != 0;
        ev_used = 0;
        for (int i = 0; i < 8; i++) { kt_ms[i] = 0; kt_n[i] = 0; }
        return MRMG_OK;
    }
    int kernel_stats(mrmg_kstat* out, int max, int* n) override
    {
        static const char* names[K_NCLASS] = {"row_fwd", "col", "row_inv", "grain1", "grain2",
                                              "col_fwd", "prod", "col_inv"};
        *n = K_NCLASS;
        for (int i = 0; i < K_NCLASS && i < max; i++) {
            memset(out[i].name, 0, sizeof(out[i].name));
            strncpy(out[i].name, names[i], sizeof(out[i].name) - 1);
            out[i].launches = kt_n[i];
            out[i].total_ms = kt_ms[i];
        }
        return MRMG_OK;
    }

    int readback(const mrmg_sensor* sn, int n_pos, const double* xpos, const double* ypos, double* out) override
    {
        const int nx = c->nx, ny = c->ny, nl = c->nl, nxy = nx * ny;
        const size_t CP = (size_t)nl * nxy;
        RbWeights wk{};
        for (int k = 0; k < nl; k++) wk.w[k] = (sn->w ? sn->w[k] : 1.0) * c->dz[k] * c->dx * c->dy;
        if (rb_plane.n < (size_t)nxy) RCHK(rb_plane.alloc(nxy));
        if (rb_pos.n < 2 * (size_t)n_pos) RCHK(rb_pos.alloc(2 * (size_t)n_pos));
        if (rb_out.n < (size_t)n_pos) RCHK(rb_out.alloc(n_pos));
        std::vector<double> pos(2 * (size_t)n_pos);
        memcpy(pos.data(), xpos, sizeof(double) * n_pos);
        memcpy(pos.data() + n_pos, ypos, sizeof(double) * n_pos);
        HIPCHK(hipMemcpyAsync(rb_pos.p, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice, c->stream));
        RCHK(scatter(ma.p));   // the cells hold the current state
        hipLaunchKernelGGL((k_rb_plane<T>), dim3((nxy + 255) / 256), dim3(256), 0, c->stream, Mcell.p, CP, nxy, nl,
                           sn->dir[0], sn->dir[1], sn->dir[2], wk, rb_plane.p);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_rb_scan, dim3(n_pos), dim3(256), sizeof(double) * (nx + ny), c->stream, rb_plane.p, nx,
                           ny, c->dx, c->dy, sn->sigma_x, sn->sigma_y, rb_pos.p, n_pos, rb_out.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(out, rb_out.p, sizeof(double) * n_pos, hipMemcpyDeviceToHost));
        return MRMG_OK;
    }
    int switch_arm(int enable, const double* dir, double thr) override
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
    int switch_mark(long long step, double t)
    {
        hipLaunchKernelGGL((k_switch<S>), dim3((G + 255) / 256), dim3(256), 0, c->stream, gdata(), ma.p, sw_dir[0],
                           sw_dir[1], sw_dir[2], sw_thr, step, t, sw_step.p, sw_t.p, sw_h.p);
        HIPCHK(hipGetLastError());
        return MRMG_OK;
    }
    int switch_read(int k, int64_t* step, double* t, double* hap) override
    {
        if (sw_step.n < (size_t)G) return set_err(MRMG_ERR_STATE, "switching record was never armed");
        const int ng = c->layers[k].ng;
        HIPCHK(hipStreamSynchronize(c->stream));
        if (step) {
            std::vector<long long> v(ng);
            HIPCHK(hipMemcpy(v.data(), sw_step.p + off[k], sizeof(long long) * ng, hipMemcpyDeviceToHost));
            for (int n = 0; n < ng; n++) step[n] = (int64_t)v[n];
        }
        if (t) HIPCHK(hipMemcpy(t, sw_t.p + off[k], sizeof(double) * ng, hipMemcpyDeviceToHost));
        if (hap) {
            std::vector<double> v(3 * (size_t)ng);
            for (int cc = 0; cc < 3; cc++)
                HIPCHK(hipMemcpy(v.data() + (size_t)cc * ng, sw_h.p + (size_t)cc * G + off[k], sizeof(double) * ng,
                                 hipMemcpyDeviceToHost));
            for (int n = 0; n < ng; n++)
                for (int cc = 0; cc < 3; cc++) hap[3 * (size_t)n + cc] = v[(size_t)cc * ng + n];
        }
        return MRMG_OK;
    }

    int demag_field(double* h_cell, double* h_grain) override
    {
        const size_t CP = (size_t)c->nl * c->nx * c->ny;
        if (demag_on()) {
            RCHK(scatter(ma.p));
            RCHK(demag());
        } else {
            HIPCHK(hipMemsetAsync(Hcell.p, 0, Hcell.bytes(), c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (h_cell) {
            std::vector<T> t(3 * CP);
            HIPCHK(hipMemcpy(t.data(), Hcell.p, t.size() * sizeof(T), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < t.size(); i++) h_cell[i] = (double)t[i];
        }
        if (h_grain) {
            std::vector<T> t(3 * CP);
            HIPCHK(hipMemcpy(t.data(), Hcell.p, t.size() * sizeof(T), hipMemcpyDeviceToHost));
            int g = 0;
            for (int k = 0; k < c->nl; k++) {
                const HostLayer& H = c->layers[k];
                for (int n = 0; n < H.ng; n++, g++) {
                    double y[3] = {0, 0, 0};
                    for (int e = H.g2gi[n]; e < H.g2gi[n + 1]; e++) {
                        const size_t ci = k * (size_t)c->nx * c->ny + (size_t)(H.g2g[2 * e + 1] - 1) * c->nx + (H.g2g[2 * e] - 1);
                        for (int cc = 0; cc < 3; cc++) y[cc] += (double)t[cc * CP + ci];
                    }
                    const int cnt = H.g2gi[n + 1] - H.g2gi[n];
                    for (int cc = 0; cc < 3; cc++) h_grain[3 * g + cc] = y[cc] / cnt;
                }
            }
        }
        return MRMG_OK;
    }
};

// ------------------------------------------------------------------------- C ABI
// (the definitions take C linkage from their declarations in mrmg.h)

const char* mrmg_last_error(void) { return g_err.c_str(); }

int mrmg_device_count(int* n)
{
    if (!n) return set_err(MRMG_ERR_ARG, "null argument");
    HIPCHK(hipGetDeviceCount(n));
    return MRMG_OK;
}

int mrmg_create(const mrmg_geom* g, mrmg_ctx** out)
{
    if (!g || !out) return set_err(MRMG_ERR_ARG, "null argument");
    *out = nullptr;
    if (g->nx < 2 || g->ny < 1 || g->n_lay < 1 || g->n_lay > MRMG_MAX_LAYERS)
        return set_err(MRMG_ERR_ARG, "bad grid %d x %d x %d (n_lay <= %d)", g->nx, g->ny, g->n_lay, MRMG_MAX_LAYERS);
    if (g->precision != MRMG_FP64 && g->precision != MRMG_FP32) return set_err(MRMG_ERR_ARG, "bad precision");
    if (!g->dz) return set_err(MRMG_ERR_ARG, "dz is NULL");
    std::unique_ptr<mrmg_ctx> c(new mrmg_ctx());
    c->nx = g->nx; c->ny = g->ny; c->nl = g->n_lay; c->nxh = g->nx / 2 + 1;
    c->precision = g->precision; c->demag = g->demag;
    c->dx = g->dx; c->dy = g->dy;
    c->dz.assign(g->dz, g->dz + g->n_lay);
    c->il.assign(g->n_lay, 0.0);
    if (g->il_space) c->il.assign(g->il_space, g->il_space + g->n_lay);
    if (!factorize(c->nx, c->px) || !factorize(c->ny, c->py))
        return set_err(MRMG_ERR_UNSUPPORTED, "nx=%d, ny=%d must factor into 2, 3, 5", g->nx, g->ny);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return set_err(MRMG_ERR_HIP, "no HIP device");
    HIPCHK(hipStreamCreate(&c->stream));
    c->layers.resize(c->nl);
    if (c->precision == MRMG_FP64) c->eng.reset(new Engine<double>(c.get()));
    else c->eng.reset(new Engine<float>(c.get()));
    int rc = c->precision == MRMG_FP64 ? static_cast<Engine<double>*>(c->eng.get())->init_fft()
                                       : static_cast<Engine<float>*>(c->eng.get())->init_fft();
    if (rc != MRMG_OK) {
        c->eng.reset();
        (void)hipStreamDestroy(c->stream);
        return rc;
    }
    *out = c.release();
    return MRMG_OK;
}

int mrmg_destroy(mrmg_ctx* c)
{
    if (!c) return MRMG_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->eng.reset();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MRMG_OK;
}

static int ensure_built(mrmg_ctx* c)
{
    if (!c->dirty) return MRMG_OK;
    for (int k = 0; k < c->nl; k++)
        if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set (mrmg_set_layer)", k);
    RCHK(c->eng->build(c->layers));
    c->dirty = false;
    return MRMG_OK;
}

static int tensor_sym(const mrmg_ctx* c, const double* n)
{
    const size_t plane = (size_t)c->nx * c->ny;
    for (int pr = 0; pr < c->nl * c->nl; pr++)
        for (int a = 0; a < 3; a++)
            for (int b = a + 1; b < 3; b++) {
                const double* p = n + plane * (9 * (size_t)pr + 3 * a + b);
                const double* q = n + plane * (9 * (size_t)pr + 3 * b + a);
                if (memcmp(p, q, plane * sizeof(double)) != 0) return 0;
            }
    return 1;
}

template <typename T>
static int set_tensor_T(mrmg_ctx* c, const double* d_real, bool sym)
{
    Engine<T>* e = static_cast<Engine<T>*>(c->eng.get());
    RCHK(e->set_slots(sym));
    return e->set_tensor_planes(d_real);
}

int mrmg_set_tensor(mrmg_ctx* c, const double* n_real)
{
    if (!c || !n_real) return set_err(MRMG_ERR_ARG, "null argument");
    const size_t count = (size_t)c->nx * c->ny * 9 * c->nl * c->nl;
    const bool sym = tensor_sym(c, n_real);
    DBuf<double> d;
    RCHK(d.upload(n_real, count, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return c->precision == MRMG_FP64 ? set_tensor_T<double>(c, d.p, sym) : set_tensor_T<float>(c, d.p, sym);
}

int mrmg_build_tensor(mrmg_ctx* c, int num_images, double* n_real_out)
{
    return mrmg_build_tensor_images(c, num_images, 1.0, 1.0, 8.e-7, 40.e-7, n_real_out);
}

int mrmg_build_tensor_images(mrmg_ctx* c, int num_images, double sul_mu, double head_mu, double fly_height,
                             double space_to_sul, double* n_real_out)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (num_images < 0) return set_err(MRMG_ERR_ARG, "num_images < 0");
    const int nx = c->nx, ny = c->ny, nl = c->nl;
    const size_t plane = (size_t)nx * ny, count = plane * 9 * nl * nl;
    DBuf<double> d;
    RCHK(d.alloc(count));
    // image strengths (mu-1)/(mu+1) of the soft underlayer and the head (tensor.f90:125-126)
    const bool sul = sul_mu > 1.0, head = head_mu > 1.0;
    const double mus = sul ? (sul_mu - 1.0) / (sul_mu + 1.0) : 0.0;
    const double mush = head ? (head_mu - 1.0) / (head_mu + 1.0) : 0.0;
    // distance from the head plane to the centre of each layer (tensor.f90:141-147)
    std::vector<double> hdz(nl);
    hdz[0] = fly_height;
    for (int ks = 1; ks < nl; ks++) hdz[ks] = hdz[ks - 1] + c->dz[ks - 1] + c->il[ks];
    for (int ks = 0; ks < nl; ks++) hdz[ks] = hdz[ks] + 0.5 * c->dz[ks];
    // calculate_tensor layer loop (tensor.f90:148-200); space_to_sul cancels in the direct rz
    const dim3 gp((unsigned)((plane + 127) / 128)), bp(128);
    double zsb = space_to_sul;
    for (int ks = nl - 1; ks >= 0; ks--) {
        const double zst = zsb + c->dz[ks];
        double zob = space_to_sul;
        for (int ko = nl - 1; ko >= 0; ko--) {
            const double zot = zob + c->dz[ko];
            const double rz = 0.5 * (zst - zot + zsb - zob);
            const double norm = c->dx * c->dy * c->dz[ko] * nx * ny;
            double* blk = d.p + plane * 9 * ((size_t)ko + (size_t)nl * ks);
            hipLaunchKernelGGL(k_plane_to_plane, gp, bp, 0, c->stream, nx, ny, c->dx, c->dy, c->dz[ks], c->dz[ko], rz,
                               num_images, norm, blk, (sul || head) ? 1 : 0, 0.0, 0.0);
            HIPCHK(hipGetLastError());
            if (sul) {
                hipLaunchKernelGGL(k_plane_to_plane, gp, bp, 0, c->stream, nx, ny, c->dx, c->dy, c->dz[ks], c->dz[ko],
                                   0.5 * (-zst - zot - zsb - zob), num_images, norm, blk, 2, -mus, mus);
                HIPCHK(hipGetLastError());
            }
            if (head) {
                hipLaunchKernelGGL(k_plane_to_plane, gp, bp, 0, c->stream, nx, ny, c->dx, c->dy, c->dz[ks], c->dz[ko],
                                   hdz[ks] + hdz[ko], num_images, norm, blk, 2, -mush, mush);
                HIPCHK(hipGetLastError());
            }
            if (sul || head) {
                hipLaunchKernelGGL(k_tensor_div, dim3((unsigned)((plane * 9 + 255) / 256)), dim3(256), 0, c->stream,
                                   blk, plane * 9, norm);
                HIPCHK(hipGetLastError());
            }
            zob = zot + c->il[ko];
        }
        zsb = zst + c->il[ks];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> host(count);
    HIPCHK(hipMemcpy(host.data(), d.p, count * sizeof(double), hipMemcpyDeviceToHost));
    if (n_real_out) memcpy(n_real_out, host.data(), count * sizeof(double));
    const bool sym = tensor_sym(c, host.data());
    return c->precision == MRMG_FP64 ? set_tensor_T<double>(c, d.p, sym) : set_tensor_T<float>(c, d.p, sym);
}

template <typename V> static void copy_vec(std::vector<V>& dst, const V* src, size_t n)
{
    if (src) dst.assign(src, src + n);
    else dst.clear();
}

// before a rebuild: the device state of every layer back into the host copies
static int save_device_state(mrmg_ctx* c)
{
    if (c->dirty) return MRMG_OK;
    for (int q = 0; q < c->nl; q++) {
        HostLayer& H = c->layers[q];
        if (!H.set) continue;
        H.m.resize(3 * (size_t)H.ng);
        H.mag.resize(H.ng);
        RCHK(c->eng->get_state(q, H.m.data(), nullptr));
        RCHK(c->eng->get_mag(q, H.mag.data()));
    }
    return MRMG_OK;
}

int mrmg_set_layer(mrmg_ctx* c, int k, const mrmg_layer* in)
{
    if (!c || !in) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (in->ng < 1) return set_err(MRMG_ERR_ARG, "ng must be >= 1");
    if (!in->grid2grain || !in->grain2gridindex || !in->grain2grid || !in->grain_vol || !in->ms || !in->hk)
        return set_err(MRMG_ERR_ARG, "layer %d: required array is NULL", k);
    RCHK(save_device_state(c));
    HostLayer H;
    H.set = true;
    const int ng = in->ng;
    H.ng = ng;
    const size_t P = (size_t)c->nx * c->ny;
    copy_vec(H.grid2grain, in->grid2grain, P);
    copy_vec(H.g2gi, in->grain2gridindex, ng + 1);
    if (H.g2gi[0] != 0) return set_err(MRMG_ERR_ARG, "grain2gridindex(0) must be 0");
    for (int n = 0; n < ng; n++)
        if (H.g2gi[n + 1] < H.g2gi[n]) return set_err(MRMG_ERR_ARG, "grain2gridindex not monotone");
    H.ncell = H.g2gi[ng];
    if ((size_t)H.ncell > P) return set_err(MRMG_ERR_ARG, "more grain cells than grid cells");
    copy_vec(H.g2g, in->grain2grid, 2 * (size_t)H.ncell);
    copy_vec(H.vol, in->grain_vol, ng);
    copy_vec(H.ms, in->ms, ng);
    copy_vec(H.hk, in->hk, 2 * (size_t)ng);
    H.has_uni = in->uni_axis != nullptr;
    if (H.has_uni) copy_vec(H.uni, in->uni_axis, 3 * (size_t)ng);
    H.has_cub = in->cubic_axis != nullptr;
    if (H.has_cub) copy_vec(H.cub, in->cubic_axis, 9 * (size_t)ng);
    auto csr = [&](const int* idx, const int* gr, const double* w, std::vector<int>& vi, std::vector<int>& vg,
                   std::vector<double>& vw, bool& has) -> int {
        has = idx != nullptr;
        if (!has) return MRMG_OK;
        if (!gr || !w) return set_err(MRMG_ERR_ARG, "layer %d: exchange CSR incomplete", k);
        copy_vec(vi, idx, ng + 1);
        if (vi[0] != 0) return set_err(MRMG_ERR_ARG, "layer %d: exchange CSR offset(0) must be 0", k);
        for (int n = 0; n < ng; n++)
            if (vi[n + 1] < vi[n]) return set_err(MRMG_ERR_ARG, "layer %d: exchange CSR not monotone", k);
        copy_vec(vg, gr, vi[ng]);
        copy_vec(vw, w, vi[ng]);
        return MRMG_OK;
    };
    RCHK(csr(in->hxgrainindex, in->hxgrain, in->hx, H.hxi, H.hxg, H.hx, H.has_hx));
    RCHK(csr(in->uhxgrainindex, in->uhxgrain, in->uhx, H.uhxi, H.uhxg, H.uhx, H.has_uhx));
    RCHK(csr(in->lhxgrainindex, in->lhxgrain, in->lhx, H.lhxi, H.lhxg, H.lhx, H.has_lhx));
    H.ms_scale.assign(ng, 1.0); if (in->ms_scale) copy_vec(H.ms_scale, in->ms_scale, ng);
    H.alpha_scale.assign(ng, 1.0); if (in->alpha_scale) copy_vec(H.alpha_scale, in->alpha_scale, ng);
    H.hk_scale.assign(ng, 1.0); if (in->hk_scale) copy_vec(H.hk_scale, in->hk_scale, ng);
    H.temp.assign(ng, 0.0); if (in->temp) copy_vec(H.temp, in->temp, ng);
    H.has_happ = in->h_app != nullptr;
    if (H.has_happ) copy_vec(H.happ, in->h_app, 3 * (size_t)ng);
    H.alpha = in->alpha; H.gamma = in->gamma; H.thm_alpha = in->thm_alpha; H.floor_ms = in->floor_ms;
    H.thermal = in->thermal;
    // LLB parameters: Tc defaults to 1e22 (data.f90:201), lambdaMs to
    // 3 Tc / (g muB/kB (S+1)) (grain.f90:61)
    H.spin = in->spin;
    H.tc.assign(ng, 1.e22); if (in->tc) copy_vec(H.tc, in->tc, ng);
    H.lambda_ms.resize(ng);
    for (int n = 0; n < ng; n++)
        H.lambda_ms[n] = in->lambda_ms ? in->lambda_ms[n] : 3.0 * H.tc[n] / (1.343427768e-4 * (H.spin + 1.0));
    // find_temp_scaling (temp_scaling.f90:365-431): nothing when tc_nom > 1e10; the
    // power laws of the (static) grain temperature are applied here, the |m| form at
    // every field evaluation on the device
    H.use_mag = false;
    H.tscale = in->tc_nom <= 1.e10;
    H.msexp = in->msexp;
    H.hkexp = in->hkexp;
    {
        const int nt[3] = {in->n_ms_tab, in->n_alpha_tab, in->n_hk_tab};
        const double* tt[3] = {in->ms_tab_t, in->alpha_tab_t, in->hk_tab_t};
        const double* ts[3] = {in->ms_tab_s, in->alpha_tab_s, in->hk_tab_s};
        for (int i = 0; i < 3; i++) {
            if (nt[i] == 0) continue;
            if (nt[i] < 2 || !tt[i] || !ts[i]) return set_err(MRMG_ERR_ARG, "scaling table %d: need >= 2 nodes", i);
            for (int j = 1; j < nt[i]; j++)
                if (!(tt[i][j] > tt[i][j - 1])) return set_err(MRMG_ERR_ARG, "scaling table %d: T/Tc not increasing", i);
            H.tab_t[i].assign(tt[i], tt[i] + nt[i]);
            H.tab_s[i].assign(ts[i], ts[i] + nt[i]);
        }
    }
    if (in->tc_nom <= 1.e10) {
        H.use_mag = in->use_mag != 0;
        auto cl = [](double x) { return x < 1.0 ? (x > 0.0 ? x : 0.0) : 1.0; };
        auto tab = [&](int i, int n) {
            return table_eval(H.tab_t[i].data(), H.tab_s[i].data(), (int)H.tab_t[i].size(), H.temp[n] / H.tc[n]);
        };
        for (int n = 0; n < ng; n++) {
            if (!H.use_mag) {
                if (in->msexp >= 0.0) H.ms_scale[n] = pow(cl(1.0 - H.temp[n] / H.tc[n]), in->msexp);
                else if (!H.tab_t[0].empty()) H.ms_scale[n] = tab(0, n);
            }
            if (!H.tab_t[1].empty()) H.alpha_scale[n] = tab(1, n);
            if (in->hkexp >= 0.0) H.hk_scale[n] = pow(cl(1.0 - H.temp[n] / H.tc[n]), in->hkexp);
            else if (!H.tab_t[2].empty()) H.hk_scale[n] = tab(2, n);
        }
    }
    // init_data_one: m = +z (data.f90:201-202), |m| = 1 (grain.f90:61)
    H.m.assign(3 * (size_t)ng, 0.0);
    for (int n = 0; n < ng; n++) H.m[3 * n + 2] = 1.0;
    H.mag.assign(ng, 1.0);
    H.uf = c->layers[k].uf;
    H.has_uf = c->layers[k].has_uf;
    for (int sl = 0; sl < 2; sl++) {
        H.src[sl] = c->layers[k].src[sl];
        H.src_x[sl] = c->layers[k].src_x[sl];
        H.src_y[sl] = c->layers[k].src_y[sl];
        H.src_v[sl] = c->layers[k].src_v[sl];
    }
    c->layers[k] = std::move(H);
    c->dirty = true;
    return MRMG_OK;
}

int mrmg_set_ufield(mrmg_ctx* c, int k, const mrmg_ufield* uf)
{
    if (!c || !uf) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    RCHK(check_tfun(uf->func, uf->t_start, uf->t_end, uf->amp, uf->gamma, uf->param));
    c->layers[k].uf = *uf;
    c->layers[k].has_uf = true;
    if (!c->dirty) return c->eng->update_ufield(c->layers);
    return MRMG_OK;
}

int mrmg_set_source(mrmg_ctx* c, int k, int slot, const mrmg_source* src)
{
    if (!c || !src) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (slot != 0 && slot != 1) return set_err(MRMG_ERR_ARG, "source slot must be 0 (field) or 1 (temperature)");
    const int kind = src->kind;
    if (kind != MRMG_SRC_NONE) {
        if (slot == 0 && kind != MRMG_SRC_FIELD_TABLE)
            return set_err(MRMG_ERR_ARG, "slot 0 takes a field table (MRMG_SRC_FIELD_TABLE)");
        if (slot == 1 && kind != MRMG_SRC_TEMP_TABLE && kind != MRMG_SRC_TEMP_SG)
            return set_err(MRMG_ERR_ARG, "slot 1 takes a temperature table or super-Gaussian");
        RCHK(check_tfun(src->func, src->t_start, src->t_end, src->amp, src->gamma, src->param));
        if (kind == MRMG_SRC_TEMP_SG) {
            if (src->expn < 1) return set_err(MRMG_ERR_ARG, "super-Gaussian power must be >= 1");
        } else {
            if (src->nxp < 2 || src->nyp < 2 || !src->xpos || !src->ypos || !src->vals)
                return set_err(MRMG_ERR_ARG, "a table needs at least 2 x 2 nodes and its arrays");
            for (int i = 1; i < src->nxp; i++)
                if (!(src->xpos[i] > src->xpos[i - 1])) return set_err(MRMG_ERR_ARG, "table x nodes must increase");
            for (int i = 1; i < src->nyp; i++)
                if (!(src->ypos[i] > src->ypos[i - 1])) return set_err(MRMG_ERR_ARG, "table y nodes must increase");
        }
    }
    RCHK(save_device_state(c));
    HostLayer& H = c->layers[k];
    H.src[slot] = *src;
    H.src_x[slot].clear(); H.src_y[slot].clear(); H.src_v[slot].clear();
    if (kind == MRMG_SRC_FIELD_TABLE || kind == MRMG_SRC_TEMP_TABLE) {
        const int nv = kind == MRMG_SRC_FIELD_TABLE ? 3 : 1;
        H.src_x[slot].assign(src->xpos, src->xpos + src->nxp);
        H.src_y[slot].assign(src->ypos, src->ypos + src->nyp);
        H.src_v[slot].assign(src->vals, src->vals + (size_t)nv * src->nxp * src->nyp);
    }
    H.src[slot].xpos = H.src[slot].ypos = H.src[slot].vals = nullptr;
    c->dirty = true;
    return MRMG_OK;
}

int mrmg_set_state(mrmg_ctx* c, int k, const double* m)
{
    if (!c || !m) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        c->layers[k].m.assign(m, m + 3 * (size_t)c->layers[k].ng);
        return MRMG_OK;
    }
    return c->eng->set_state(k, m);
}

int mrmg_set_mag(mrmg_ctx* c, int k, const double* mag)
{
    if (!c || !mag) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        c->layers[k].mag.assign(mag, mag + c->layers[k].ng);
        return MRMG_OK;
    }
    return c->eng->set_mag(k, mag);
}

int mrmg_get_mag(mrmg_ctx* c, int k, double* mag)
{
    if (!c || !mag) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        memcpy(mag, c->layers[k].mag.data(), sizeof(double) * c->layers[k].ng);
        return MRMG_OK;
    }
    return c->eng->get_mag(k, mag);
}

int mrmg_get_state(mrmg_ctx* c, int k, double* m, double* h)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        const HostLayer& H = c->layers[k];
        if (m) memcpy(m, H.m.data(), sizeof(double) * 3 * H.ng);
        if (h) memset(h, 0, sizeof(double) * 3 * H.ng);
        return MRMG_OK;
    }
    return c->eng->get_state(k, m, h);
}

int mrmg_step(mrmg_ctx* c, const mrmg_step_params* p, const double* noise, mrmg_diag* diag)
{
    if (!c || !p) return set_err(MRMG_ERR_ARG, "null argument");
    if (p->n_steps < 0) return set_err(MRMG_ERR_ARG, "n_steps < 0");
    if (c->demag && !c->eng->has_tensor())
        return set_err(MRMG_ERR_STATE, "demag enabled but no tensor (mrmg_set_tensor / mrmg_build_tensor)");
    RCHK(ensure_built(c));
    return c->eng->step(p, noise, diag);
}

int mrmg_philox_normals(mrmg_ctx* c, uint64_t seed, uint64_t step0, int n_steps, double* out)
{
    if (!c || !out) return set_err(MRMG_ERR_ARG, "null argument");
    RCHK(ensure_built(c));
    int G = 0;
    for (auto& L : c->layers) G += L.ng;
    const size_t n = (size_t)n_steps * G;
    DBuf<double> d;
    RCHK(d.alloc(3 * n));
    hipLaunchKernelGGL(k_philox_dump, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       (unsigned long long)seed, (unsigned long long)step0, n_steps, G, d.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, d.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    return MRMG_OK;
}

int mrmg_demag_field(mrmg_ctx* c, double* h_cell, double* h_grain)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (c->demag && !c->eng->has_tensor()) return set_err(MRMG_ERR_STATE, "no tensor");
    RCHK(ensure_built(c));
    return c->eng->demag_field(h_cell, h_grain);
}

int mrmg_readback(mrmg_ctx* c, const mrmg_sensor* s, int n_pos, const double* xpos, const double* ypos,
                  double* out)
{
    if (!c || !s) return set_err(MRMG_ERR_ARG, "null argument");
    if (n_pos < 0) return set_err(MRMG_ERR_ARG, "n_pos < 0");
    if (n_pos > 0 && (!xpos || !ypos || !out)) return set_err(MRMG_ERR_ARG, "null position or output array");
    if (!(s->sigma_x > 0.0) || !(s->sigma_y > 0.0)) return set_err(MRMG_ERR_ARG, "sensor widths must be > 0");
    if ((size_t)(c->nx + c->ny) * sizeof(double) > 64 * 1024 - 64)   // with the kernel's static reduction array
        return set_err(MRMG_ERR_UNSUPPORTED, "readback: nx + ny too large for one workgroup's table");
    RCHK(ensure_built(c));
    if (n_pos == 0) return MRMG_OK;
    return c->eng->readback(s, n_pos, xpos, ypos, out);
}

int mrmg_switch_arm(mrmg_ctx* c, int enable, const double* dir, double thr)
{
    if (!c || (enable && !dir)) return set_err(MRMG_ERR_ARG, "null argument");
    RCHK(ensure_built(c));
    return c->eng->switch_arm(enable, dir, thr);
}

int mrmg_switch_read(mrmg_ctx* c, int k, int64_t* step, double* t, double* h_app)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    RCHK(ensure_built(c));
    return c->eng->switch_read(k, step, t, h_app);
}

int mrmg_profile(mrmg_ctx* c, int enable)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->profile(enable);
}

int mrmg_tune(mrmg_ctx* c, const char* key, int value)
{
    if (!c || !key) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->tune(key, value);
}

int mrmg_get_tune(mrmg_ctx* c, const char* key, int* value)
{
    if (!c || !key || !value) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->get_tune(key, value);
}

int mrmg_kernel_stats(mrmg_ctx* c, mrmg_kstat* out, int max, int* n)
{
    if (!c || !out || !n) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->kernel_stats(out, max, n);
}

int mrmg_time_kernel(mrmg_ctx* c, const char* name, int reps, double* ms)
{
    if (!c || !name || !ms) return set_err(MRMG_ERR_ARG, "null argument");
    if (reps < 1) return set_err(MRMG_ERR_ARG, "reps must be >= 1");
    static const char* names[] = {"row_fwd", "col", "row_inv", "", "", "col_fwd", "prod", "col_inv"};
    int cls = -1;
    for (int i = 0; i < 8; i++)
        if (names[i][0] && strcmp(names[i], name) == 0) cls = i;
    if (cls < 0) return set_err(MRMG_ERR_ARG, "unknown demag kernel class '%s'", name);
    RCHK(ensure_built(c));
    return c->eng->time_class(cls, reps, ms);
}

int mrmg_copy_probe(mrmg_ctx* c, size_t bytes, int reps, double* ms)
{
    if (!c || !ms || reps < 1) return set_err(MRMG_ERR_ARG, "bad argument");
    const size_t n = bytes / 16;
    if (n == 0) return set_err(MRMG_ERR_ARG, "bytes < 16");
    DBuf<uint4> a, b;
    RCHK(a.alloc(n));
    RCHK(b.alloc(n));
    HIPCHK(hipMemsetAsync(a.p, 1, n * 16, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c->stream, a.p, b.p, n);   // warm
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c->stream, a.p, b.p, n);
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MRMG_OK;
}

int mrmg_footprint(mrmg_ctx* c, size_t* bytes)
{
    if (!c || !bytes) return set_err(MRMG_ERR_ARG, "null argument");
    *bytes = c->eng->footprint();
    return MRMG_OK;
}

