// tensor.hip - the FP64 demag tensor: Newell's volume-averaged blocks with periodic images
// (tensor_element.f90), their FFT and the packing into the stored spectra (demag.f90:200-286).
#include "engine.h"

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
                                 double fz = 0.0, int onx = 0, int ony = 0)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nx * ny) return;
    const int oi = p % nx, oj = p / nx;
    int is[2], js[2], ni = 0, nj = 0;
    if (onx > 0) {
        // open boundaries (nx = 2 onx, ny = 2 ony, nimg = 0): index oi holds the one
        // displacement -oi (oi < onx) or nx - oi (oi > onx) of the media grid; the middle
        // line oi = onx is no displacement of it and stays zero
        if (oi != onx) { is[0] = oi < onx ? -oi : nx - oi; ni = 1; }
        if (oj != ony) { js[0] = oj < ony ? -oj : ny - oj; nj = 1; }
    } else {
        if (nx % 2 == 0 && oi == nx / 2) { is[0] = -nx / 2; is[1] = nx / 2; ni = 2; }
        else { is[0] = (oi <= nx / 2) ? -oi : nx - oi; ni = 1; }
        if (ny % 2 == 0 && oj == ny / 2) { js[0] = -ny / 2; js[1] = ny / 2; nj = 2; }
        else { js[0] = (oj <= ny / 2) ? -oj : ny - oj; nj = 1; }
    }
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
    extern __shared__ __align__(16) unsigned char smem_raw[];
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
    extern __shared__ __align__(16) unsigned char smem_raw[];
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

// Regroup the tensor spectra into Tspec[kx][q][slot][ky] and round to T: q runs over the
// stored blocks (tensor_block) and real[q] is the real block (9 planes of TC) it comes from.
template <typename T>
__global__ void k_tensor_pack(const C2<double>* __restrict__ TC, C2<T>* __restrict__ Tspec, int ny, int nxh,
                              int nq, int nslot, const int* __restrict__ slot2c, const int* __restrict__ real)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nxh * nq * nslot * ny;
    if (i >= total) return;
    size_t r = i;
    const int ky = r % ny; r /= ny;
    const int s = r % nslot; r /= nslot;
    const int q = r % nq; r /= nq;
    const int kx = (int)r;
    const size_t plane = (size_t)slot2c[s] + 9 * (size_t)real[q];
    const C2<double> v = TC[plane * nxh * ny + (size_t)ky * nxh + kx];
    Tspec[i] = C2<T>{(T)v.x, (T)v.y};
}

// L1 distance of every layer-pair block (ko + nl ks, 9 planes of n elements each) of the
// real tensor from the block of its z-offset representative (ko', ks') = (max(d,0),
// max(d,0) - d), d = ko - ks: out[2p] += sum |R_p - R_rep|, out[2p+1] += sum |R_rep|.
__global__ void __launch_bounds__(256) k_block_l1(const double* __restrict__ R, size_t n, int nl,
                                                  double* __restrict__ out)
{
    __shared__ double sa[4], sb[4];
    const int p = blockIdx.y, ko = p % nl, ks = p / nl, d = ko - ks;
    const int rko = d > 0 ? d : 0, rks = rko - d;
    const double* a = R + (size_t)p * n;
    const double* b = R + (size_t)(rko + nl * rks) * n;
    double s0 = 0.0, s1 = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        s0 += fabs(a[i] - b[i]);
        s1 += fabs(b[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
    }
    if ((threadIdx.x & 63) == 0) { sa[threadIdx.x >> 6] = s0; sb[threadIdx.x >> 6] = s1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(out + 2 * p, sa[0] + sa[1] + sa[2] + sa[3]);
        atomicAdd(out + 2 * p + 1, sb[0] + sb[1] + sb[2] + sb[3]);
    }
}

// elements where component (a,b) differs from (b,a), over nblk blocks of 9 planes
__global__ void k_sym_count(const double* __restrict__ R, size_t plane, int nblk, unsigned long long* __restrict__ cnt)
{
    const size_t total = plane * nblk;
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / plane, e = i - b * plane;
        const double* q = R + 9 * plane * b + e;
        c += (q[1 * plane] != q[3 * plane]) + (q[2 * plane] != q[6 * plane]) + (q[5 * plane] != q[7 * plane]);
    }
    if (c) atomicAdd(cnt, c);
}

template <typename T>
int Engine<T>::set_tensor_planes(const double* d_real, bool toeplitz)
{
    // d_real: device FP64 real-space blocks of 9 (nx,ny) planes: the nl^2 layer pairs
    // (ko + nl ks, the (nx,ny,9,nl,nl) array of calculate_tensor), or with toeplitz the
    // 2 nl - 1 z offsets (block d + nl - 1 for d = ko - ks); planes of the FFT grid
    const int nx = c->fx, ny = c->fy, nxh = c->nxh, nl = c->nl;
    const int nq = toeplitz ? 2 * nl - 1 : nl * nl;
    const int P = 9 * nq;
    DBuf<C2<double>> TC, tdx, tdy;
    RCHK(TC.alloc((size_t)P * nxh * ny));
    auto wx = plan_twiddles(c->px), wy = plan_twiddles(c->py);
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
    // symmetry decides 6 or 9 slots (set_slots, from the caller's check of the real tensor)
    std::vector<int> real(nq);
    for (int q = 0; q < nq; q++) real[q] = toeplitz ? q : (q / nl) + nl * (q % nl);   // q = ko*nl + ks
    DBuf<int> d_real_of;
    RCHK(d_real_of.upload(real.data(), nq, c->stream));
    const size_t total = (size_t)nxh * nq * nslot * ny;
    RCHK(Tspec.alloc(total));
    hipLaunchKernelGGL(k_tensor_pack<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, TC.p,
                       Tspec.p, ny, nxh, nq, nslot, slot2c.p, d_real_of.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    toep = toeplitz;
    tensor_ok = true;
    return MRMG_OK;
}


int launch_plane_to_plane(hipStream_t s, int nx, int ny, double dx, double dy, double dzs, double dzo, double rz,
                          int nimg, double norm, double* blk, int mode, double fxy, double fz, int onx, int ony)
{
    const size_t plane = (size_t)nx * ny;
    hipLaunchKernelGGL(k_plane_to_plane, dim3((unsigned)((plane + 127) / 128)), dim3(128), 0, s, nx, ny, dx, dy, dzs,
                       dzo, rz, nimg, norm, blk, mode, fxy, fz, onx, ony);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

int launch_tensor_div(hipStream_t s, double* blk, size_t n, double norm)
{
    hipLaunchKernelGGL(k_tensor_div, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, blk, n, norm);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

int tensor_toeplitz_l1(hipStream_t s, const double* R, size_t block_elems, int nl, double* d_out)
{
    HIPCHK(hipMemsetAsync(d_out, 0, sizeof(double) * 2 * nl * nl, s));
    const unsigned gx = (unsigned)std::min<size_t>((block_elems + 255) / 256, 1024);
    hipLaunchKernelGGL(k_block_l1, dim3(gx, nl * nl), dim3(256), 0, s, R, block_elems, nl, d_out);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

int tensor_sym_count(hipStream_t s, const double* R, size_t plane, int nblk, unsigned long long* d_cnt)
{
    HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), s));
    const unsigned g = (unsigned)std::min<size_t>((plane * nblk + 255) / 256, 8192);
    hipLaunchKernelGGL(k_sym_count, dim3(g), dim3(256), 0, s, R, plane, nblk, d_cnt);
    HIPCHK(hipGetLastError());
    return MRMG_OK;
}

template int Engine<float>::set_tensor_planes(const double*, bool);
template int Engine<double>::set_tensor_planes(const double*, bool);
