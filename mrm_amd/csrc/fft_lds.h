// fft_lds.h - workgroup-cooperative complex FFTs held in LDS, for gfx950.
//
// Data sit in LDS as separate real and imaginary arrays (8-byte lanes for FP64, 4-byte
// for FP32), B independent transforms of length N back to back.  Each pass is a
// Stockham autosort radix-R step (R = 2, 3, 4, 5): every butterfly reads its R inputs
// and twiddles into registers, the workgroup synchronises, then writes the R outputs to
// their self-sorted positions in the same arrays.  Twiddles w_N^e = exp(-2*pi*i*e/N)
// come from a table generated on the host in FP64 (rounded once for FP32).
// Transforms are unnormalised, forward = exp(-i...), inverse = exp(+i...), the FFTW
// convention the reference relies on (fftw.f90:37-93, demag.f90:259-267).
#pragma once
#include <hip/hip_runtime.h>

#define MRMG_MAX_FAC 24

struct FFTPlan {
    int n;
    int nfac;
    int fac[MRMG_MAX_FAC];
};

template <typename T> struct C2 { T x, y; };

template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// in-register DFT of length R; INV selects exp(+2*pi*i/R)
template <typename T, int R, bool INV> struct Dft;

template <typename T, bool INV> struct Dft<T, 2, INV> {
    __device__ __forceinline__ static void run(C2<T>* v)
    {
        C2<T> a = v[0], b = v[1];
        v[0] = {a.x + b.x, a.y + b.y};
        v[1] = {a.x - b.x, a.y - b.y};
    }
};

template <typename T, bool INV> struct Dft<T, 4, INV> {
    __device__ __forceinline__ static void run(C2<T>* v)
    {
        C2<T> s0 = {v[0].x + v[2].x, v[0].y + v[2].y};
        C2<T> s1 = {v[0].x - v[2].x, v[0].y - v[2].y};
        C2<T> s2 = {v[1].x + v[3].x, v[1].y + v[3].y};
        C2<T> s3 = {v[1].x - v[3].x, v[1].y - v[3].y};
        v[0] = {s0.x + s2.x, s0.y + s2.y};
        v[2] = {s0.x - s2.x, s0.y - s2.y};
        if (!INV) {  // y1 = s1 - i s3, y3 = s1 + i s3
            v[1] = {s1.x + s3.y, s1.y - s3.x};
            v[3] = {s1.x - s3.y, s1.y + s3.x};
        } else {
            v[1] = {s1.x - s3.y, s1.y + s3.x};
            v[3] = {s1.x + s3.y, s1.y - s3.x};
        }
    }
};

template <typename T, bool INV> struct Dft<T, 3, INV> {
    __device__ __forceinline__ static void run(C2<T>* v)
    {
        const T h = (T)0.86602540378443864676;  // sin(2 pi/3)
        C2<T> t1 = {v[1].x + v[2].x, v[1].y + v[2].y};
        C2<T> t2 = {v[1].x - v[2].x, v[1].y - v[2].y};
        C2<T> m = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
        v[0] = {v[0].x + t1.x, v[0].y + t1.y};
        if (!INV) {
            v[1] = {m.x + h * t2.y, m.y - h * t2.x};
            v[2] = {m.x - h * t2.y, m.y + h * t2.x};
        } else {
            v[1] = {m.x - h * t2.y, m.y + h * t2.x};
            v[2] = {m.x + h * t2.y, m.y - h * t2.x};
        }
    }
};

template <typename T, bool INV> struct Dft<T, 5, INV> {
    __device__ __forceinline__ static void run(C2<T>* v)
    {
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        C2<T> a1 = {v[1].x + v[4].x, v[1].y + v[4].y}, b1 = {v[1].x - v[4].x, v[1].y - v[4].y};
        C2<T> a2 = {v[2].x + v[3].x, v[2].y + v[3].y}, b2 = {v[2].x - v[3].x, v[2].y - v[3].y};
        C2<T> p1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        C2<T> p2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        C2<T> q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
        C2<T> q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
        v[0] = {v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
        if (!INV) {  // y1 = p1 - i q1, y4 = p1 + i q1, y2 = p2 - i q2, y3 = p2 + i q2
            v[1] = {p1.x + q1.y, p1.y - q1.x};
            v[4] = {p1.x - q1.y, p1.y + q1.x};
            v[2] = {p2.x + q2.y, p2.y - q2.x};
            v[3] = {p2.x - q2.y, p2.y + q2.x};
        } else {
            v[1] = {p1.x - q1.y, p1.y + q1.x};
            v[4] = {p1.x + q1.y, p1.y - q1.x};
            v[2] = {p2.x - q2.y, p2.y + q2.x};
            v[3] = {p2.x + q2.y, p2.y - q2.x};
        }
    }
};

// One Stockham pass over B transforms of length N with span Ns.
template <typename T, int R, int MAXB, bool INV>
__device__ __forceinline__ void fft_pass(T* re, T* im, int N, int Ns, int B, const C2<T>* __restrict__ tw)
{
    const int NR = N / R;
    const int total = B * NR;
    const int tstep = N / (Ns * R);
    C2<T> v[MAXB][R];
    int dst[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int idx = threadIdx.x + q * blockDim.x;
        dst[q] = -1;
        if (idx < total) {
            const int b = idx / NR;
            const int j = idx - b * NR;
            const int k = j % Ns;
            const int base = b * N;
#pragma unroll
            for (int r = 0; r < R; r++) {
                v[q][r].x = re[base + j + r * NR];
                v[q][r].y = im[base + j + r * NR];
            }
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = tw[k * r * tstep];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            Dft<T, R, INV>::run(v[q]);
            dst[q] = base + (j / Ns) * Ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                re[dst[q] + r * Ns] = v[q][r].x;
                im[dst[q] + r * Ns] = v[q][r].y;
            }
        }
    }
    __syncthreads();
}

// Full transform: every thread of the workgroup must call it (it synchronises).
template <typename T, int MAXB, bool INV>
__device__ void lds_fft(T* re, T* im, int B, const FFTPlan& plan, const C2<T>* __restrict__ tw)
{
    int Ns = 1;
    const int N = plan.n;
    for (int p = 0; p < plan.nfac; p++) {
        const int R = plan.fac[p];
        if (R == 4) fft_pass<T, 4, MAXB, INV>(re, im, N, Ns, B, tw);
        else if (R == 2) fft_pass<T, 2, MAXB, INV>(re, im, N, Ns, B, tw);
        else if (R == 3) fft_pass<T, 3, MAXB, INV>(re, im, N, Ns, B, tw);
        else fft_pass<T, 5, MAXB, INV>(re, im, N, Ns, B, tw);
        Ns *= R;
    }
}
