// fft_lds.h - workgroup-cooperative complex FFTs held in LDS, for gfx950.
//
// Data sit in LDS as separate real and imaginary arrays (8-byte lanes for FP64, 4-byte
// for FP32), B independent transforms of length N back to back.  Each pass is a
// Stockham autosort radix-R step (R = 16, 8, 4, 2, 3, 5; a 256-point transform takes two
// radix-16 passes): every butterfly reads its R inputs
// and twiddles into registers, the workgroup synchronises, then writes the R outputs to
// their self-sorted positions in the same arrays.  Twiddles w_N^e = exp(-2*pi*i*e/N)
// come from a table generated on the host in FP64 (rounded once for FP32) and staged in
// LDS by the calling kernel.
// Transforms are unnormalised, forward = exp(-i...), inverse = exp(+i...), the FFTW
// convention the reference relies on (fftw.f90:37-93, demag.f90:259-267).
#pragma once
#include <hip/hip_runtime.h>

#define MRMG_MAX_FAC 24

// LDS index padding: one spare slot per 16 elements, so the stride-R writes of a
// Stockham pass (R = 8, 16) from 16 consecutive lanes land in distinct banks for 8-byte
// elements (and 32 lanes for 4-byte elements) instead of piling onto a few.
#define LP(i) ((i) + ((i) >> 4))
// elements of an LDS array holding n padded entries
#define LPN(n) ((n) + ((n) >> 4) + 1)

// Pass p has radix fac[p] and span ns[p] (the product of the earlier radices).  Its
// twiddles w_n^(k r tstep[p]), k < ns[p], 0 < r < fac[p], sit in the plan's twiddle table at
// toff[p] + (r-1) ns[p] + k: the lanes of a pass read consecutive entries (no LDS bank
// conflicts, coalesced global loads).  The divisions of a pass's index arithmetic use
// multipliers fixed when the plan is made:
// x / d == umulhi(x, mnr or mns) for the indices of one workgroup (x * d < 2^32), d == 1
// taken separately.  Reading them per pass also keeps the compiler from hoisting the
// index arithmetic of every radix branch out of the pass loop.
struct FFTPlan {
    int n;
    int nfac;
    int fac[MRMG_MAX_FAC];
    int ns[MRMG_MAX_FAC];
    int tstep[MRMG_MAX_FAC];        // twiddle stride n / (ns * fac)
    int toff[MRMG_MAX_FAC];         // the pass's twiddles in the plan's table
    unsigned mnr[MRMG_MAX_FAC];     // multiplier of n / fac
    unsigned mns[MRMG_MAX_FAC];     // multiplier of ns
};

__host__ __device__ inline unsigned fdiv_mul(int d) { return d > 1 ? (unsigned)((0x100000000ull + d - 1) / d) : 0u; }
__device__ __forceinline__ int fdiv(int x, int d, unsigned m) { return d == 1 ? x : (int)__umulhi((unsigned)x, m); }

// fill the per-pass fields of a plan whose n, nfac and fac are set
inline void plan_passes(FFTPlan& p)
{
    int ns = 1, off = 0;
    for (int i = 0; i < p.nfac; i++) {
        p.ns[i] = ns;
        p.tstep[i] = p.n / (ns * p.fac[i]);
        p.toff[i] = off;                      // n - fac[0] entries in all, fewer than n
        if (ns > 1) off += ns * (p.fac[i] - 1);
        p.mnr[i] = fdiv_mul(p.n / p.fac[i]);
        p.mns[i] = fdiv_mul(ns);
        ns *= p.fac[i];
    }
}

template <typename T> struct C2 { T x, y; };
template <typename T> struct C4 { C2<T> a, b; };

// Global accesses of the demag's streams (spectra, tensor: each element read or written
// once per pass, the arrays far larger than the caches), non-temporal when NT.
template <bool NT, typename T>
__device__ __forceinline__ C2<T> ld_s(const C2<T>* p)
{
    if (NT) return C2<T>{__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y)};
    return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_s(C2<T>* p, C2<T> v)
{
    if (NT) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
    } else {
        *p = v;
    }
}
template <bool NT, typename T>
__device__ __forceinline__ void st_s(C4<T>* p, C4<T> v)
{
    if (NT) {
        __builtin_nontemporal_store(v.a.x, &p->a.x);
        __builtin_nontemporal_store(v.a.y, &p->a.y);
        __builtin_nontemporal_store(v.b.x, &p->b.x);
        __builtin_nontemporal_store(v.b.y, &p->b.y);
    } else {
        *p = v;
    }
}

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

// w_16^e = exp(-2 pi i e/16), e = 0..15 (cos, sin); the forward twiddle is (c, -s)
template <typename T> struct W16 {
    __device__ __forceinline__ static C2<T> get(int e, bool inv)
    {
        const T c[16] = {(T)1, (T)0.92387953251128675613, (T)0.70710678118654752440, (T)0.38268343236508977173,
                         (T)0, (T)-0.38268343236508977173, (T)-0.70710678118654752440, (T)-0.92387953251128675613,
                         (T)-1, (T)-0.92387953251128675613, (T)-0.70710678118654752440, (T)-0.38268343236508977173,
                         (T)0, (T)0.38268343236508977173, (T)0.70710678118654752440, (T)0.92387953251128675613};
        const T s[16] = {(T)0, (T)0.38268343236508977173, (T)0.70710678118654752440, (T)0.92387953251128675613,
                         (T)1, (T)0.92387953251128675613, (T)0.70710678118654752440, (T)0.38268343236508977173,
                         (T)0, (T)-0.38268343236508977173, (T)-0.70710678118654752440, (T)-0.92387953251128675613,
                         (T)-1, (T)-0.92387953251128675613, (T)-0.70710678118654752440, (T)-0.38268343236508977173};
        return C2<T>{c[e & 15], inv ? s[e & 15] : -s[e & 15]};
    }
};

// Length 8 = 4 x 2 and 16 = 4 x 4 in registers: X[p + A r] = sum_q w_R^{qp} w_B^{qr} U_q[p],
// U_q = DFT_A of v[q + B m] over m (R = A B, A = 4).
template <typename T, bool INV> struct Dft<T, 8, INV> {
    __device__ __forceinline__ static void run(C2<T>* v)
    {
        C2<T> u[2][4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            C2<T> t[4] = {v[q], v[q + 2], v[q + 4], v[q + 6]};
            Dft<T, 4, INV>::run(t);
#pragma unroll
            for (int p = 0; p < 4; p++) u[q][p] = (q && p) ? cmul(t[p], W16<T>::get(2 * p, INV)) : t[p];
        }
#pragma unroll
        for (int p = 0; p < 4; p++) {
            C2<T> t[2] = {u[0][p], u[1][p]};
            Dft<T, 2, INV>::run(t);
            v[p] = t[0];
            v[p + 4] = t[1];
        }
    }
};

template <typename T, bool INV> struct Dft<T, 16, INV> {
    __device__ __forceinline__ static void run(C2<T>* v)
    {
        C2<T> u[4][4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            C2<T> t[4] = {v[q], v[q + 4], v[q + 8], v[q + 12]};
            Dft<T, 4, INV>::run(t);
#pragma unroll
            for (int p = 0; p < 4; p++) u[q][p] = (q && p) ? cmul(t[p], W16<T>::get(q * p, INV)) : t[p];
        }
#pragma unroll
        for (int p = 0; p < 4; p++) {
            C2<T> t[4] = {u[0][p], u[1][p], u[2][p], u[3][p]};
            Dft<T, 4, INV>::run(t);
#pragma unroll
            for (int r = 0; r < 4; r++) v[p + 4 * r] = t[r];
        }
    }
};

// Ordering point between the read and write halves of a wave-local pass: the wave's own
// LDS operations are complete (s_waitcnt lgkmcnt(0)) and the compiler moves no memory
// access across it.  Other waves are not involved.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0), vmcnt and expcnt left alone
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham pass over B transforms of length N with span Ns, by nt threads numbered t
// (the workgroup, or one wave when WAVE: then only that wave synchronises).
template <typename T, int R, int MAXB, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft_pass(T* re, T* im, int off, int N, int Ns, int B, const C2<T>* __restrict__ tw,
                                         int t, int nt, int toff, unsigned mnr, unsigned mns)
{
    const int NR = N / R;
    const int total = B * NR;
    C2<T> v[MAXB][R];
    int dst[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int idx = t + q * nt;
        dst[q] = -1;
        if (idx < total) {
            const int b = fdiv(idx, NR, mnr);
            const int j = idx - b * NR;
            const int jq = fdiv(j, Ns, mns);
            const int k = j - jq * Ns;
            const int base = off + b * N;
            if ((NR & 15) == 0) {
                // a stride that is a multiple of 16 moves the padded index by NR + NR/16
                const int l0 = LP(base + j), ls = NR + (NR >> 4);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    v[q][r].x = re[l0 + r * ls];
                    v[q][r].y = im[l0 + r * ls];
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    v[q][r].x = re[LP(base + j + r * NR)];
                    v[q][r].y = im[LP(base + j + r * NR)];
                }
            }
            if (Ns > 1) {
                const C2<T>* twk = tw + toff + k;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = twk[(r - 1) * Ns];   // w_N^(k r tstep)
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            Dft<T, R, INV>::run(v[q]);
            dst[q] = base + jq * Ns * R + k;
        }
    }
    if (WAVE) wave_sync();
    else __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        if (dst[q] >= 0) {
            if ((Ns & 15) == 0) {
                const int l0 = LP(dst[q]), ls = Ns + (Ns >> 4);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    re[l0 + r * ls] = v[q][r].x;
                    im[l0 + r * ls] = v[q][r].y;
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    re[LP(dst[q] + r * Ns)] = v[q][r].x;
                    im[LP(dst[q] + r * Ns)] = v[q][r].y;
                }
            }
        }
    }
    if (WAVE) wave_sync();
    else __syncthreads();
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// One pass of the power-of-two form (lds_fft_p2): fft_pass with the radix R, the span NS,
// the length N and the first radix R0 fixed at compile time.  The read stride N/R and the
// span are multiples of 16 (or the span is 1), so with the LDS padding one element's
// neighbours in a butterfly sit at constant distances from its own padded slot: a pass
// computes one base address per butterfly and reads, twiddles and writes through
// immediate offsets.  The arithmetic is fft_pass's, operation for operation.
template <typename T, int N, int R, int NS, int R0, int MAXB, bool INV>
__device__ __forceinline__ void fft_pass_p2(T* re, T* im, int off, int B, const C2<T>* __restrict__ tw, int t, int nt)
{
    constexpr int NR = N / R, LNR = ilog2c(NR), LNS = ilog2c(NS), LR = ilog2c(R), LN = ilog2c(N);
    static_assert(NR % 16 == 0 && (NS == 1 || NS % 16 == 0), "padded strides must be constant");
    constexpr int RS = NR + NR / 16, WS = NS == 1 ? 1 : NS + NS / 16, TOFF = NS == 1 ? 0 : NS - R0;
    const int total = B << LNR;
    C2<T> v[MAXB][R];
    int dst[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int idx = t + q * nt;
        dst[q] = -1;
        if (idx < total) {
            const int b = idx >> LNR, j = idx & (NR - 1);
            const int jq = j >> LNS, k = j & (NS - 1);
            const int base = off + (b << LN);
            const int l0 = LP(base + j);
#pragma unroll
            for (int r = 0; r < R; r++) {
                v[q][r].x = re[l0 + r * RS];
                v[q][r].y = im[l0 + r * RS];
            }
            if (NS > 1) {
                const C2<T>* twk = tw + TOFF + k;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = twk[(r - 1) * NS];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            Dft<T, R, INV>::run(v[q]);
            dst[q] = LP(base + (jq << (LNS + LR)) + k);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                re[dst[q] + r * WS] = v[q][r].x;
                im[dst[q] + r * WS] = v[q][r].y;
            }
        }
    }
    __syncthreads();
}

template <typename T, int N, int NS, int E, bool INV>
__device__ __forceinline__ void p2_passes(T* re, T* im, int off, int B, const C2<T>* __restrict__ tw, int t, int nt)
{
    if constexpr (NS < N) {
        constexpr int REM = N / NS, R = REM % 16 == 0 ? 16 : REM;
        fft_pass_p2<T, N, R, NS, 16, (E + R - 1) / R, INV>(re, im, off, B, tw, t, nt);
        p2_passes<T, N, NS * R, E, INV>(re, im, off, B, tw, t, nt);
    }
}

// lds_fft for the plans of factorize(n, 16) with n a power of two from 256 to 4096 (radix-16
// passes, then one of 8, 4 or 2): the same transform, bit for bit, with the index
// arithmetic fixed at compile time.  Returns false (and does nothing) for other plans or
// an offset that is not a multiple of 16; every thread of the workgroup takes the same
// branch.
template <typename T, int E, bool INV>
__device__ __forceinline__ bool lds_fft_p2(T* re, T* im, int B, const FFTPlan& plan, const C2<T>* __restrict__ tw,
                                           int off = 0)
{
    if ((off & 15) != 0 || plan.fac[0] != 16) return false;
    const int t = threadIdx.x, nt = blockDim.x;
    switch (plan.n) {
    case 256: if constexpr (256 <= E * 256) { p2_passes<T, 256, 1, E, INV>(re, im, off, B, tw, t, nt); return true; }
        break;
    case 512: if constexpr (512 <= E * 256) { p2_passes<T, 512, 1, E, INV>(re, im, off, B, tw, t, nt); return true; }
        break;
    case 1024: if constexpr (1024 <= E * 256) { p2_passes<T, 1024, 1, E, INV>(re, im, off, B, tw, t, nt); return true; }
        break;
    case 2048: if constexpr (2048 <= E * 256) { p2_passes<T, 2048, 1, E, INV>(re, im, off, B, tw, t, nt); return true; }
        break;
    case 4096: if constexpr (4096 <= E * 256) { p2_passes<T, 4096, 1, E, INV>(re, im, off, B, tw, t, nt); return true; }
        break;
    default: break;
    }
    return false;
}

// Full transform: every thread of the workgroup must call it (it synchronises).
// Element i of the B*N array sits at LDS slot LP(i).
// E bounds the complex elements one thread holds in a pass: B*N <= E*blockDim.x.
// off: element index of the first transform within the padded LDS arrays.
// P2: take lds_fft_p2 where it applies.
template <typename T, int E, bool INV, bool P2 = false>
__device__ __forceinline__ void lds_fft(T* re, T* im, int B, const FFTPlan& plan, const C2<T>* __restrict__ tw,
                                        int off = 0)
{
    if constexpr (P2)
        if (lds_fft_p2<T, E, INV>(re, im, B, plan, tw, off)) return;
    const int N = plan.n;
    const int t = threadIdx.x, nt = blockDim.x;
    for (int p = 0; p < plan.nfac; p++) {
        const int R = plan.fac[p], Ns = plan.ns[p], ts = plan.toff[p];
        const unsigned a = plan.mnr[p], c = plan.mns[p];
        if (R == 16) fft_pass<T, 16, (E + 15) / 16, INV>(re, im, off, N, Ns, B, tw, t, nt, ts, a, c);
        else if (R == 8) fft_pass<T, 8, (E + 7) / 8, INV>(re, im, off, N, Ns, B, tw, t, nt, ts, a, c);
        else if (R == 4) fft_pass<T, 4, (E + 3) / 4, INV>(re, im, off, N, Ns, B, tw, t, nt, ts, a, c);
        else if (R == 2) fft_pass<T, 2, (E + 1) / 2, INV>(re, im, off, N, Ns, B, tw, t, nt, ts, a, c);
        else if (R == 3) fft_pass<T, 3, (E + 2) / 3, INV>(re, im, off, N, Ns, B, tw, t, nt, ts, a, c);
        else fft_pass<T, 5, (E + 4) / 5, INV>(re, im, off, N, Ns, B, tw, t, nt, ts, a, c);
    }
}

// Wave-local form: the B transforms are dealt out to the waves in contiguous runs of
// ceil(B / waves), and each wave transforms its run alone, with no workgroup barrier
// (the runs are disjoint).  E bounds the elements one lane holds: ceil(B/waves)*N <= 64*E.
// The caller synchronises the workgroup before any wave reads another wave's transforms.
template <typename T, int E, bool INV>
__device__ __forceinline__ void lds_fft_wave(T* re, T* im, int B, const FFTPlan& plan,
                                             const C2<T>* __restrict__ tw, int off = 0)
{
    const int nw = (int)blockDim.x >> 6, w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int bpw = (B + nw - 1) / nw;
    const int b0 = w * bpw;
    const int nb = B - b0 < bpw ? B - b0 : bpw;
    if (nb <= 0) return;   // uniform over the wave
    const int N = plan.n;
    const int o = off + b0 * N;
    for (int p = 0; p < plan.nfac; p++) {
        const int R = plan.fac[p], Ns = plan.ns[p], ts = plan.toff[p];
        const unsigned a = plan.mnr[p], c = plan.mns[p];
        if (R == 16) fft_pass<T, 16, (E + 15) / 16, INV, true>(re, im, o, N, Ns, nb, tw, lane, 64, ts, a, c);
        else if (R == 8) fft_pass<T, 8, (E + 7) / 8, INV, true>(re, im, o, N, Ns, nb, tw, lane, 64, ts, a, c);
        else if (R == 4) fft_pass<T, 4, (E + 3) / 4, INV, true>(re, im, o, N, Ns, nb, tw, lane, 64, ts, a, c);
        else if (R == 2) fft_pass<T, 2, (E + 1) / 2, INV, true>(re, im, o, N, Ns, nb, tw, lane, 64, ts, a, c);
        else if (R == 3) fft_pass<T, 3, (E + 2) / 3, INV, true>(re, im, o, N, Ns, nb, tw, lane, 64, ts, a, c);
        else fft_pass<T, 5, (E + 4) / 5, INV, true>(re, im, o, N, Ns, nb, tw, lane, 64, ts, a, c);
    }
}

// Global -> LDS staging of a strided loop over [0, total): each thread issues the loads of
// R of its elements before it stores any of them, so it waits for memory once per R
// elements instead of once per element (a load followed by its LDS store in one loop
// iteration serialises on the load's latency).  ld(idx) loads element idx and must be
// unconditional (an index past the end is clamped to the last element and its value
// dropped); st(idx, v) applies any masking and stores.  tw_dst/tw_src/n_tw: a twiddle
// table copied alongside, its loads issued after the first chunk's.  No barrier.
template <int R, typename T, typename Ld, typename St>
__device__ __forceinline__ void staged_load(int total, Ld ld, St st, C2<T>* tw_dst = nullptr,
                                            const C2<T>* __restrict__ tw_src = nullptr, int n_tw = 0)
{
    const int nt = (int)blockDim.x, t0 = (int)threadIdx.x;
    bool tw_done = tw_dst == nullptr;
    for (int base = 0; base < total || !tw_done; base += R * nt) {
        decltype(ld(0)) v[R];
        if (base < total) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = base + r * nt + t0;
                v[r] = ld(idx < total ? idx : total - 1);
            }
        }
        if (!tw_done) {
            for (int i = t0; i < n_tw; i += nt) tw_dst[i] = tw_src[i];
            tw_done = true;
        }
        if (base < total) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = base + r * nt + t0;
                if (idx < total) st(idx, v[r]);
            }
        }
    }
}
