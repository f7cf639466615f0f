// device_common.h - device helpers shared by the kernel translation units: vector
// products, the Philox4x32-10 counter-based generator and the XCD-aware block remap.
#pragma once
#include <hip/hip_runtime.h>

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


// XCD-aware remap: hardware places block b on XCD b % 8; give the blocks of one XCD a
// contiguous range of logical ids so neighbouring columns share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb8)
{
    const int per = nb8 >> 3;
    return (b & 7) * per + (b >> 3);
}
