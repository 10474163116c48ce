// Device-side helpers shared by the gfx950 kernels of libertm_hip.so.
//
// Arithmetic contract: the library is compiled with -ffp-contract=off,
// correctly rounded float32 division and denormals preserved, so every
// float32 expression below rounds exactly where the reference's NumPy
// float32 array expression rounds (operation by operation).  That makes the
// GPU engine bit-identical to the reference on the same inputs, and makes
// every kernel deterministic (no atomics on floating-point data).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ertm_hip.h"

namespace ertm {

// Global-address-space views of device pointers.  Pointers read from the
// descriptor tables are generic (flat) to the compiler; flat memory
// operations also count against the LDS counter, so every LDS barrier would
// wait for all in-flight global loads.  Casting restores global_load/store.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float gcfloat;
__device__ __forceinline__ gfloat* gp(float* p) { return (gfloat*)p; }
__device__ __forceinline__ gcfloat* gp(const float* p) { return (gcfloat*)p; }

// Correctly rounded float32 quotient s/h without a division instruction.
// q0 = s * RN64(1/h) is within 2^-52 (relative) of s/h.
//   * normal results: s/h is never a float32 rounding midpoint (a midpoint's
//     significand is an odd 25-bit integer M, and M * odd(h) has more than 24
//     bits, so it is no float32 s), and a non-midpoint quotient stays at
//     least ~2^-50 (relative) away from every midpoint (s - m*h is a nonzero
//     multiple of 2^(exponent(s) - 48)); so rounding q0 to float32 already
//     gives RN32(s/h);
//   * subnormal results (|q0| < 2^-125 covers them): s/h can sit exactly on
//     a midpoint of the 2^-149 grid.  One float64 Newton step with the exact
//     FMA residual r = s - q0*h brings q1 within ~2^-104 of s/h; a midpoint
//     there is a float64 number, q1 equals it, and the final rounding applies
//     ties-to-even like the IEEE division.  The step runs only when some lane
//     of the wave needs it (the branch is on the wave's ballot).  A zero
//     quotient (s = 0: q0 = s * rh cannot underflow in float64) is exact
//     and skips it; zeros fill every region the wave has not reached, and
//     the empty asm keeps the step a real branch instead of code that every
//     wave executes.
// Non-finite inputs give the plain product's rounding (inf, NaN).
__device__ __forceinline__ float div_h(float s, float h, double rh) {
    const double q0 = (double)s * rh;
    float q = (float)q0;
    const bool tiny = q0 != 0.0 && __builtin_fabs(q0) < 0x1p-125;  // false for NaN
    if (__builtin_amdgcn_ballot_w64(tiny)) {
        asm volatile("");
        const double r = fma(-q0, (double)h, (double)s);
        const double q1 = fma(r, rh, q0);
        if (tiny) q = (float)q1;
    }
    return q;
}

// Half-stencil K = order/2.  Staggered first derivative, ertm/numerics.py:68-117:
//   plus : sum_k c_k (f[i+k]   - f[i-k+1]) / h   (value at i+1/2)
//   minus: sum_k c_k (f[i+k-1] - f[i-k])   / h   (value at i-1/2)
// accumulated in k order, then divided; zero outside [K, n-K).  The
// reference starts the sum from 0; 0 + t == t (up to the sign of a zero
// result), so the sum starts from the first term.
template <int K, typename F>
__device__ __forceinline__ float dplus(const F& f, int64_t idx, int64_t s, const float* c, float h,
                                       double rh) {
    float acc = c[0] * (f[idx + s] - f[idx]);
#pragma unroll
    for (int k = 2; k <= K; ++k) {
        float d = f[idx + k * s] - f[idx - (k - 1) * s];
        float t = c[k - 1] * d;
        acc = acc + t;
    }
    return div_h(acc, h, rh);
}

template <int K, typename F>
__device__ __forceinline__ float dminus(const F& f, int64_t idx, int64_t s, const float* c,
                                        float h, double rh) {
    float acc = c[0] * (f[idx] - f[idx - s]);
#pragma unroll
    for (int k = 2; k <= K; ++k) {
        float d = f[idx + (k - 1) * s] - f[idx - k * s];
        float t = c[k - 1] * d;
        acc = acc + t;
    }
    return div_h(acc, h, rh);
}

// CPML memory update and corrected derivative, ertm/numerics.py:221-227:
//   psi = psi*b ; psi = psi + a*D ; return D + psi
__device__ __forceinline__ float cpml(float d, float psi_in, float a, float b, float* psi_out) {
    float m = psi_in * b;
    float ad = a * d;
    m = m + ad;
    *psi_out = m;
    return d + m;
}

// Same recursion, returning the new memory instead of storing it (used where
// halo cells are recomputed and only the owner stores).
__device__ __forceinline__ float cpml_r(float d, float psi_in, float a, float b, float& psi) {
    float m = psi_in * b;
    float ad = a * d;
    psi = m + ad;
    return d + psi;
}

// Offsets of the compact CPML memories (see ErtmProp): x-memory of column j
// in row i, z-memory of row i in column j.  Return -1 outside the strips.
__device__ __forceinline__ int64_t psix_index(const ErtmProp& P, int i, int j) {
    if (P.wx <= 0) return -1;
    int jj;
    if (j < P.wx) jj = j;
    else if (j >= P.nxp - P.wx) jj = j - P.nxp + 2 * P.wx;
    else return -1;
    return (int64_t)i * (2 * P.wx) + jj;
}

__device__ __forceinline__ int64_t psiz_index(const ErtmProp& P, int i, int j) {
    int ii;
    if (i < P.wzt) ii = i;
    else if (P.wzb > 0 && i >= P.nzp - P.wzb) ii = i - P.nzp + P.wzt + P.wzb;
    else return -1;
    return (int64_t)ii * P.ld + j;
}

// Face averages, ertm/propagation.py:169-187 (0.5*(a+b) in float32; the last
// column / row is copied).
__device__ __forceinline__ float avg_x(const float* __restrict__ a, const ErtmProp& P, int i,
                                       int j) {
    int64_t o = (int64_t)i * P.ld + j;
    if (j < P.nxp - 1) {
        float s = a[o] + a[o + 1];
        return 0.5f * s;
    }
    return a[o];
}

__device__ __forceinline__ float avg_z(const float* __restrict__ a, const ErtmProp& P, int i,
                                       int j) {
    int64_t o = (int64_t)i * P.ld + j;
    if (i < P.nzp - 1) {
        float s = a[o] + a[o + P.ld];
        return 0.5f * s;
    }
    return a[o];
}

// mu on cell corners: mean_z(mean_x(mu)), ertm/propagation.py:185-187.
__device__ __forceinline__ float avg_xz(const float* __restrict__ a, const ErtmProp& P, int i,
                                        int j) {
    float m0 = avg_x(a, P, i, j);
    if (i < P.nzp - 1) {
        float m1 = avg_x(a, P, i + 1, j);
        float s = m0 + m1;
        return 0.5f * s;
    }
    return m0;
}

}  // namespace ertm
