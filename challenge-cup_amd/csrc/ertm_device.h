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

// Half-stencil K = order/2.  Staggered first derivative, ertm/numerics.py:68-117:
//   plus : sum_k c_k (f[i+k]   - f[i-k+1]) / h   (value at i+1/2)
//   minus: sum_k c_k (f[i+k-1] - f[i-k])   / h   (value at i-1/2)
// accumulated from 0 in k order, then divided; zero outside [K, n-K).
template <int K>
__device__ __forceinline__ float dplus(const float* __restrict__ f, int64_t idx, int64_t s,
                                       const float* c, float h) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 1; k <= K; ++k) {
        float d = f[idx + k * s] - f[idx - (k - 1) * s];
        float t = c[k - 1] * d;
        acc = acc + t;
    }
    return acc / h;
}

template <int K>
__device__ __forceinline__ float dminus(const float* __restrict__ f, int64_t idx, int64_t s,
                                        const float* c, float h) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 1; k <= K; ++k) {
        float d = f[idx + (k - 1) * s] - f[idx - k * s];
        float t = c[k - 1] * d;
        acc = acc + t;
    }
    return acc / h;
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
