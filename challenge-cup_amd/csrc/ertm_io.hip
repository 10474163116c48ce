// Receiver sampling/injection, image-field extraction and imaging-condition
// accumulation, reductions, the standalone derivative and the pieces of the
// viscoacoustic stabiliser.  All deterministic: injection is an ordered
// per-cell gather (never float atomics), reductions use order-independent max.
#include "ertm_kernels.h"

namespace ertm {

// ---------------------------------------------------------------------------
// sampling: ertm/numerics.py:478-511 via ertm/propagation.py:371-391
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sample_value(const ErtmProp& P, const float* S, int field,
                                              int64_t off) {
    if (field == -2) {  // elastic pressure 0.5*(sxx+szz), ertm/propagation.py:945-947
        float s = S[(int64_t)ERTM_EL_SXX * P.fstride + off] + S[(int64_t)ERTM_EL_SZZ * P.fstride + off];
        return 0.5f * s;
    }
    return S[(int64_t)field * P.fstride + off];
}

__global__ void k_sample(const ErtmProp* __restrict__ props, const ErtmRecv* __restrict__ recv,
                         LaunchArgs la) {
    const int slot = la_prop(la, blockIdx.y);
    const ErtmRecv& R = recv[slot];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= R.nsamp) return;
    const ErtmProp& P = props[slot];
    const float* S = P.state[la_cur(la, blockIdx.y)];
    const int t = la_time(la, blockIdx.y);
    const int field = R.s_field[e];
    double v = 0.0;
    if (field != -1 && R.s_valid[e]) {
        const int64_t o = (int64_t)R.s_i0[e] * P.ld + R.s_j0[e];
        const double tz = R.s_tz[e], tx = R.s_tx[e];
        const double wz0 = 1.0 - tz, wx0 = 1.0 - tx;
        const double v00 = sample_value(P, S, field, o);
        const double v01 = sample_value(P, S, field, o + 1);
        const double v10 = sample_value(P, S, field, o + P.ld);
        const double v11 = sample_value(P, S, field, o + P.ld + 1);
        double a = v00 * wz0;
        a = a * wx0;
        double b = v01 * wz0;
        b = b * tx;
        double c = v10 * tz;
        c = c * wx0;
        double d = v11 * tz;
        d = d * tx;
        v = a + b;
        v = v + c;
        v = v + d;
    }
    R.records[R.s_rec[e]][(int64_t)t * R.nrec + R.s_col[e]] = (float)v;
}

void launch_sample(const ErtmProp* dprops, const ErtmRecv* drecv, const LaunchArgs& la,
                   int max_samp, hipStream_t s) {
    if (max_samp <= 0) return;
    hipLaunchKernelGGL(k_sample, dim3((max_samp + 255) / 256, la.n), dim3(256), 0, s, dprops,
                       drecv, la);
}

// ---------------------------------------------------------------------------
// injection: ordered np.add.at semantics, ertm/propagation.py:331-369, :472-494
// ---------------------------------------------------------------------------
__global__ void k_inject(const ErtmProp* __restrict__ props, const ErtmRecv* __restrict__ recv,
                         LaunchArgs la) {
    const int slot = la_prop(la, blockIdx.y);
    const ErtmRecv& R = recv[slot];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= R.ncells) return;
    const ErtmProp& P = props[slot];
    float* S = P.state[la_cur(la, blockIdx.y)];
    const int t = la_time(la, blockIdx.y);
    float* f = S + (int64_t)R.cell_field[u] * P.fstride + R.cell_off[u];
    float x = *f;
    const int e1 = R.cell_ptr[u + 1];
    for (int e = R.cell_ptr[u]; e < e1; ++e) {
        const double v = (double)R.records[R.e_comp[e]][(int64_t)t * R.nrec + R.e_rec[e]];
        double c = v * R.e_wa[e];
        if (R.inject_mode == 0) c = c * R.e_wb[e];
        x = (float)((double)x + c);
    }
    *f = x;
}

void launch_inject(const ErtmProp* dprops, const ErtmRecv* drecv, const LaunchArgs& la,
                   int max_cells, hipStream_t s) {
    if (max_cells <= 0) return;
    hipLaunchKernelGGL(k_inject, dim3((max_cells + 255) / 256, la.n), dim3(256), 0, s, dprops,
                       drecv, la);
}

// ---------------------------------------------------------------------------
// image fields (ertm/propagation.py:496-497, :945-977) and imaging conditions
// (ertm/imaging.py:279-324, :450-474)
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ float wave_value(const ErtmProp& P, const float* S, int sep, int wave,
                                            int64_t o) {
    if (P.engine != ERTM_ELASTIC) return S[(int64_t)ERTM_AC_P * P.fstride + o];
    const float* vx = S + (int64_t)ERTM_EL_VX * P.fstride;
    const float* vz = S + (int64_t)ERTM_EL_VZ * P.fstride;
    if (sep == ERTM_SEP_STRESS) {
        if (wave == ERTM_WAVE_S) return S[(int64_t)ERTM_EL_SXZ * P.fstride + o];
        float s = S[(int64_t)ERTM_EL_SXX * P.fstride + o] + S[(int64_t)ERTM_EL_SZZ * P.fstride + o];
        return 0.5f * s;
    }
    // divcurl: plain derivatives, no CPML memory (ertm/propagation.py:949-963)
    if (wave == ERTM_WAVE_P) {
        float a = dminus<K>(vx, o, 1, P.coef, P.dx, P.rdx);
        float b = dminus<K>(vz, o, P.ld, P.coef, P.dz, P.rdz);
        return a + b;
    }
    float a = dminus<K>(vx, o, P.ld, P.coef, P.dz, P.rdz);
    float b = dminus<K>(vz, o, 1, P.coef, P.dx, P.rdx);
    return a - b;
}

template <int K>
__global__ void k_image(const ErtmProp* __restrict__ props, const ErtmImage* __restrict__ imgs,
                        LaunchArgs la) {
    const int slot = la_prop(la, blockIdx.y);
    const ErtmProp& P = props[slot];
    const ErtmImage& I = imgs[slot];
    const int64_t ncell = (int64_t)P.nz * P.nx;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const int i = (int)(q / P.nx), j = (int)(q - (int64_t)i * P.nx);
    const int64_t o = (int64_t)(i + P.pad) * P.ld + (j + P.pad);
    const float* S = P.state[la_cur(la, blockIdx.y)];
    const int t = la_time(la, blockIdx.y);
    const int t0 = la_t0(la, blockIdx.y, I);
    switch (la.op) {
        case ERTM_IMG_EXTRACT: {
            I.rbuf[(int64_t)(t - t0) * ncell + q] = wave_value<K>(P, S, I.separation, I.swave, o);
            break;
        }
        case ERTM_IMG_CORRELATE: {
            const float sv = I.rbuf[(int64_t)(t - t0) * ncell + q];
            const bool cross = t >= I.start_step;
            if (I.illum & ERTM_ILLUM_SRC) {
                float ss = sv * sv;
                I.ill[q] = I.ill[q] + (double)ss;
            }
            for (int w = 0; w < I.nrw; ++w) {
                const float rv = wave_value<K>(P, S, I.separation, I.rwave[w], o);
                if (I.illum & ERTM_ILLUM_REC) {
                    float rr = rv * rv;
                    I.rill[w][q] = I.rill[w][q] + (double)rr;
                }
                if (cross) {
                    float sr = sv * rv;
                    I.num[w][q] = I.num[w][q] + (double)sr;
                }
            }
            break;
        }
        case ERTM_IMG_EXC_PEAK: {
            // excitation maps use the P field, ertm/imaging.py:465-473
            const float a = fabsf(wave_value<K>(P, S, I.separation, ERTM_WAVE_P, o));
            if ((double)a > I.running[q]) {
                I.running[q] = (double)a;
                I.peak_amp[q] = (double)a;
                I.peak_step[q] = t;
            }
            break;
        }
        case ERTM_IMG_EXC_APPLY: {
            if (t >= I.start_step && I.peak_step[q] == t) {
                const float rv = wave_value<K>(P, S, I.separation, I.rwave[0], o);
                I.num[0][q] = I.num[0][q] + I.peak_amp[q] * (double)rv;
                if (I.illum & ERTM_ILLUM_REC) {
                    float rr = rv * rv;
                    I.rill[0][q] = I.rill[0][q] + (double)rr;
                }
            }
            break;
        }
        default:
            break;
    }
}

// Fold `count` backward steps t = t_hi, t_hi-1, ... (la.t = t_hi per
// propagator, la.op = count) from the receiver-field ring into the
// accumulators: per cell the additions happen in the same order, with the
// same float32 products, as `count` ERTM_IMG_CORRELATE passes.  With rscale
// (viscoacoustic stabiliser) ring entry t of table entry p is multiplied by
// rscale[(t mod R) * rs_n + p] first, in float32 as the stabiliser scales the
// state, which makes it the stabilised field the correlation reads.
__global__ void k_image_flush(const ErtmProp* __restrict__ props,
                              const ErtmImage* __restrict__ imgs, LaunchArgs la,
                              const float* __restrict__ rscale, int rs_n) {
    const int slot = la_prop(la, blockIdx.y);
    const ErtmProp& P = props[slot];
    const ErtmImage& I = imgs[slot];
    const int64_t ncell = (int64_t)P.nz * P.nx;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const int t_hi = la_time(la, blockIdx.y), count = la.op;
    const int t0 = la_t0(la, blockIdx.y, I);
    const bool src = (I.illum & ERTM_ILLUM_SRC) != 0, rec = (I.illum & ERTM_ILLUM_REC) != 0;
    const int nrw = I.nrw;
    double ill = src ? I.ill[q] : 0.0;
    double n0 = I.num[0][q], n1 = nrw > 1 ? I.num[1][q] : 0.0;
    double r0 = rec ? I.rill[0][q] : 0.0, r1 = (rec && nrw > 1) ? I.rill[1][q] : 0.0;
    // ring rows are on the state's pitch: cell (i, j) at row i, column j + pad
    const int64_t nring = (int64_t)P.nz * P.ld;
    const int64_t qi = q / P.nx;
    const int64_t qr = qi * P.ld + (q - qi * P.nx) + P.pad;
    for (int j = 0; j < count; ++j) {
        const int t = t_hi - j;
        const float sv = I.rbuf[(int64_t)(t - t0) * ncell + q];
        const float* ring = I.rring + (int64_t)(t % I.ring_len) * nrw * nring;
        if (src) {
            float ss = sv * sv;
            ill = ill + (double)ss;
        }
        const bool cross = t >= I.start_step;
        const float rs = rscale ? rscale[(int64_t)(t % I.ring_len) * rs_n + slot] : 1.0f;
        float rv0 = ring[qr];
        if (rscale) rv0 = rv0 * rs;
        if (rec) {
            float rr = rv0 * rv0;
            r0 = r0 + (double)rr;
        }
        if (cross) {
            float sr = sv * rv0;
            n0 = n0 + (double)sr;
        }
        if (nrw > 1) {
            float rv1 = ring[nring + qr];
            if (rscale) rv1 = rv1 * rs;
            if (rec) {
                float rr = rv1 * rv1;
                r1 = r1 + (double)rr;
            }
            if (cross) {
                float sr = sv * rv1;
                n1 = n1 + (double)sr;
            }
        }
    }
    if (src) I.ill[q] = ill;
    I.num[0][q] = n0;
    if (nrw > 1) I.num[1][q] = n1;
    if (rec) {
        I.rill[0][q] = r0;
        if (nrw > 1) I.rill[1][q] = r1;
    }
}

void launch_image_flush(const ErtmProp* dprops, const ErtmImage* dimg, const LaunchArgs& la,
                        int max_cells, const float* rscale, int rs_n, hipStream_t s) {
    if (max_cells <= 0 || la.op <= 0) return;
    hipLaunchKernelGGL(k_image_flush, dim3((max_cells + 255) / 256, la.n), dim3(256), 0, s, dprops,
                       dimg, la, rscale, rs_n);
}

void launch_image(const ErtmProp* dprops, const ErtmImage* dimg, const LaunchArgs& la, int order,
                  int max_cells, hipStream_t s) {
    if (max_cells <= 0) return;
    dim3 g((max_cells + 255) / 256, la.n);
    switch (order) {
        case 2: hipLaunchKernelGGL(k_image<1>, g, dim3(256), 0, s, dprops, dimg, la); break;
        case 4: hipLaunchKernelGGL(k_image<2>, g, dim3(256), 0, s, dprops, dimg, la); break;
        case 6: hipLaunchKernelGGL(k_image<3>, g, dim3(256), 0, s, dprops, dimg, la); break;
        default: hipLaunchKernelGGL(k_image<4>, g, dim3(256), 0, s, dprops, dimg, la); break;
    }
}

// ---------------------------------------------------------------------------
// max |x| (order-independent, hence deterministic)
// ---------------------------------------------------------------------------
__global__ void k_zero_u32(unsigned int* p) { *p = 0u; }

__global__ void k_absmax(const float* __restrict__ x, int64_t n, unsigned int* out) {
    float m = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        float v = fabsf(x[k]);
        m = (v > m || v != v) ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_down(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    __shared__ float red[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = (red[w] > r || red[w] != red[w]) ? red[w] : r;
        atomicMax(out, __float_as_uint(r));
    }
}

void launch_absmax(const float* x, int64_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, (unsigned int*)out);
    int64_t blocks = (n + 1023) / 1024;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)blocks), dim3(1024), 0, s, x, n,
                       (unsigned int*)out);
}

// max |field| of every member of a launch in one launch: member k's padded
// field slab (nzp rows of ld floats, as k_absmax reads it) reduces into
// out[k] (float bits, which order as unsigned integers for |x| and put NaN
// above infinity); out[0 .. la.n) is zeroed by k_zero_n first.
__global__ void k_zero_n(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_absmax_batch(const ErtmProp* __restrict__ props, LaunchArgs la, int field,
                               unsigned int* out) {
    const int k = blockIdx.y;
    const ErtmProp& P = props[la_prop(la, k)];
    const float* x = P.state[la_cur(la, k)] + (int64_t)field * P.fstride;
    const int64_t n = (int64_t)P.nzp * P.ld;
    float m = 0.0f;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        float v = fabsf(x[q]);
        m = (v > m || v != v) ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_down(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    __shared__ float red[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = (red[w] > r || red[w] != red[w]) ? red[w] : r;
        atomicMax(out + k, __float_as_uint(r));
    }
}

void launch_absmax_batch(const ErtmProp* dprops, const LaunchArgs& la, int field, int64_t max_n,
                         float* out, hipStream_t s) {
    if (la.n <= 0) return;
    hipLaunchKernelGGL(k_zero_n, dim3((la.n + 255) / 256), dim3(256), 0, s, (unsigned int*)out,
                       la.n);
    // about 2048 blocks over the launch, at least 8 per member
    int64_t per = (max_n + 1023) / 1024;
    per = std::min<int64_t>(per, std::max<int64_t>(8, 2048 / la.n));
    if (per < 1) per = 1;
    hipLaunchKernelGGL(k_absmax_batch, dim3((unsigned)per, la.n), dim3(1024), 0, s, dprops, la,
                       field, (unsigned int*)out);
}

// ---------------------------------------------------------------------------
// standalone staggered derivative of a dense array (ertm/numerics.py:68-137)
// ---------------------------------------------------------------------------
template <int K, typename T>
__global__ void k_derivative(const T* __restrict__ f, T* __restrict__ out, int n0, int n1,
                             int axis, int plus, DerivCoef c, double hd) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)n0 * n1) return;
    const int i0 = (int)(q / n1), i1 = (int)(q - (int64_t)i0 * n1);
    const int idx = axis == 0 ? i0 : i1;
    const int n = axis == 0 ? n0 : n1;
    const int64_t s = axis == 0 ? n1 : 1;
    const T h = (T)hd;
    T acc = (T)0;
    if (idx >= K && idx < n - K) {
#pragma unroll
        for (int k = 1; k <= K; ++k) {
            T d = plus ? (f[q + k * s] - f[q - (k - 1) * s]) : (f[q + (k - 1) * s] - f[q - k * s]);
            T t = (T)c.c64[k - 1] * d;
            if (sizeof(T) == 4) t = (T)c.c[k - 1] * d;
            acc = acc + t;
        }
        acc = acc / h;
    }
    out[q] = acc;
}

template <int K>
static void launch_derivative_k(const void* f, void* out, int n0, int n1, int axis, int plus,
                                const DerivCoef& c, double h, int is_f64, hipStream_t s) {
    const int64_t n = (int64_t)n0 * n1;
    dim3 g((unsigned)((n + 255) / 256));
    if (is_f64)
        hipLaunchKernelGGL((k_derivative<K, double>), g, dim3(256), 0, s, (const double*)f,
                           (double*)out, n0, n1, axis, plus, c, h);
    else
        hipLaunchKernelGGL((k_derivative<K, float>), g, dim3(256), 0, s, (const float*)f,
                           (float*)out, n0, n1, axis, plus, c, h);
}

void launch_derivative(const void* f, void* out, int n0, int n1, int axis, int plus, int order,
                       const DerivCoef& c, double h, int is_f64, hipStream_t s) {
    switch (order) {
        case 2: launch_derivative_k<1>(f, out, n0, n1, axis, plus, c, h, is_f64, s); break;
        case 4: launch_derivative_k<2>(f, out, n0, n1, axis, plus, c, h, is_f64, s); break;
        case 6: launch_derivative_k<3>(f, out, n0, n1, axis, plus, c, h, is_f64, s); break;
        default: launch_derivative_k<4>(f, out, n0, n1, axis, plus, c, h, is_f64, s); break;
    }
}

// ---------------------------------------------------------------------------
// viscoacoustic stabiliser, ertm/propagation.py:726-749
// ---------------------------------------------------------------------------
// rscale (may be null): the factor the receiver-field ring entry of this step
// needs to equal the stabilised pressure, for members 0..n-1: the clip scale,
// or 1 at low-pass steps (the low-pass rewrites the entry after filtering).
__global__ void k_stab_control(const float* __restrict__ peaks, double* __restrict__ ctl,
                               int n, int step_index, double cap, int lowpass_now,
                               float* __restrict__ scales, float* __restrict__ rscale) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double peak = (double)peaks[k];
    double* c = ctl + 4 * k;
    float sc = 1.0f;
    if (step_index < 30) {
        c[0] = c[0] > peak ? c[0] : peak;
    } else {
        const double ref = c[0] > 1e-30 ? c[0] : 1e-30;
        const double limit = cap * ref;
        if (peak > limit) {
            sc = (float)(limit / peak);
            c[1] += 1.0;
        }
        if (lowpass_now) c[2] += 1.0;
    }
    scales[k] = sc;
    if (rscale) rscale[k] = (step_index >= 30 && lowpass_now) ? 1.0f : sc;
}

void launch_stab_control(const ErtmProp*, const LaunchArgs& la, const float* peaks, double* ctl,
                         int step_index, double cap, int lowpass_now, float* scales,
                         float* rscale, hipStream_t s) {
    hipLaunchKernelGGL(k_stab_control, dim3(1), dim3(64), 0, s, peaks, ctl, la.n, step_index, cap,
                       lowpass_now, scales, rscale);
}

// Clip scaling of the viscoacoustic state (ertm/propagation.py:740-747):
// every float of p, vx, vz, eps and r_0..r_{L-1} times the shot's factor.
// Four cells per thread as float4 (rows and field slabs are multiples of 64
// floats), and all fields' loads issued before the stores: the slabs do not
// overlap, so this is the same arithmetic as field after field, with up to
// 4 + L loads in flight per thread instead of one.
__global__ void k_stab_scale(const ErtmProp* __restrict__ props, LaunchArgs la,
                             const float* __restrict__ scales) {
    const int slot = la_prop(la, blockIdx.y);
    const float sc = scales[blockIdx.y];
    if (sc == 1.0f) return;
    const ErtmProp& P = props[slot];
    float* S1 = P.state[la_cur(la, blockIdx.y)];
    const int nf = 4 + P.nmech;  // p, vx, vz, eps, r_0..r_{L-1}
    if (((uintptr_t)S1 & 15) != 0 || (P.ld & 3) != 0 || (P.fstride & 3) != 0) {
        // a caller's state that is not float4-aligned: one float at a time
        const int64_t cells = (int64_t)P.nzp * P.ld;
        for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < cells;
             q += (int64_t)gridDim.x * blockDim.x)
            for (int f = 0; f < nf; ++f) {
                float* x = S1 + (int64_t)f * P.fstride + q;
                *x = *x * sc;
            }
        return;
    }
    float4* S = (float4*)S1;
    const int64_t cells4 = (int64_t)P.nzp * P.ld / 4;
    const int64_t fs4 = P.fstride / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < cells4;
         q += (int64_t)gridDim.x * blockDim.x) {
        float4 v[4 + ERTM_MAX_MECH];
#pragma unroll
        for (int f = 0; f < 4 + ERTM_MAX_MECH; ++f)
            if (f < nf) v[f] = S[f * fs4 + q];
#pragma unroll
        for (int f = 0; f < 4 + ERTM_MAX_MECH; ++f) {
            if (f < nf) {
                float4 w = v[f];
                w.x = w.x * sc;
                w.y = w.y * sc;
                w.z = w.z * sc;
                w.w = w.w * sc;
                S[f * fs4 + q] = w;
            }
        }
    }
}

void launch_stab_scale(const ErtmProp* dprops, const LaunchArgs& la, const float* scales,
                       int max_tiles, hipStream_t s) {
    (void)max_tiles;
    hipLaunchKernelGGL(k_stab_scale, dim3(512, la.n), dim3(256), 0, s, dprops, la, scales);
}

// k-space low-pass of padded fields (ertm/numerics.py:424-442), batched over
// the p and eps fields of every member of an LpArgs group (transform y =
// 2 * member + field, fields of one shape).  The input is real and the radial
// ramp is even in kz and kx, so the filtered spectrum is Hermitian and the
// real part of the full inverse transform equals the real-to-complex round
// trip: each field is widened to float64, transformed by the caller with a
// D2Z plan, weighted on the half spectrum (nxp / 2 + 1 columns), transformed
// back with a Z2D plan and narrowed with the 1/N factor.
__device__ __forceinline__ const ErtmProp& lp_prop(const ErtmProp* props, const LpArgs& a, int y) {
    return props[a.prop[y >> 1]];
}

__device__ __forceinline__ float* lp_field(const ErtmProp& P, const LpArgs& a, int y) {
    return P.state[a.cur] + (int64_t)((y & 1) ? ERTM_VI_EPS : ERTM_VI_P) * P.fstride;
}

__global__ void k_lp_prep(const ErtmProp* __restrict__ props, LpArgs a, double* __restrict__ r) {
    const int y = blockIdx.y;
    const int64_t cells = (int64_t)a.nzp * a.nxp;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cells) return;
    const ErtmProp& P = lp_prop(props, a, y);
    const int i = (int)(q / a.nxp), j = (int)(q - (int64_t)i * a.nxp);
    r[(int64_t)y * cells + q] = (double)lp_field(P, a, y)[(int64_t)i * P.ld + j];
}

__device__ __forceinline__ double fftfreq(int k, int n, double d) {
    const int half = (n - 1) / 2 + 1;
    const double v = 1.0 / (n * d);
    return (k < half ? (double)k : (double)(k - n)) * v;
}

__global__ void k_lp_response(const ErtmProp* __restrict__ props, LpArgs a,
                              double2* __restrict__ c) {
    const int y = blockIdx.y;
    const int nh = a.nxp / 2 + 1;
    const int64_t cells = (int64_t)a.nzp * nh;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cells) return;
    const ErtmProp& P = lp_prop(props, a, y);
    const int i = (int)(q / nh), j = (int)(q - (int64_t)i * nh);
    const double dx = (double)P.dx, dz = (double)P.dz;
    double kzm = 0.5 / dz, kxm = 0.5 / dx;
    kzm = kzm > 1e-12 ? kzm : 1e-12;
    kxm = kxm > 1e-12 ? kxm : 1e-12;
    // column j of the half spectrum is frequency index j of the full one
    const double kz = fftfreq(i, a.nzp, dz) / kzm;
    const double kx = fftfreq(j, a.nxp, dx) / kxm;
    const double rr = sqrt(kz * kz + kx * kx);
    double w = 1.0 - (rr - a.ratio[y >> 1]) / 0.15;
    w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    double2* v = c + (int64_t)y * cells + q;
    v->x *= w;
    v->y *= w;
}

// With a.ring_slot >= 0 the filtered pressure of each physical cell is also
// written to that receiver-field ring slot (rows on the state's pitch: cell
// (i, j) of the physical grid at row i, column j + pad).
__global__ void k_lp_finish(const ErtmProp* __restrict__ props,
                            const ErtmImage* __restrict__ imgs, LpArgs a,
                            const double* __restrict__ r) {
    const int y = blockIdx.y;
    const int64_t cells = (int64_t)a.nzp * a.nxp;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cells) return;
    const ErtmProp& P = lp_prop(props, a, y);
    const int i = (int)(q / a.nxp), j = (int)(q - (int64_t)i * a.nxp);
    const float v = (float)(r[(int64_t)y * cells + q] / ((double)a.nzp * (double)a.nxp));
    lp_field(P, a, y)[(int64_t)i * P.ld + j] = v;
    if (a.ring_slot >= 0 && (y & 1) == 0) {
        const ErtmImage& I = imgs[a.prop[y >> 1]];
        const int ii = i - P.pad, jj = j - P.pad;
        if (ii >= 0 && ii < P.nz && jj >= 0 && jj < P.nx) {
            const int64_t nring = (int64_t)P.nz * P.ld;
            I.rring[(int64_t)a.ring_slot * I.nrw * nring + (int64_t)ii * P.ld + j] = v;
        }
    }
}

void launch_lowpass_prep(const ErtmProp* dprops, const LpArgs& a, double* real, hipStream_t s) {
    const int64_t n = (int64_t)a.nzp * a.nxp;
    hipLaunchKernelGGL(k_lp_prep, dim3((unsigned)((n + 255) / 256), 2 * a.n), dim3(256), 0, s,
                       dprops, a, real);
}

void launch_lowpass_response(const ErtmProp* dprops, const LpArgs& a, void* cplx, hipStream_t s) {
    const int64_t n = (int64_t)a.nzp * (a.nxp / 2 + 1);
    hipLaunchKernelGGL(k_lp_response, dim3((unsigned)((n + 255) / 256), 2 * a.n), dim3(256), 0, s,
                       dprops, a, (double2*)cplx);
}

void launch_lowpass_finish(const ErtmProp* dprops, const ErtmImage* dimg, const LpArgs& a,
                           const double* real, hipStream_t s) {
    const int64_t n = (int64_t)a.nzp * a.nxp;
    hipLaunchKernelGGL(k_lp_finish, dim3((unsigned)((n + 255) / 256), 2 * a.n), dim3(256), 0, s,
                       dprops, dimg, a, real);
}

// ---------------------------------------------------------------------------
// image post-processing on the device (ertm/imaging.py:326-331, :476-502):
// normalisation, 5-point Laplacian with edge padding (ertm/numerics.py:398-404),
// radial wavenumber low-pass (ertm/numerics.py:424-442), PML mask, float32.
// All float64; normalisation, Laplacian and mask round exactly like NumPy,
// the FFT differs from pocketfft in rounding only.
// ---------------------------------------------------------------------------
__global__ void k_post_norm(const double* __restrict__ num, const double* __restrict__ ill,
                            const double* __restrict__ rill, int64_t n, double eps, int div_ill,
                            int div_rill, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double v = num[q];
    if (div_ill) v = v / (ill[q] + eps);
    if (div_rill) v = v / (rill[q] + eps);
    out[q] = v;
}

// Laplacian of a (nz, nx) float64 image; writes float64 (to_complex = 0) or
// the real parts of a complex buffer (to_complex = 1)
__global__ void k_post_lap(const double* __restrict__ a, int nz, int nx, double dxx, double dzz,
                           int to_complex, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nz * nx) return;
    const int i = (int)(q / nx), j = (int)(q - (int64_t)i * nx);
    const double c = a[q];
    const double xl = a[(int64_t)i * nx + (j > 0 ? j - 1 : 0)];
    const double xr = a[(int64_t)i * nx + (j < nx - 1 ? j + 1 : nx - 1)];
    const double zu = a[(int64_t)(i > 0 ? i - 1 : 0) * nx + j];
    const double zd = a[(int64_t)(i < nz - 1 ? i + 1 : nz - 1) * nx + j];
    const double c2 = 2.0 * c;
    double hx = xr - c2;
    hx = hx + xl;
    double hz = zd - c2;
    hz = hz + zu;
    const double v = hx / dxx + hz / dzz;
    if (to_complex) {
        out[2 * q] = v;
        out[2 * q + 1] = 0.0;
    } else {
        out[q] = v;
    }
}

__global__ void k_post_to_complex(const double* __restrict__ a, int64_t n, double* __restrict__ c) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    c[2 * q] = a[q];
    c[2 * q + 1] = 0.0;
}

__global__ void k_post_response(double2* __restrict__ c, int nz, int nx, double dx, double dz,
                                double frac) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nz * nx) return;
    const int i = (int)(q / nx), j = (int)(q - (int64_t)i * nx);
    const double kz = fftfreq(i, nz, dz) / fmax(0.5 / dz, 1e-12);
    const double kx = fftfreq(j, nx, dx) / fmax(0.5 / dx, 1e-12);
    const double r = sqrt(kz * kz + kx * kx);
    double resp = 1.0 - (r - frac) / 0.15;
    resp = resp < 0.0 ? 0.0 : (resp > 1.0 ? 1.0 : resp);
    c[q].x = c[q].x * resp;
    c[q].y = c[q].y * resp;
}

// real part (complex input, divided by nz*nx) or the float64 image itself,
// times the PML mask, to float32
__global__ void k_post_finish(const double* __restrict__ src, int from_complex, int nz, int nx,
                              int npml, int mask_top, float* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)nz * nx;
    if (q >= n) return;
    const int i = (int)(q / nx), j = (int)(q - (int64_t)i * nx);
    const double v = from_complex ? src[2 * q] / (double)n : src[q];
    const bool zero = j < npml || j >= nx - npml || i >= nz - npml || (mask_top && i < npml);
    const double m = zero ? 0.0 : 1.0;
    out[q] = (float)(v * m);
}

static dim3 grid_of(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_post_norm(const double* num, const double* ill, const double* rill, int64_t n,
                      double eps, int div_ill, int div_rill, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_post_norm, grid_of(n), dim3(256), 0, s, num, ill, rill, n, eps, div_ill,
                       div_rill, out);
}

void launch_post_lap(const double* a, int nz, int nx, double dxx, double dzz, int to_complex,
                     double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_post_lap, grid_of((int64_t)nz * nx), dim3(256), 0, s, a, nz, nx, dxx, dzz,
                       to_complex, out);
}

void launch_post_to_complex(const double* a, int64_t n, double* c, hipStream_t s) {
    hipLaunchKernelGGL(k_post_to_complex, grid_of(n), dim3(256), 0, s, a, n, c);
}

void launch_post_response(double* c, int nz, int nx, double dx, double dz, double frac,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_post_response, grid_of((int64_t)nz * nx), dim3(256), 0, s, (double2*)c,
                       nz, nx, dx, dz, frac);
}

void launch_post_finish(const double* src, int from_complex, int nz, int nx, int npml,
                        int mask_top, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_post_finish, grid_of((int64_t)nz * nx), dim3(256), 0, s, src,
                       from_complex, nz, nx, npml, mask_top, out);
}

// ---------------------------------------------------------------------------
// record preparation on the device (ertm/preprocess.py:36-63, :121-189):
// float64 arithmetic in the reference's order, float32 after each step
// ---------------------------------------------------------------------------
__global__ void k_prep_mute_taper(float* __restrict__ rec, int nt, int nrec,
                                  const double* __restrict__ t, const double* __restrict__ t_mute,
                                  double taper, int w_end) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nt * nrec) return;
    const int it = (int)(q / nrec), ir = (int)(q - (int64_t)it * nrec);
    double x = (double)rec[q];
    if (t_mute) {
        const double start = t_mute[ir] - taper, end = t_mute[ir] + taper;
        double ramp = (t[it] - start) / (end - start);
        ramp = ramp < 0.0 ? 0.0 : (ramp > 1.0 ? 1.0 : ramp);
        const double w = 0.5 - 0.5 * cos(M_PI * ramp);
        x = (double)(float)(x * w);
    }
    if (w_end > 0) {
        // ramp[i] = 0.5 - 0.5 cos((pi * (i + 0.5)) / w), NumPy's evaluation order
        if (it < w_end) x = x * (0.5 - 0.5 * cos((M_PI * ((double)it + 0.5)) / w_end));
        if (it >= nt - w_end) {
            const int j = nt - 1 - it;
            x = x * (0.5 - 0.5 * cos((M_PI * ((double)j + 0.5)) / w_end));
        }
    }
    rec[q] = (float)x;
}

__global__ void k_prep_rotate(float* __restrict__ vx, float* __restrict__ vz, int64_t n, double c,
                              double s) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double a = (double)vx[q], b = (double)vz[q];
    vx[q] = (float)(c * a + s * b);
    vz[q] = (float)(-s * a + c * b);
}

void launch_prep_mute_taper(float* rec, int nt, int nrec, const double* t, const double* t_mute,
                            double taper, int w_end, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_mute_taper, grid_of((int64_t)nt * nrec), dim3(256), 0, s, rec, nt,
                       nrec, t, t_mute, taper, w_end);
}

// Sinc resampling of a record (ertm/numerics.py:337-392) followed by the
// length fit of ertm/imaging.py:416-423: out row k < n_out is
// sum_q w[k][q] * in[clip(off[k][q], 0, n_in - 1)] over the taps q in order,
// in float64 from zero with every product and sum rounded once (the order
// NumPy's einsum takes for records of two or more traces), then rounded to
// float32; rows n_out .. out_rows-1 are zero.  Thread = (row, trace), traces
// fastest: coalesced gathers of input rows.
__global__ void k_resample(const float* __restrict__ in, int n_in, int nrec,
                           const double* __restrict__ w, const int32_t* __restrict__ off,
                           int n_out, int taps, float* __restrict__ out, int out_rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)out_rows * nrec) return;
    const int k = (int)(i / nrec);
    const int r = (int)(i - (int64_t)k * nrec);
    float v = 0.0f;
    if (k < n_out) {
        double acc = 0.0;
        for (int q = 0; q < taps; ++q) {
            const int j = min(max(off[(int64_t)k * taps + q], 0), n_in - 1);
            const double p = w[(int64_t)k * taps + q] * (double)in[(int64_t)j * nrec + r];
            acc = acc + p;
        }
        v = (float)acc;
    }
    out[i] = v;
}

void launch_resample(const float* in, int n_in, int nrec, const double* w, const int32_t* off,
                     int n_out, int taps, float* out, int out_rows, hipStream_t s) {
    hipLaunchKernelGGL(k_resample, grid_of((int64_t)out_rows * nrec), dim3(256), 0, s, in, n_in,
                       nrec, w, off, n_out, taps, out, out_rows);
}

void launch_prep_rotate(float* vx, float* vz, int64_t n, double c, double s_, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_rotate, grid_of(n), dim3(256), 0, s, vx, vz, n, c, s_);
}

// ---------------------------------------------------------------------------
// device clock of a captured time loop
// ---------------------------------------------------------------------------
__global__ void k_tick(int32_t* clk, int steps) {
    if (blockIdx.x == 0 && threadIdx.x == 0) clk[0] = clk[0] + steps;
}

void launch_tick(int32_t* clk, int steps, hipStream_t s) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, clk, steps);
}

// Checkpoint save whose slot depends on the device clock; blocks of
// propagators whose time is no checkpoint time return at once.  States are
// whole multiples of 64 floats and 256-byte aligned: float4 copies.
__global__ void k_checkpoint(const ErtmProp* __restrict__ props, LaunchArgs la,
                             float* const* __restrict__ arenas, int64_t stride, int origin,
                             int interval) {
    const int k = blockIdx.y;
    const int t = la_time(la, k);
    if (t < origin || (t - origin) % interval != 0) return;
    const ErtmProp& P = props[la_prop(la, k)];
    const int64_t slot = (t - origin) / interval;
    const float4* src = (const float4*)P.state[la_cur(la, k)];
    float4* dst = (float4*)(arenas[la_prop(la, k)] + slot * stride);
    const int64_t n4 = P.state_floats / 4;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Eager checkpoint saves and restores of a whole batch in one launch
// (blockIdx.y = propagator): one hipMemcpyAsync per shot ran a 64-shot
// 600x1200 elastic save (2.3 GB) at a few hundred GB/s.
__global__ void k_copy_batch(CopyArgs a) {
    const int k = blockIdx.y;
    const float4* __restrict__ src = a.src[k];
    float4* __restrict__ dst = a.dst[k];
    const int64_t n4 = a.n4[k];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy_batch(const CopyArgs& a, hipStream_t s) {
    int64_t most = 1;
    for (int k = 0; k < a.n; ++k) most = std::max<int64_t>(most, a.n4[k]);
    // about 8192 blocks over the batch: 32 per CU, enough loads in flight to
    // stream HBM, and never more blocks than a propagator has float4 groups
    const int64_t per = std::max<int64_t>(1, 8192 / std::max(a.n, 1));
    const int64_t gx = std::min<int64_t>(per, (most + 255) / 256);
    hipLaunchKernelGGL(k_copy_batch, dim3((unsigned)gx, a.n), dim3(256), 0, s, a);
}

void launch_checkpoint(const ErtmProp* dprops, const LaunchArgs& la, float* const* darenas,
                       int64_t stride, int origin, int interval, int64_t max_floats,
                       hipStream_t s) {
    // 64 blocks per propagator: enough to stream a due checkpoint, few to
    // retire on the (far more common) steps where nothing is due
    const int64_t blocks = std::min<int64_t>(64, (max_floats / 4 + 255) / 256);
    hipLaunchKernelGGL(k_checkpoint, dim3((unsigned)std::max<int64_t>(blocks, 1), la.n), dim3(256),
                       0, s, dprops, la, darenas, stride, origin, interval);
}

}  // namespace ertm
