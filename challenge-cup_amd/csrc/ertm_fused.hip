// Fused single-launch time step for the elastic P-SV engine.
//
// One launch advances velocities AND stresses by one step (52 B of HBM
// traffic per cell: 5 state reads + 5 writes + lambda, mu, 1/rho), instead of
// the two launches of ertm_stencil.hip (72 B).
//
// Work decomposition: a block of NT threads owns a band of TX = NT - 4K
// columns and a chunk of ZC rows.  Thread t handles column x0 - 2K + t and
// walks the chunk row by row (z-streaming):
//   * old stresses are loaded one row ahead of use (row k+2K at iteration k)
//     into per-thread register queues that hold the 2K+1 rows a z-derivative
//     needs; new velocities go to register queues too;
//   * x-neighbours are exchanged through one LDS row per field;
//   * the velocity of row k+K is computed for columns [x0-K, x0+TX+K) (the
//     halo columns and the K rows above/below the chunk are recomputed, not
//     stored), then the stresses of row k for the owned columns.
// The per-cell arithmetic is the same inline code as the two-launch kernels
// (ertm_device.h), so results are bit-identical to them and to the
// reference.  Ping-pong state buffers make the recomputed halos race-free.
//
// Block order is XCD-aware: the hardware deals consecutive workgroups to the
// 8 XCDs round-robin; the remap gives each XCD a contiguous run of logical
// blocks, i.e. vertically adjacent chunks of one band, whose halo rows then
// hit in that XCD's L2.
#include "ertm_fused_common.h"

namespace ertm {

template <int K, int QXL>
struct ElState {
    float q_sxx[2 * K + 1], q_szz[2 * K + 1], q_sxz[2 * K + 1];  // rows k .. k+2K*D
    float q_vx[QXL];   // q_vx[i] holds row k + (i - (K - 1 + QXL - 2K)) * D
    float q_vz[2 * K + 1];                                      // rows k-K*D .. k+K*D
    float ir_c, mu_c;  // own column, carried from the previous row (see el_iter)
    float n_sxx, n_szz, n_sxz, n_vx, n_vz, n_ir, n_mu, n_lam;   // inputs of this row
    unsigned i_ir, i_mu, i_lam;  // with a medium palette: their indices, looked up in el_iter
};

// Load the inputs of iteration k: stress row k+2K*D, velocity row m = k+K*D,
// 1/rho row m+1 (down) / m (up), mu row k+1 (down) / k (up), lambda row k.
// Medium palette index of padded row offset `rowb` (bytes of a float row) at
// the lane's column: one byte per cell, rows on the float rows' pitch / 4.
// Rows outside the grid read 0, the zero medium.
template <class Blk>
__device__ __forceinline__ unsigned mindex(const Blk& B, int rowb) {
    return __builtin_amdgcn_raw_buffer_load_b8(B.rMI, B.c + (rowb >> 2), 0, 0);
}

template <int K, int QXL, int D, bool PAL>
__device__ __forceinline__ void el_fetch(const ElBlock& B, ElState<K, QXL>& X, int k) {
    const int ld4 = B.ld4;
    const int rs = (k + 2 * K * D) * ld4, rv = (k + K * D) * ld4;
    const int ri = (k + K * D + (D > 0 ? 1 : 0)) * ld4, rm = (k + (D > 0 ? 1 : 0)) * ld4;
    X.n_sxx = bload(B.rS, B.c4 + (rs + B.fXX));
    X.n_szz = bload(B.rS, B.c4 + (rs + B.fZZ));
    X.n_sxz = bload(B.rS, B.c4 + (rs + B.fXZ));
    X.n_vx = bload(B.rS, B.c4 + (rv + B.fVX));
    X.n_vz = bload(B.rS, B.c4 + (rv + B.fVZ));
    if constexpr (PAL) {
        X.i_ir = mindex(B, ri);
        X.i_mu = mindex(B, rm);
        X.i_lam = mindex(B, k * ld4);
    } else {
        X.n_ir = bload(B.rIR, B.c4 + ri);
        X.n_mu = bload(B.rMU, B.c4 + rm);
        X.n_lam = bload(B.rLAM, B.c4 + k * ld4);
    }
}

// vx queue length: Dz+vx at row k reads rows k-K+1 .. k+K, i.e. K rows ahead
// of a downward stream but K rows behind an upward one; streaming down the
// queue ends at the velocity front m = k+K (2K rows), streaming up it must
// reach back to row k+K (2K+1 rows).  S imaging (Dz-vx, rows k-K .. k+K-1)
// needs 2K+1 rows either way.
template <int K, int IMG, int D>
constexpr int qx_len() { return (!img_may_s<IMG>() && D > 0) ? 2 * K : 2 * K + 1; }

// One row iteration.  STRESS = false: velocity only (the 2K prologue rows
// of a chunk, whose stress rows lie outside the chunk).
//
// D = +1 streams the chunk top to bottom (rows increasing), D = -1 bottom to
// top.  Register queues are ordered along the stream: q[j] holds row
// (first row) + j*D, so a z-derivative at queue index i is dplus/dminus with
// stride D -- the same operands in the same order as in row space, hence the
// same rounding.  Velocity runs K rows ahead of stress along the stream:
// m = k + D*K.  1/rho is needed at rows m and m+1 and mu at rows k and k+1:
// the row reached first along the stream is carried, the other one loaded.
// L = this iteration's LDS buffer (7 rows of NT floats).
template <int K, int NT, bool STRIP, bool STRESS, int IMG, int D, bool PAL>
__device__ __forceinline__ void el_iter(const ElBlock& B, ElState<K, qx_len<K, IMG, D>()>& X,
                                        int k, float* L) {
    constexpr int QS = 2 * K + 1, QX = qx_len<K, IMG, D>(), QZ = 2 * K + 1;
    constexpr int XO = QX - 2 * K;  // q_vx[i] holds row k + (i - (K - 1 + XO)) * D
    const int tid = B.tid, ld4 = B.ld4, nzp = B.nzp;
    const int m = k + D * K;
    if constexpr (PAL) {  // this row's medium from the palette (the same float values)
        X.n_ir = B.ptab[X.i_ir];
        X.n_mu = B.ptab[256 + X.i_mu];
        X.n_lam = B.ptab[512 + X.i_lam];
    }
    const float vx_old = X.n_vx, vz_old = X.n_vz, lam_k = X.n_lam;
    const float ir_m = D > 0 ? X.ir_c : X.n_ir;   // 1/rho, row m
    const float ir_m1 = D > 0 ? X.n_ir : X.ir_c;  // row m+1
    const float mu_k = D > 0 ? X.mu_c : X.n_mu;   // mu, row k
    const float mu_k1 = D > 0 ? X.n_mu : X.mu_c;  // row k+1
#pragma unroll
    for (int q = 0; q < QS - 1; ++q) {
        X.q_sxx[q] = X.q_sxx[q + 1];
        X.q_szz[q] = X.q_szz[q + 1];
        X.q_sxz[q] = X.q_sxz[q + 1];
    }
    X.q_sxx[QS - 1] = X.n_sxx;
    X.q_szz[QS - 1] = X.n_szz;
    X.q_sxz[QS - 1] = X.n_sxz;
    // prefetch the next iteration's inputs; in flight during this row
    el_fetch<K, QX, D, PAL>(B, X, k + D);

    // ---- LDS rows of this iteration, one barrier: stress row m and 1/rho
    // row m (velocity x-derivatives); velocity row k, computed K iterations
    // ago, and mu rows k, k+1 (stress x-derivatives).  The buffers alternate
    // between iterations, so the next iteration's writes cannot overtake a
    // wave still reading these.
    float* l_sxx = L;
    float* l_sxz = L + NT;
    float* l_ir = L + 2 * NT;
    float* l_vx = L + 3 * NT;
    float* l_vz = L + 4 * NT;
    float* l_mu0 = L + 5 * NT;
    float* l_mu1 = L + 6 * NT;
    l_sxx[tid] = X.q_sxx[K];
    l_sxz[tid] = X.q_sxz[K];
    l_ir[tid] = ir_m;
    if (STRESS) {
        l_vx[tid] = X.q_vx[K + XO];  // row k (the queue is pushed below)
        l_vz[tid] = X.q_vz[K + 1];
        l_mu0[tid] = mu_k;
        l_mu1[tid] = mu_k1;
    }
    __syncthreads();

    // velocity of row m on every lane (lanes outside the band and rows
    // outside the grid compute values nobody uses)
    const bool inz_m = m >= K && m < nzp - K;
    const bool row_m = m >= B.z0 && m < B.z1;  // velocity row m is this block's
    const bool own_m = B.own && row_m;
    float a = dplus<K>(l_sxx, tid, 1, B.cf, B.hx, B.rhx);
    float b = dminus<K>(X.q_sxz, K, D, B.cf, B.hz, B.rhz);
    float cc = dminus<K>(l_sxz, tid, 1, B.cf, B.hx, B.rhx);
    float d = dplus<K>(X.q_szz, K, D, B.cf, B.hz, B.rhz);
    a = B.inx ? a : 0.0f;
    cc = B.inx ? cc : 0.0f;
    b = inz_m ? b : 0.0f;
    d = inz_m ? d : 0.0f;
    if (STRIP) {
        if (B.velc && m >= 0 && m < nzp) {
            int pzo = -1;  // byte offset of (m, c) inside a z-memory
            if (m < B.wzt) pzo = m * ld4 + B.c4;
            else if (B.wzb > 0 && m >= nzp - B.wzb) pzo = (m - nzp + B.wzt + B.wzb) * ld4 + B.c4;
            const int pxo = B.jx4 >= 0 ? m * B.xrow4 + B.jx4 : -1;
            float pn;
            if (pxo >= 0) {
                a = cpml_r(a, bload(B.rS, B.px0 + pxo), B.axc, B.bxc, pn);
                bstore_l2(B.rT, own_m ? B.px0 + pxo : ERTM_NOWRITE, pn);
                cc = cpml_r(cc, bload(B.rS, B.px1 + pxo), B.axc, B.bxc, pn);
                bstore_l2(B.rT, own_m ? B.px1 + pxo : ERTM_NOWRITE, pn);
            }
            if (pzo >= 0) {
                const float az = B.AZ[m], bz = B.BZ[m];
                b = cpml_r(b, bload(B.rS, B.pz0 + pzo), az, bz, pn);
                bstore_l2(B.rT, own_m ? B.pz0 + pzo : ERTM_NOWRITE, pn);
                d = cpml_r(d, bload(B.rS, B.pz1 + pzo), az, bz, pn);
                bstore_l2(B.rT, own_m ? B.pz1 + pzo : ERTM_NOWRITE, pn);
            }
        }
    }
    float irx, irz;
    {
        float s = ir_m + l_ir[tid + 1];
        irx = (B.c < B.nxp - 1) ? 0.5f * s : ir_m;
        float t = ir_m + ir_m1;
        irz = (m < nzp - 1) ? 0.5f * t : ir_m;
    }
    const float bxv = B.dt * irx;
    const float bzv = B.dt * irz;
    const float sx = a + b;
    const float sz = cc + d;
    const float ux = bxv * sx;
    const float uz = bzv * sz;
    const float vx_new = vx_old + ux;
    const float vz_new = vz_old + uz;
    // force_z source: added to the stored vz after the stresses of this
    // step have used the pre-source value (ertm/propagation.py:870-871)
    float vzs = vz_new;
    if (B.src_col && m == B.src_iz && B.mech == ERTM_SRC_FORCE_Z) vzs = vzs + B.amp;
    bstore(B.rT, B.c4own + (row_m ? m * ld4 + B.fVX : ERTM_BAD), vx_new);
    bstore(B.rT, B.c4own + (row_m ? m * ld4 + B.fVZ : ERTM_BAD), vzs);
#pragma unroll
    for (int q = 0; q < QX - 1; ++q) X.q_vx[q] = X.q_vx[q + 1];
    X.q_vx[QX - 1] = vx_new;
#pragma unroll
    for (int q = 0; q < QZ - 1; ++q) X.q_vz[q] = X.q_vz[q + 1];
    X.q_vz[QZ - 1] = vz_new;

    // ---- stress of row k
    if (STRESS) {
        const bool inz_k = k >= K && k < nzp - K;
        const bool own_k = B.own;
        float ex = dminus<K>(l_vx, tid, 1, B.cf, B.hx, B.rhx);
        float ez = dminus<K>(X.q_vz, K, D, B.cf, B.hz, B.rhz);
        float e = dplus<K>(X.q_vx, K - 1 + XO, D, B.cf, B.hz, B.rhz);
        const float ex_plain = ex, ez_plain = ez;
        float f = dplus<K>(l_vz, tid, 1, B.cf, B.hx, B.rhx);
        ex = B.inx ? ex : 0.0f;
        f = B.inx ? f : 0.0f;
        ez = inz_k ? ez : 0.0f;
        e = inz_k ? e : 0.0f;
        if (STRIP) {
            if (own_k) {
                int pzo = -1;
                if (k < B.wzt) pzo = k * ld4 + B.c4;
                else if (B.wzb > 0 && k >= nzp - B.wzb) pzo = (k - nzp + B.wzt + B.wzb) * ld4 + B.c4;
                const int pxo = B.jx4 >= 0 ? k * B.xrow4 + B.jx4 : -1;
                float pn;
                if (pxo >= 0) {
                    ex = cpml_r(ex, bload(B.rS, B.px2 + pxo), B.axc, B.bxc, pn);
                    bstore_l2(B.rT, B.px2 + pxo, pn);
                    f = cpml_r(f, bload(B.rS, B.px3 + pxo), B.axc, B.bxc, pn);
                    bstore_l2(B.rT, B.px3 + pxo, pn);
                }
                if (pzo >= 0) {
                    const float az = B.AZ[k], bz = B.BZ[k];
                    ez = cpml_r(ez, bload(B.rS, B.pz2 + pzo), az, bz, pn);
                    bstore_l2(B.rT, B.pz2 + pzo, pn);
                    e = cpml_r(e, bload(B.rS, B.pz3 + pzo), az, bz, pn);
                    bstore_l2(B.rT, B.pz3 + pzo, pn);
                }
            }
        }
        const float two_mu = 2.0f * mu_k;
        const float l2m = lam_k + two_mu;
        float mx0, mx1;  // mu on the cell corner: mean_z(mean_x(mu))
        {
            float s0 = mu_k + l_mu0[tid + 1];
            float s1 = mu_k1 + l_mu1[tid + 1];
            const bool xin = B.c < B.nxp - 1;
            mx0 = xin ? 0.5f * s0 : mu_k;
            mx1 = xin ? 0.5f * s1 : mu_k1;
        }
        float muxz;
        {
            float s = mx0 + mx1;
            muxz = (k < nzp - 1) ? 0.5f * s : mx0;
        }
        float t1 = l2m * ex, t2 = lam_k * ez;
        const float sxx_rate = t1 + t2;
        float t3 = lam_k * ex, t4 = l2m * ez;
        const float szz_rate = t3 + t4;
        const float sxx_inc = B.dt * sxx_rate;
        const float szz_inc = B.dt * szz_rate;
        float sxx = X.q_sxx[0] + sxx_inc;
        float szz = X.q_szz[0] + szz_inc;
        const float dtmu = B.dt * muxz;
        const float ef = e + f;
        const float sxz_inc = dtmu * ef;
        float sxz = X.q_sxz[0] + sxz_inc;
        if (B.src_col && k == B.src_iz) {
            if (B.mech == ERTM_SRC_SHEAR) {
                sxz = sxz + B.amp;
            } else if (B.mech == ERTM_SRC_EXPLOSIVE) {
                sxx = sxx + B.amp;
                szz = szz + B.amp;
            }
        }
        if (k == B.fs_row) {
            szz = 0.0f;
            sxz = 0.0f;
        }
        const int rk = k * ld4;
        bstore(B.rT, B.c4own + (rk + B.fXX), sxx);
        bstore(B.rT, B.c4own + (rk + B.fZZ), szz);
        bstore(B.rT, B.c4own + (rk + B.fXZ), sxz);
        if constexpr (IMG != ERTM_FIMG_NONE) {
            // physical cell (k - pad, jcol) of an owned lane
            const int i = k - B.pad;
            const bool cell = own_k && B.jcol >= 0 && i >= 0 && i < B.nzph;
            const int q = i * B.nxph + B.jcol;
            float fp, fs = 0.0f;
            if (B.sep == ERTM_SEP_STRESS) {
                const float ss = sxx + szz;
                fp = 0.5f * ss;
                fs = sxz;
            } else {
                fp = ex_plain + ez_plain;
                if (img_may_s<IMG>() && B.need_s) {
                    const float dzvx = dminus<K>(X.q_vx, K + XO - 1, D, B.cf, B.hz, B.rhz);
                    const float dxvz = dminus<K>(l_vz, tid, 1, B.cf, B.hx, B.rhx);
                    fs = dzvx - dxvz;
                }
            }
            if constexpr (img_op<IMG>() == ERTM_FIMG_EXTRACT) {
                bstore_l2(B.rR, cell ? 4 * q : ERTM_NOWRITE, B.swave == ERTM_WAVE_S ? fs : fp);
            } else if constexpr (img_op<IMG>() == ERTM_FIMG_RING) {
                const int r4 = cell ? i * ld4 + B.c4 : ERTM_NOWRITE;
                bstore_l2(B.rR, r4, B.rw0 == ERTM_WAVE_S ? fs : fp);
                if (B.nrw > 1)
                    bstore_l2(B.rR, cell ? B.o_num1 + r4 : ERTM_NOWRITE,
                              B.rw1 == ERTM_WAVE_S ? fs : fp);
            } else {
                const int q4 = cell ? 4 * q : ERTM_NOWRITE, q8 = 8 * q;
                const float sv = bload(B.rR, q4);
                if (B.illum & ERTM_ILLUM_SRC) {
                    const int o = cell ? B.o_ill + q8 : ERTM_NOWRITE;
                    const float sq = sv * sv;
                    bstore64(B.rA, o, bload64(B.rA, o) + (double)sq);
                }
                const float rv0 = B.rw0 == ERTM_WAVE_S ? fs : fp;
                if (B.illum & ERTM_ILLUM_REC) {
                    const int o = cell ? B.o_rill0 + q8 : ERTM_NOWRITE;
                    const float rr = rv0 * rv0;
                    bstore64(B.rA, o, bload64(B.rA, o) + (double)rr);
                }
                if (B.cross) {
                    const int o = cell ? B.o_num0 + q8 : ERTM_NOWRITE;
                    const float sr = sv * rv0;
                    bstore64(B.rA, o, bload64(B.rA, o) + (double)sr);
                }
                if (B.nrw > 1) {
                    const float rv1 = B.rw1 == ERTM_WAVE_S ? fs : fp;
                    if (B.illum & ERTM_ILLUM_REC) {
                        const int o = cell ? B.o_rill1 + q8 : ERTM_NOWRITE;
                        const float rr = rv1 * rv1;
                        bstore64(B.rA, o, bload64(B.rA, o) + (double)rr);
                    }
                    if (B.cross) {
                        const int o = cell ? B.o_num1 + q8 : ERTM_NOWRITE;
                        const float sr = sv * rv1;
                        bstore64(B.rA, o, bload64(B.rA, o) + (double)sr);
                    }
                }
            }
        }
    }
    X.ir_c = D > 0 ? ir_m1 : ir_m;  // the next row's m+1 (down) / m (up) is this one's
    X.mu_c = D > 0 ? mu_k1 : mu_k;
}

// One chunk: 2K prologue rows (velocity only), then the chunk's rows (both
// updates), streaming in direction D.
template <int K, int NT, bool STRIP, int IMG, int D, bool PAL>
__device__ __forceinline__ void el_loop(const ElBlock& B, float* lds) {
    constexpr int QX = qx_len<K, IMG, D>(), QZ = 2 * K + 1;
    const int ld4 = B.ld4;
    ElState<K, QX> X;
#pragma unroll
    for (int q = 0; q < QX; ++q) X.q_vx[q] = 0.0f;
#pragma unroll
    for (int q = 0; q < QZ; ++q) X.q_vz[q] = 0.0f;
    // first stress row of the stream (the 2K prologue rows lie outside the chunk)
    const int kbeg = D > 0 ? B.z0 - 2 * K : B.z1 - 1 + 2 * K;
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) {  // stress rows kbeg + q*D -> slots 1..2K
        const int r = (kbeg + q * D) * ld4;
        X.q_sxx[q + 1] = bload(B.rS, B.c4 + (r + B.fXX));
        X.q_szz[q + 1] = bload(B.rS, B.c4 + (r + B.fZZ));
        X.q_sxz[q + 1] = bload(B.rS, B.c4 + (r + B.fXZ));
    }
    const int ri0 = (kbeg + D * K + (D > 0 ? 0 : 1)) * ld4, rm0 = (kbeg + (D > 0 ? 0 : 1)) * ld4;
    if constexpr (PAL) {
        X.ir_c = B.ptab[mindex(B, ri0)];
        X.mu_c = B.ptab[256 + mindex(B, rm0)];
    } else {
        X.ir_c = bload(B.rIR, B.c4 + ri0);
        X.mu_c = bload(B.rMU, B.c4 + rm0);
    }
    el_fetch<K, QX, D, PAL>(B, X, kbeg);
    int k = kbeg;
    int par = 0;
    for (int p = 0; p < 2 * K; ++p, k += D, par ^= 1)
        el_iter<K, NT, STRIP, false, IMG, D, PAL>(B, X, k, lds + par * 7 * NT);
    for (int p = B.z0; p < B.z1; ++p, k += D, par ^= 1)
        el_iter<K, NT, STRIP, true, IMG, D, PAL>(B, X, k, lds + par * 7 * NT);
}

// The plain step and the no-S imaging variants fit 96 VGPRs, i.e. 5 waves
// per SIMD instead of 4; the variants with the S field would spill there and
// keep 4.
template <int K, int NT, int IMG, bool PAL>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(img_may_s<IMG>() ? 4 : 5)))
k_elastic_fused(const ErtmProp* __restrict__ props,
                                                      const ErtmImage* __restrict__ imgs,
                                                      LaunchArgs la, int zc) {
    constexpr int TX = NT - 4 * K;
    // The hardware deals workgroups to the XCDs by linear id (x fastest), so
    // the remap works on the linear id of the whole (blocks x propagators)
    // grid: each XCD gets a contiguous run of one propagator's blocks.
    const int lin = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int ys = lin / gridDim.x;
    const int bid = lin - ys * gridDim.x;
    const int slot = la_prop(la, ys);
    const ErtmProp& P = props[slot];
    const int nzp = P.nzp, nxp = P.nxp;
    const int nbands = (nxp + TX - 1) / TX;
    const int nchunks = (nzp + zc - 1) / zc;
    if (bid >= nbands * nchunks) return;
    const int band = bid / nchunks, chunk = bid - band * nchunks;
    ElBlock B;
    const int x0 = band * TX;
    B.nzp = nzp;
    B.nxp = nxp;
    B.z0 = chunk * zc;
    B.z1 = min(B.z0 + zc, nzp);
    B.tid = threadIdx.x;
    B.c = x0 - 2 * K + B.tid;
    B.c4 = 4 * B.c;
    B.colv = B.c >= 0 && B.c < nxp;
    B.velc = B.colv && B.tid >= K && B.tid < NT - K;
    B.own = B.colv && B.tid >= 2 * K && B.tid < NT - 2 * K;
    B.c4own = B.own ? B.c4 : ERTM_BAD;
    B.ld4 = 4 * P.ld;
    const int fs4 = (int)(4 * P.fstride);  // bytes between field slabs
    const int64_t sbytes = 4 * P.state_floats;
    const int cur = la_cur(la, ys);
    B.rS = make_rsrc(P.state[cur], sbytes);
    B.rT = make_rsrc(P.state[cur ^ 1], sbytes);
    const int64_t mbytes = (int64_t)nzp * P.ld * 4;
    B.rIR = make_rsrc(P.irho, mbytes);
    B.rLAM = make_rsrc(P.lam, mbytes);
    B.rMU = make_rsrc(P.mu, mbytes);
    B.fVX = ERTM_EL_VX * fs4;
    B.fVZ = ERTM_EL_VZ * fs4;
    B.fXX = ERTM_EL_SXX * fs4;
    B.fZZ = ERTM_EL_SZZ * fs4;
    B.fXZ = ERTM_EL_SXZ * fs4;
    B.AZ = gp(P.az);
    B.BZ = gp(P.bz);
    B.px0 = (int)(4 * P.psix[0]);
    B.px1 = (int)(4 * P.psix[1]);
    B.px2 = (int)(4 * P.psix[2]);
    B.px3 = (int)(4 * P.psix[3]);
    B.pz0 = (int)(4 * P.psiz[0]);
    B.pz1 = (int)(4 * P.psiz[1]);
    B.pz2 = (int)(4 * P.psiz[2]);
    B.pz3 = (int)(4 * P.psiz[3]);
    B.wx = P.wx;
    B.wzt = P.wzt;
    B.wzb = P.wzb;
    B.xrow4 = 8 * P.wx;  // bytes per x-memory row
    B.src_iz = P.src_iz;
    B.mech = P.mechanism;
    B.fs_row = P.free_surface ? P.pad : -1;
    B.dt = P.dt;
    B.hx = P.dx;
    B.hz = P.dz;
    B.rhx = P.rdx;
    B.rhz = P.rdz;
#pragma unroll
    for (int q = 0; q < 4; ++q) B.cf[q] = P.coef[q];
    B.amp = la_amp(la, ys);
    B.src_col = B.amp != 0.0f && B.c == P.src_ix;
    B.inx = B.c >= K && B.c < nxp - K;
    B.jx4 = -1;
    if (B.wx > 0 && B.colv) {
        if (B.c < B.wx) B.jx4 = B.c4;
        else if (B.c >= nxp - B.wx) B.jx4 = 4 * (B.c - nxp + 2 * B.wx);
    }
    B.axc = B.colv ? gp(P.ax)[B.c] : 0.0f;
    B.bxc = B.colv ? gp(P.bx)[B.c] : 1.0f;

    if constexpr (IMG != ERTM_FIMG_NONE) {
        const ErtmImage& I = imgs[slot];
        const int t = la_time(la, ys);
        B.pad = P.pad;
        B.nzph = P.nz;
        B.nxph = P.nx;
        B.jcol = (B.own && B.c >= P.pad && B.c < P.pad + P.nx) ? B.c - P.pad : -1;
        const int64_t ncell = (int64_t)P.nz * P.nx;
        if (img_op<IMG>() == ERTM_FIMG_RING) {
            // ring slot t mod R: nrw fields of nz rows on the state's pitch,
            // lane column c; o_num1 = second field
            const int64_t nring = (int64_t)P.nz * P.ld;
            B.rR = make_rsrc(I.rring + (int64_t)(t % I.ring_len) * I.nrw * nring,
                             4 * nring * I.nrw);
            B.o_num1 = (int)(4 * nring);
        } else {
            B.rR = make_rsrc(I.rbuf + (int64_t)(t - la_t0(la, ys, I)) * ncell, 4 * ncell);
        }
        B.sep = I.separation;
        B.swave = I.swave;
        B.nrw = I.nrw;
        B.rw0 = I.rwave[0];
        B.rw1 = I.rwave[1];
        B.illum = I.illum;
        B.cross = t >= I.start_step;
        B.need_s = img_op<IMG>() == ERTM_FIMG_EXTRACT
                       ? I.swave == ERTM_WAVE_S
                       : (I.rwave[0] == ERTM_WAVE_S || (I.nrw > 1 && I.rwave[1] == ERTM_WAVE_S));
        (void)ncell;
        if (IMG == ERTM_FIMG_CORRELATE) {
            // the accumulators are one allocation starting at num[0] (checked on the host)
            const char* a0 = (const char*)I.num[0];
            B.rA = make_rsrc(a0, 8 * ncell * (2 * I.nrw + 1));
            B.o_num0 = 0;
            B.o_num1 = I.nrw > 1 ? (int)((const char*)I.num[1] - a0) : 0;
            B.o_ill = I.ill ? (int)((const char*)I.ill - a0) : 0;
            B.o_rill0 = I.rill[0] ? (int)((const char*)I.rill[0] - a0) : 0;
            B.o_rill1 = (I.nrw > 1 && I.rill[1]) ? (int)((const char*)I.rill[1] - a0) : 0;
        }
    }

    __shared__ float lds[2 * 7 * NT];  // two buffers of 7 rows (see el_iter)
    // the medium palette: 1/rho, mu, lambda of each index (entry 0: zeros)
    __shared__ float ptab[PAL ? 768 : 1];
    if constexpr (PAL) {
        B.rMI = make_rsrc(P.midx, (int64_t)nzp * P.ld);
        gcfloat* tab = gp(P.mtab);
        for (int i = B.tid; i < 768; i += NT) ptab[i] = tab[i];
        B.ptab = ptab;
        __syncthreads();
    }
    // Does the block (columns x0-2K .. x0+TX+2K, rows z0-2K .. z1+2K) reach a
    // CPML strip?  Uniform per block.
    const int cx0 = x0 - 2 * K, cx1 = x0 + TX + 2 * K;
    const int rz0 = B.z0 - 2 * K, rz1 = B.z1 + 2 * K;
    const bool strip = (B.wx > 0 && (cx0 < B.wx || cx1 > nxp - B.wx)) ||
                       (B.wzt > 0 && rz0 < B.wzt) || (B.wzb > 0 && rz1 > nzp - B.wzb);
    // Alternate chunks stream in opposite directions, so the two blocks on
    // either side of a chunk seam read the seam's halo rows at about the same
    // time (both at their start, or both at their end) and the second read
    // hits in L2.
    const bool up = (chunk & 1) != 0;
    if (strip) {
        if (up) el_loop<K, NT, true, IMG, -1, PAL>(B, lds);
        else el_loop<K, NT, true, IMG, 1, PAL>(B, lds);
    } else {
        if (up) el_loop<K, NT, false, IMG, -1, PAL>(B, lds);
        else el_loop<K, NT, false, IMG, 1, PAL>(B, lds);
    }
}

// ---------------------------------------------------------------------------
// Fused acoustic step (ertm/propagation.py:430-454): the same streaming
// scheme as the elastic kernel, with one pressure field.  With VL = L > 0
// the same kernel is the viscoacoustic GSLS step (ertm/propagation.py:686-723,
// L relaxation mechanisms): the velocity half is identical, the pressure half
// updates eps and r_1..r_L of row k and sets p = K_R*eps + sum r_l.
//   velocity of row m = k + K*D: vx += (dt*bx) * Dx+ p ; vz += (dt*bz) * Dz+ p
//   pressure of row k:           p += (dt*kappa) * (Dx- vx + Dz- vz)
// Register queues: pressure rows k .. k+2K*D (Dz+ at row m and the old
// pressure of row k), new vz rows k-K*D .. k+K*D (Dz- at row k) and new vx
// rows k .. k+(K-1)*D (vx is only differentiated in x, at row k, K rows
// after it was computed).  LDS rows per iteration: pressure and 1/rho of row m, new
// vx of row k.  Per cell 32 B of HBM traffic: p, vx, vz read and written,
// kappa and 1/rho read.
// ---------------------------------------------------------------------------
struct AcBlock {
    rsrc_t rS, rT, rIR, rKA;
    rsrc_t rMI;         // medium palette indices (acoustic; ErtmProp.midx)
    const float* ptab;  // the palette in LDS: 1/rho at 0, kappa at 256
    gcfloat* AZ;
    gcfloat* BZ;
    int nzp, nxp, z0, z1, c, c4, c4own, tid, ld4;
    int fP, fVX, fVZ;
    int px0, px1, pz0, pz1;
    int wx, wzt, wzb, xrow4, jx4;
    int src_iz, fs_row;
    bool colv, velc, own, inx, src_col;
    float dt, hx, hz, amp, axc, bxc;
    double rhx, rhz;
    float cf[4];
    // imaging (IMG != 0): replay slot or ring slot of this propagator
    rsrc_t rR;
    int pad, nzph, nxph, jcol, nrw, o_f1;
    // viscoacoustic (VL > 0): eps and r_l slabs, Q class per cell and the
    // per-class tables (a = exp(-dt/tau) in float32, delta and tau in float64)
    rsrc_t rQ;
    int fEPS, fR0, fs4;
    const __attribute__((address_space(1))) float* qa;
    const __attribute__((address_space(1))) double* qd;
    const __attribute__((address_space(1))) double* qt;
    double dt64;
};

template <int K, int VL>
struct AcState {
    float q_p[2 * K + 1];   // pressure rows k + j*D
    float q_vz[2 * K + 1];  // new vz; after the push q_vz[j] holds row k + (j - K)*D
    float q_vx[K];          // new vx; before the push q_vx[j] holds row k + j*D
    float ir_c;             // 1/rho carried from the previous row (see ac_iter)
    float n_p, n_vx, n_vz, n_ir, n_ka;  // inputs of this row
    float n_eps, n_r[VL > 0 ? VL : 1];  // viscoacoustic inputs of row k
    int n_q;
    unsigned i_ir, i_ka;  // with a medium palette: indices of n_ir, n_ka (looked up in ac_iter)
};

// Inputs of iteration k: pressure row k+2K*D, old velocities of row
// m = k+K*D, 1/rho row m+1 (down) / m (up), kappa row k.
template <int K, int VL, int D, bool PAL>
__device__ __forceinline__ void ac_fetch(const AcBlock& B, AcState<K, VL>& X, int k) {
    const int ld4 = B.ld4;
    const int rp = (k + 2 * K * D) * ld4, rv = (k + K * D) * ld4;
    const int ri = (k + K * D + (D > 0 ? 1 : 0)) * ld4;
    X.n_p = bload(B.rS, B.c4 + (rp + B.fP));
    X.n_vx = bload(B.rS, B.c4 + (rv + B.fVX));
    X.n_vz = bload(B.rS, B.c4 + (rv + B.fVZ));
    if constexpr (PAL) {
        X.i_ir = mindex(B, ri);
        X.i_ka = mindex(B, k * ld4);
    } else {
        X.n_ir = bload(B.rIR, B.c4 + ri);
        X.n_ka = bload(B.rKA, B.c4 + k * ld4);
    }
    if constexpr (VL > 0) {
        const int rk = k * ld4;
        X.n_eps = bload(B.rS, B.c4 + (rk + B.fEPS));
#pragma unroll
        for (int l = 0; l < VL; ++l) X.n_r[l] = bload(B.rS, B.c4 + (rk + B.fR0 + l * B.fs4));
        X.n_q = __builtin_bit_cast(int, bload(B.rQ, B.c4 + rk));
    }
}

// IMG: ERTM_FIMG_NONE, ERTM_FIMG_EXTRACT (the new pressure into the replay
// slot) or ERTM_FIMG_RING (into ring slot t mod R, every receiver field):
// the acoustic image field is the pressure (k_image's wave_value).
template <int K, int NT, bool STRIP, bool PRES, int IMG, int VL, int D, bool PAL>
__device__ __forceinline__ void ac_iter(const AcBlock& B, AcState<K, VL>& X, int k, float* L) {
    constexpr int QP = 2 * K + 1;
    const int tid = B.tid, ld4 = B.ld4, nzp = B.nzp;
    const int m = k + D * K;
    if constexpr (PAL) {  // this row's medium from the palette (the same float values)
        X.n_ir = B.ptab[X.i_ir];
        X.n_ka = B.ptab[256 + X.i_ka];
    }
    const float vx_old = X.n_vx, vz_old = X.n_vz, ka_k = X.n_ka;
    float eps_k = 0.0f, r_k[VL > 0 ? VL : 1];
    int q_k = 0;
    if constexpr (VL > 0) {
        eps_k = X.n_eps;
#pragma unroll
        for (int l = 0; l < VL; ++l) r_k[l] = X.n_r[l];
        q_k = X.n_q;
    }
    const float ir_m = D > 0 ? X.ir_c : X.n_ir;
    const float ir_m1 = D > 0 ? X.n_ir : X.ir_c;
#pragma unroll
    for (int q = 0; q < QP - 1; ++q) X.q_p[q] = X.q_p[q + 1];
    X.q_p[QP - 1] = X.n_p;
    ac_fetch<K, VL, D, PAL>(B, X, k + D);

    float* l_p = L;
    float* l_ir = L + NT;
    float* l_vx = L + 2 * NT;
    l_p[tid] = X.q_p[K];
    l_ir[tid] = ir_m;
    if (PRES) l_vx[tid] = X.q_vx[0];  // row k
    __syncthreads();

    // ---- velocity of row m
    const bool inz_m = m >= K && m < nzp - K;
    const bool row_m = m >= B.z0 && m < B.z1;
    const bool own_m = B.own && row_m;
    float dx = dplus<K>(l_p, tid, 1, B.cf, B.hx, B.rhx);
    float dz = dplus<K>(X.q_p, K, D, B.cf, B.hz, B.rhz);
    dx = B.inx ? dx : 0.0f;
    dz = inz_m ? dz : 0.0f;
    if (STRIP) {
        if (B.velc && m >= 0 && m < nzp) {
            int pzo = -1;
            if (m < B.wzt) pzo = m * ld4 + B.c4;
            else if (B.wzb > 0 && m >= nzp - B.wzb) pzo = (m - nzp + B.wzt + B.wzb) * ld4 + B.c4;
            const int pxo = B.jx4 >= 0 ? m * B.xrow4 + B.jx4 : -1;
            float pn;
            if (pxo >= 0) {
                dx = cpml_r(dx, bload(B.rS, B.px0 + pxo), B.axc, B.bxc, pn);
                bstore_l2(B.rT, own_m ? B.px0 + pxo : ERTM_NOWRITE, pn);
            }
            if (pzo >= 0) {
                const float az = B.AZ[m], bz = B.BZ[m];
                dz = cpml_r(dz, bload(B.rS, B.pz0 + pzo), az, bz, pn);
                bstore_l2(B.rT, own_m ? B.pz0 + pzo : ERTM_NOWRITE, pn);
            }
        }
    }
    float irx, irz;
    {
        float sx = ir_m + l_ir[tid + 1];
        irx = (B.c < B.nxp - 1) ? 0.5f * sx : ir_m;
        float sz = ir_m + ir_m1;
        irz = (m < nzp - 1) ? 0.5f * sz : ir_m;
    }
    const float bxv = B.dt * irx;
    const float bzv = B.dt * irz;
    const float ux = bxv * dx;
    const float uz = bzv * dz;
    const float vx_new = vx_old + ux;
    const float vz_new = vz_old + uz;
    bstore(B.rT, B.c4own + (row_m ? m * ld4 + B.fVX : ERTM_BAD), vx_new);
    bstore(B.rT, B.c4own + (row_m ? m * ld4 + B.fVZ : ERTM_BAD), vz_new);
#pragma unroll
    for (int q = 0; q < K - 1; ++q) X.q_vx[q] = X.q_vx[q + 1];
    X.q_vx[K - 1] = vx_new;
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) X.q_vz[q] = X.q_vz[q + 1];
    X.q_vz[2 * K] = vz_new;

    // ---- pressure of row k
    if (PRES) {
        const bool inz_k = k >= K && k < nzp - K;
        float dxv = dminus<K>(l_vx, tid, 1, B.cf, B.hx, B.rhx);
        float dzv = dminus<K>(X.q_vz, K, D, B.cf, B.hz, B.rhz);
        dxv = B.inx ? dxv : 0.0f;
        dzv = inz_k ? dzv : 0.0f;
        if (STRIP) {
            if (B.own) {
                int pzo = -1;
                if (k < B.wzt) pzo = k * ld4 + B.c4;
                else if (B.wzb > 0 && k >= nzp - B.wzb) pzo = (k - nzp + B.wzt + B.wzb) * ld4 + B.c4;
                const int pxo = B.jx4 >= 0 ? k * B.xrow4 + B.jx4 : -1;
                float pn;
                if (pxo >= 0) {
                    dxv = cpml_r(dxv, bload(B.rS, B.px1 + pxo), B.axc, B.bxc, pn);
                    bstore_l2(B.rT, B.px1 + pxo, pn);
                }
                if (pzo >= 0) {
                    const float az = B.AZ[k], bz = B.BZ[k];
                    dzv = cpml_r(dzv, bload(B.rS, B.pz1 + pzo), az, bz, pn);
                    bstore_l2(B.rT, B.pz1 + pzo, pn);
                }
            }
        }
        const float div = dxv + dzv;
        float p;
        if constexpr (VL == 0) {
            const float dk = B.dt * ka_k;
            const float inc = dk * div;
            p = X.q_p[0] + inc;
            if (B.src_col && k == B.src_iz) p = p + B.amp;
            if (k == B.fs_row) p = 0.0f;
        } else {
            // eps += dt*div ; r_l = r_l*a_l + b_l*inc ; p = K_R*eps + sum r_l, with
            // b_l = f32(K_R * delta * f32(1 - a) * tau / dt) in float64
            const float inc = B.dt * div;
            float eps = eps_k + inc;
            float tot = 0.0f;
            const int qb = (q_k < 0 ? 0 : q_k) * VL;  // lanes off the grid read class 0
#pragma unroll
            for (int l = 0; l < VL; ++l) {
                const float a = B.qa[qb + l];
                const double one_minus_a = (double)(1.0f - a);
                double bd = (double)ka_k * B.qd[qb + l];
                bd = bd * one_minus_a;
                bd = bd * B.qt[qb + l];
                bd = bd / B.dt64;
                const float bl = (float)bd;
                float r = r_k[l] * a;
                const float bi = bl * inc;
                r = r + bi;
                bstore(B.rT, B.c4own + (k * ld4 + B.fR0 + l * B.fs4), r);
                tot = (l == 0) ? r : (tot + r);
            }
            const float ke = ka_k * eps;
            p = ke + tot;
            if (B.src_col && k == B.src_iz) {
                eps = eps + (float)((double)B.amp / (double)ka_k);
                p = p + B.amp;
            }
            bstore(B.rT, B.c4own + (k * ld4 + B.fEPS), eps);
        }
        bstore(B.rT, B.c4own + (k * ld4 + B.fP), p);
        if constexpr (IMG != ERTM_FIMG_NONE) {
            const int i = k - B.pad;
            const bool cell = B.own && B.jcol >= 0 && i >= 0 && i < B.nzph;
            const int q4 = !cell ? ERTM_NOWRITE
                           : IMG == ERTM_FIMG_RING ? i * B.ld4 + B.c4 : 4 * (i * B.nxph + B.jcol);
            bstore_l2(B.rR, q4, p);
            if (IMG == ERTM_FIMG_RING && B.nrw > 1)
                bstore_l2(B.rR, cell ? B.o_f1 + q4 : ERTM_NOWRITE, p);
        }
    }
    X.ir_c = D > 0 ? ir_m1 : ir_m;
}

template <int K, int NT, bool STRIP, int IMG, int VL, int D, bool PAL>
__device__ __forceinline__ void ac_loop(const AcBlock& B, float* lds) {
    const int ld4 = B.ld4;
    AcState<K, VL> X;
#pragma unroll
    for (int q = 0; q < K; ++q) X.q_vx[q] = 0.0f;
#pragma unroll
    for (int q = 0; q <= 2 * K; ++q) X.q_vz[q] = 0.0f;
    const int kbeg = D > 0 ? B.z0 - 2 * K : B.z1 - 1 + 2 * K;
#pragma unroll
    for (int q = 0; q < 2 * K; ++q)  // pressure rows kbeg + q*D -> slots 1..2K
        X.q_p[q + 1] = bload(B.rS, B.c4 + ((kbeg + q * D) * ld4 + B.fP));
    const int ri0 = (kbeg + D * K + (D > 0 ? 0 : 1)) * ld4;
    if constexpr (PAL) X.ir_c = B.ptab[mindex(B, ri0)];
    else X.ir_c = bload(B.rIR, B.c4 + ri0);
    ac_fetch<K, VL, D, PAL>(B, X, kbeg);
    int k = kbeg;
    int par = 0;
    for (int p = 0; p < 2 * K; ++p, k += D, par ^= 1)
        ac_iter<K, NT, STRIP, false, IMG, VL, D, PAL>(B, X, k, lds + par * 3 * NT);
    for (int p = B.z0; p < B.z1; ++p, k += D, par ^= 1)
        ac_iter<K, NT, STRIP, true, IMG, VL, D, PAL>(B, X, k, lds + par * 3 * NT);
}

template <int K, int NT, int IMG, int VL, bool PAL>
__global__ void __launch_bounds__(NT) k_acoustic_fused(const ErtmProp* __restrict__ props,
                                                       const ErtmImage* __restrict__ imgs,
                                                       LaunchArgs la, int zc) {
    constexpr int TX = NT - 4 * K;
    const int lin = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int ys = lin / gridDim.x;
    const int bid = lin - ys * gridDim.x;
    const int slot = la_prop(la, ys);
    const ErtmProp& P = props[slot];
    const int nzp = P.nzp, nxp = P.nxp;
    const int nbands = (nxp + TX - 1) / TX;
    const int nchunks = (nzp + zc - 1) / zc;
    if (bid >= nbands * nchunks) return;
    const int band = bid / nchunks, chunk = bid - band * nchunks;
    AcBlock B;
    const int x0 = band * TX;
    B.nzp = nzp;
    B.nxp = nxp;
    B.z0 = chunk * zc;
    B.z1 = min(B.z0 + zc, nzp);
    B.tid = threadIdx.x;
    B.c = x0 - 2 * K + B.tid;
    B.c4 = 4 * B.c;
    B.colv = B.c >= 0 && B.c < nxp;
    B.velc = B.colv && B.tid >= K && B.tid < NT - K;
    B.own = B.colv && B.tid >= 2 * K && B.tid < NT - 2 * K;
    B.c4own = B.own ? B.c4 : ERTM_BAD;
    B.ld4 = 4 * P.ld;
    const int fs4 = (int)(4 * P.fstride);
    const int64_t sbytes = 4 * P.state_floats;
    const int cur = la_cur(la, ys);
    B.rS = make_rsrc(P.state[cur], sbytes);
    B.rT = make_rsrc(P.state[cur ^ 1], sbytes);
    const int64_t mbytes = (int64_t)nzp * P.ld * 4;
    B.rIR = make_rsrc(P.irho, mbytes);
    B.rKA = make_rsrc(P.kappa, mbytes);
    B.fP = ERTM_AC_P * fs4;
    B.fVX = ERTM_AC_VX * fs4;
    B.fVZ = ERTM_AC_VZ * fs4;
    if constexpr (VL > 0) {
        B.fs4 = fs4;
        B.fEPS = ERTM_VI_EPS * fs4;
        B.fR0 = ERTM_VI_R0 * fs4;
        B.rQ = make_rsrc(P.qidx, mbytes);
        B.qa = (const __attribute__((address_space(1))) float*)P.q_a;
        B.qd = (const __attribute__((address_space(1))) double*)P.q_delta;
        B.qt = (const __attribute__((address_space(1))) double*)P.q_tau;
        B.dt64 = P.dt64;
    }
    B.AZ = gp(P.az);
    B.BZ = gp(P.bz);
    B.px0 = (int)(4 * P.psix[0]);
    B.px1 = (int)(4 * P.psix[1]);
    B.pz0 = (int)(4 * P.psiz[0]);
    B.pz1 = (int)(4 * P.psiz[1]);
    B.wx = P.wx;
    B.wzt = P.wzt;
    B.wzb = P.wzb;
    B.xrow4 = 8 * P.wx;
    B.src_iz = P.src_iz;
    B.fs_row = P.free_surface ? P.pad : -1;
    B.dt = P.dt;
    B.hx = P.dx;
    B.hz = P.dz;
    B.rhx = P.rdx;
    B.rhz = P.rdz;
#pragma unroll
    for (int q = 0; q < 4; ++q) B.cf[q] = P.coef[q];
    B.amp = la_amp(la, ys);
    B.src_col = B.amp != 0.0f && B.c == P.src_ix;
    B.inx = B.c >= K && B.c < nxp - K;
    B.jx4 = -1;
    if (B.wx > 0 && B.colv) {
        if (B.c < B.wx) B.jx4 = B.c4;
        else if (B.c >= nxp - B.wx) B.jx4 = 4 * (B.c - nxp + 2 * B.wx);
    }
    B.axc = B.colv ? gp(P.ax)[B.c] : 0.0f;
    B.bxc = B.colv ? gp(P.bx)[B.c] : 1.0f;
    if constexpr (IMG != ERTM_FIMG_NONE) {
        const ErtmImage& I = imgs[slot];
        const int t = la_time(la, ys);
        const int64_t ncell = (int64_t)P.nz * P.nx;
        B.pad = P.pad;
        B.nzph = P.nz;
        B.nxph = P.nx;
        B.jcol = (B.own && B.c >= P.pad && B.c < P.pad + P.nx) ? B.c - P.pad : -1;
        B.nrw = I.nrw;
        // ring rows on the state's pitch (lane column c), replay slots (nz, nx)
        const int64_t nring = (int64_t)P.nz * P.ld;
        B.o_f1 = (int)(4 * nring);
        if (IMG == ERTM_FIMG_RING)
            B.rR = make_rsrc(I.rring + (int64_t)(t % I.ring_len) * I.nrw * nring,
                             4 * nring * I.nrw);
        else
            B.rR = make_rsrc(I.rbuf + (int64_t)(t - la_t0(la, ys, I)) * ncell, 4 * ncell);
    }

    __shared__ float lds[2 * 3 * NT];
    // the medium palette: 1/rho and kappa of each index (entry 0: zeros)
    __shared__ float ptab[PAL ? 512 : 1];
    if constexpr (PAL) {
        B.rMI = make_rsrc(P.midx, (int64_t)nzp * P.ld);
        gcfloat* tab = gp(P.mtab);
        for (int i = B.tid; i < 512; i += NT) ptab[i] = tab[i];
        B.ptab = ptab;
        __syncthreads();
    }
    const int cx0 = x0 - 2 * K, cx1 = x0 + TX + 2 * K;
    const int rz0 = B.z0 - 2 * K, rz1 = B.z1 + 2 * K;
    const bool strip = (B.wx > 0 && (cx0 < B.wx || cx1 > nxp - B.wx)) ||
                       (B.wzt > 0 && rz0 < B.wzt) || (B.wzb > 0 && rz1 > nzp - B.wzb);
    const bool up = (chunk & 1) != 0;
    if (strip) {
        if (up) ac_loop<K, NT, true, IMG, VL, -1, PAL>(B, lds);
        else ac_loop<K, NT, true, IMG, VL, 1, PAL>(B, lds);
    } else {
        if (up) ac_loop<K, NT, false, IMG, VL, -1, PAL>(B, lds);
        else ac_loop<K, NT, false, IMG, VL, 1, PAL>(B, lds);
    }
}

template <int K, int NT, int VL>
static bool launch_acoustic_k(const ErtmProp* dprops, const ErtmImage* dimg,
                              const ErtmProp* hprops, const LaunchArgs& la, int zc, int img,
                              hipStream_t s);

template <int K, int NT>
static int blocks_for(const ErtmProp& P, int zc) {
    constexpr int TX = NT - 4 * K;
    return ((P.nxp + TX - 1) / TX) * ((P.nzp + zc - 1) / zc);
}

// img: any extraction kind -> pressure into the replay slot, any ring kind ->
// pressure into the ring; the fused correlation is elastic only.
template <int K, int NT, int VL>
static bool launch_acoustic_k(const ErtmProp* dprops, const ErtmImage* dimg,
                              const ErtmProp* hprops, const LaunchArgs& la, int zc, int img,
                              hipStream_t s) {
    int nb = 0;
    bool pal = VL == 0;  // acoustic, and every member has a medium palette
    for (int k = 0; k < la.n; ++k) {
        const ErtmProp& P = hprops[la_prop(la, k)];
        nb = max(nb, blocks_for<K, NT>(P, zc));
        pal = pal && P.npal > 0 && P.npal < 256 && P.midx && P.mtab;
    }
    const dim3 g(nb, la.n);
    if (img != ERTM_FIMG_NONE && img != ERTM_FIMG_EXTRACT && img != ERTM_FIMG_EXTRACT_NOS &&
        img != ERTM_FIMG_RING && img != ERTM_FIMG_RING_NOS)
        return false;
    const int kind = img == ERTM_FIMG_NONE ? ERTM_FIMG_NONE
                   : (img == ERTM_FIMG_RING || img == ERTM_FIMG_RING_NOS) ? ERTM_FIMG_RING
                                                                            : ERTM_FIMG_EXTRACT;
    if constexpr (VL == 0) {
        if (pal) {
            if (kind == ERTM_FIMG_NONE)
                hipLaunchKernelGGL((k_acoustic_fused<K, NT, ERTM_FIMG_NONE, 0, true>), g, dim3(NT),
                                   0, s, dprops, dimg, la, zc);
            else if (kind == ERTM_FIMG_EXTRACT)
                hipLaunchKernelGGL((k_acoustic_fused<K, NT, ERTM_FIMG_EXTRACT, 0, true>), g,
                                   dim3(NT), 0, s, dprops, dimg, la, zc);
            else
                hipLaunchKernelGGL((k_acoustic_fused<K, NT, ERTM_FIMG_RING, 0, true>), g, dim3(NT),
                                   0, s, dprops, dimg, la, zc);
            return true;
        }
    }
    if (kind == ERTM_FIMG_NONE)
        hipLaunchKernelGGL((k_acoustic_fused<K, NT, ERTM_FIMG_NONE, VL, false>), g, dim3(NT), 0, s,
                           dprops, dimg, la, zc);
    else if (kind == ERTM_FIMG_EXTRACT)
        hipLaunchKernelGGL((k_acoustic_fused<K, NT, ERTM_FIMG_EXTRACT, VL, false>), g, dim3(NT), 0,
                           s, dprops, dimg, la, zc);
    else
        hipLaunchKernelGGL((k_acoustic_fused<K, NT, ERTM_FIMG_RING, VL, false>), g, dim3(NT), 0, s,
                           dprops, dimg, la, zc);
    return true;
}

// Fused launch of a pressure engine: acoustic (VL = 0) or viscoacoustic with
// L = 1..4 mechanisms (more fall back to the two-launch kernels).
template <int NT>
static bool launch_pressure(const ErtmProp* dprops, const ErtmImage* dimg, const ErtmProp* hprops,
                            const LaunchArgs& la, int zc, int img, int engine, int nmech,
                            hipStream_t s) {
    if (engine == ERTM_ACOUSTIC)
        return launch_acoustic_k<4, NT, 0>(dprops, dimg, hprops, la, zc, img, s);
    switch (nmech) {
        case 1: return launch_acoustic_k<4, NT, 1>(dprops, dimg, hprops, la, zc, img, s);
        case 2: return launch_acoustic_k<4, NT, 2>(dprops, dimg, hprops, la, zc, img, s);
        case 3: return launch_acoustic_k<4, NT, 3>(dprops, dimg, hprops, la, zc, img, s);
        case 4: return launch_acoustic_k<4, NT, 4>(dprops, dimg, hprops, la, zc, img, s);
        default: return false;
    }
}

template <int K, int NT, int IMG>
static void launch_fused_k(const ErtmProp* dprops, const ErtmImage* dimg,
                           const ErtmProp* hprops, const LaunchArgs& la, int zc, hipStream_t s) {
    int nb = 0;
    bool pal = true;  // every member has a medium palette
    for (int k = 0; k < la.n; ++k) {
        const ErtmProp& P = hprops[la_prop(la, k)];
        nb = max(nb, blocks_for<K, NT>(P, zc));
        pal = pal && P.npal > 0 && P.npal < 256 && P.midx && P.mtab;
    }
    if (pal)
        hipLaunchKernelGGL((k_elastic_fused<K, NT, IMG, true>), dim3(nb, la.n), dim3(NT), 0, s,
                           dprops, dimg, la, zc);
    else
        hipLaunchKernelGGL((k_elastic_fused<K, NT, IMG, false>), dim3(nb, la.n), dim3(NT), 0, s,
                           dprops, dimg, la, zc);
}

template <int K, int NT>
static bool launch_fused_img(const ErtmProp* dprops, const ErtmImage* dimg,
                             const ErtmProp* hprops, const LaunchArgs& la, int zc, int img,
                             hipStream_t s) {
    switch (img) {
        case ERTM_FIMG_NONE: launch_fused_k<K, NT, ERTM_FIMG_NONE>(dprops, dimg, hprops, la, zc, s); return true;
        case ERTM_FIMG_EXTRACT: launch_fused_k<K, NT, ERTM_FIMG_EXTRACT>(dprops, dimg, hprops, la, zc, s); return true;
        case ERTM_FIMG_CORRELATE: launch_fused_k<K, NT, ERTM_FIMG_CORRELATE>(dprops, dimg, hprops, la, zc, s); return true;
        case ERTM_FIMG_RING: launch_fused_k<K, NT, ERTM_FIMG_RING>(dprops, dimg, hprops, la, zc, s); return true;
        case ERTM_FIMG_EXTRACT_NOS: launch_fused_k<K, NT, ERTM_FIMG_EXTRACT_NOS>(dprops, dimg, hprops, la, zc, s); return true;
        case ERTM_FIMG_RING_NOS: launch_fused_k<K, NT, ERTM_FIMG_RING_NOS>(dprops, dimg, hprops, la, zc, s); return true;
        default: return false;
    }
}

// variant w = 1, 2, 3, 4 -> 256, 128, 64, 192 threads per block; zc = rows per chunk.
// img (ERTM_FIMG_*) != NONE also images the new state; la.t then holds each
// propagator's time index.
bool launch_step_fused(const ErtmProp* dprops, const ErtmProp* hprops, const LaunchArgs& la,
                       int order, int variant, int zc, const ErtmImage* dimg, int img,
                       hipStream_t s) {
    if (order != 8 || zc < 8) return false;
    // byte offsets, including rows up to 32 past the grid and ERTM_BAD, must
    // stay below 2^31; every state buffer below 2^30 bytes keeps ERTM_BAD out
    // of range
    for (int k = 0; k < la.n; ++k) {
        const ErtmProp& P = hprops[la_prop(la, k)];
        if (4 * P.state_floats + 128 * (int64_t)P.ld >= ((int64_t)1 << 30)) return false;
    }
    // one engine (and mechanism count) per launch
    const int engine = hprops[la_prop(la, 0)].engine, nmech = hprops[la_prop(la, 0)].nmech;
    for (int k = 0; k < la.n; ++k)
        if (hprops[la_prop(la, k)].engine != engine ||
            (engine == ERTM_VISCO && hprops[la_prop(la, k)].nmech != nmech))
            return false;
    if (engine == ERTM_ACOUSTIC || engine == ERTM_VISCO) {
        switch (variant) {
            case 1: return launch_pressure<256>(dprops, dimg, hprops, la, zc, img, engine, nmech, s);
            case 2: return launch_pressure<128>(dprops, dimg, hprops, la, zc, img, engine, nmech, s);
            case 3: return launch_pressure<64>(dprops, dimg, hprops, la, zc, img, engine, nmech, s);
            case 4: return launch_pressure<192>(dprops, dimg, hprops, la, zc, img, engine, nmech, s);
            default: return false;
        }
    }
    if (engine != ERTM_ELASTIC) return false;
    switch (variant) {
        case 1: return launch_fused_img<4, 256>(dprops, dimg, hprops, la, zc, img, s);
        case 2: return launch_fused_img<4, 128>(dprops, dimg, hprops, la, zc, img, s);
        case 3: return launch_fused_img<4, 64>(dprops, dimg, hprops, la, zc, img, s);
        case 4: return launch_fused_img<4, 192>(dprops, dimg, hprops, la, zc, img, s);
        default: return false;
    }
}

}  // namespace ertm
