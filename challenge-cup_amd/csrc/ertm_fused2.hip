// Two time steps per launch for the elastic P-SV engine (order 8).
//
// The one-step kernel (ertm_fused.hip) reads the 5 state fields and the 3
// medium fields of a cell and writes the 5 state fields back once per step.
// This kernel advances the state by two steps per launch: state[cur] (time
// t) -> state[cur ^ 1] (time t + 2), so each state field is read and written
// once per TWO steps.  The state after the first step is never stored.
//
// Block: a band of NT lanes (column c = x0 - 4K + lane) streaming a chunk of
// rows [z0, z1), as in the one-step kernel, with two pipelined stages per
// row iteration j:
//   stage A = the one-step update of the old state: velocity of row j + K*D,
//             stresses of row j; it keeps its results in registers (never
//             stores the fields);
//   stage B = the one-step update of stage A's output, 2K rows behind:
//             velocity of row j - K*D, stresses of row j - 2K*D; it stores
//             the state at t + 2 for the rows and columns the block owns.
// Stage B's z-derivatives read stage A's new stresses from register queues
// (stage A's stress row j enters them the iteration it is computed) and its
// old velocities are the oldest entries of stage A's velocity queues; its
// x-derivatives go through LDS like stage A's.  Halo: the stencil radius K
// applies four times (two half-steps per step), so the block owns lanes
// [4K, NT - 4K) and recomputes 4K rows beyond each chunk end.
//
// Validity: stage A's velocities are exact on lanes [K, NT - K), its
// stresses on [2K, NT - 2K); stage B's velocities on [3K, NT - 3K), its
// stresses on [4K, NT - 4K) -- the owned lanes.  Rows: stage A computes
// every row stage B needs from old rows it loaded (loads outside the grid
// return 0, as in the one-step kernel).
//
// CPML memories: stage A reads the old memories (state[cur]) and writes the
// memories after the first step to a scratch buffer of the batch (the CPML
// part of a state buffer), on every lane where its value is exact (velocity
// terms on [K, NT - K), stress terms on [2K, NT - 2K)); stage B reads them
// back (same lane, 2K iterations later, program order) and writes the final
// memories to state[cur ^ 1] for owned cells.  Blocks whose halos overlap
// write the same cells of the scratch, always with the same value (the same
// operands and operations), so the order of those writes does not matter.
//
// Every per-cell expression is the one-step kernel's (ertm_fused.hip,
// el_iter), operand for operand, so two steps here are bit-identical to two
// launches of the one-step kernel and to the reference.
//
// Imaging (replay pass): with IMG = ERTM_FIMG_EXTRACT[_NOS] each stage also
// writes the source-side field of its new state into the replay slot of its
// time (t - t0 and t + 1 - t0).
#include "ertm_fused_common.h"

#ifndef ERTM_2STEP_WAVES
#define ERTM_2STEP_WAVES 3
#endif

namespace ertm {

struct El2Block {
    ElBlock G;       // geometry, media, S = state[cur] (rS), T = state[cur ^ 1] (rT)
    rsrc_t rP;       // CPML memories after the first step (scratch)
    int psi0;        // byte offset of the CPML memories inside a state buffer
    bool ownA;       // lanes whose first-step stresses are exact: [2K, NT - 2K)
    bool own2;       // lanes this block owns: [4K, NT - 4K)
    int c4own2;      // c4 on owned lanes, ERTM_BAD elsewhere
    float amp1, amp2;
    bool srcA, srcB;  // source column, first / second step
    rsrc_t rR2;       // replay slot of the second step
};

template <int K>
struct El2State {
    // stage A: old stress rows j .. j + 2K*D; first-step velocity rows
    // j - K*D .. j + K*D (after the push)
    float a_sxx[2 * K + 1], a_szz[2 * K + 1], a_sxz[2 * K + 1];
    float a_vx[2 * K + 1], a_vz[2 * K + 1];
    float a_irc, a_muc;
    float n_sxx, n_szz, n_sxz, n_vx, n_vz, n_ir, n_mu, n_lam;  // stage A inputs of row j
    // stage B: first-step stress rows j - 2K*D .. j (after the push);
    // second-step velocity rows kB - K*D .. kB + K*D, kB = j - 2K*D
    float b_sxx[2 * K + 1], b_szz[2 * K + 1], b_sxz[2 * K + 1];
    float b_vx[2 * K + 1], b_vz[2 * K + 1];
    float b_irc, b_muc;
    float m_ir, m_mu, m_lam;  // stage B medium inputs of its row
};

// Stage A inputs of iteration k: stress row k + 2K*D, velocity row
// m = k + K*D, 1/rho row m + 1 (down) / m (up), mu row k + 1 (down) / k (up),
// lambda row k (as el_fetch).
template <int K, int D>
__device__ __forceinline__ void el2_fetch_a(const ElBlock& B, El2State<K>& X, int k) {
    const int ld4 = B.ld4;
    const int rs = (k + 2 * K * D) * ld4, rv = (k + K * D) * ld4;
    const int ri = (k + K * D + (D > 0 ? 1 : 0)) * ld4, rm = (k + (D > 0 ? 1 : 0)) * ld4;
    X.n_sxx = bload(B.rS, B.c4 + (rs + B.fXX));
    X.n_szz = bload(B.rS, B.c4 + (rs + B.fZZ));
    X.n_sxz = bload(B.rS, B.c4 + (rs + B.fXZ));
    X.n_vx = bload(B.rS, B.c4 + (rv + B.fVX));
    X.n_vz = bload(B.rS, B.c4 + (rv + B.fVZ));
    X.n_ir = bload(B.rIR, B.c4 + ri);
    X.n_mu = bload(B.rMU, B.c4 + rm);
    X.n_lam = bload(B.rLAM, B.c4 + k * ld4);
}

// Stage B medium inputs of its row kB (same rows relative to kB as stage A's).
template <int K, int D>
__device__ __forceinline__ void el2_fetch_b(const ElBlock& B, El2State<K>& X, int kb) {
    const int ld4 = B.ld4;
    X.m_ir = bload(B.rIR, B.c4 + (kb + K * D + (D > 0 ? 1 : 0)) * ld4);
    X.m_mu = bload(B.rMU, B.c4 + (kb + (D > 0 ? 1 : 0)) * ld4);
    X.m_lam = bload(B.rLAM, B.c4 + kb * ld4);
}

// Velocity of row m from LDS rows (stress x-neighbours, 1/rho) and the
// stress queues (z-derivatives at queue index K), el_iter's arithmetic.
// CPML: memories read from rPR at byte offset (slot offset - shR), written
// to rPW at (slot offset - shW) where wr holds.
template <int K, bool STRIP, int D>
__device__ __forceinline__ void el2_velocity(const ElBlock& B, int m, const float* l_sxx,
                                             const float* l_sxz, const float* l_ir,
                                             const float* q_sxz, const float* q_szz, float ir_m,
                                             float ir_m1, float vx_old, float vz_old,
                                             rsrc_t rPR, int shR, rsrc_t rPW, int shW, bool wr,
                                             float& vx_new, float& vz_new) {
    const int tid = B.tid, ld4 = B.ld4, nzp = B.nzp;
    const bool inz_m = m >= K && m < nzp - K;
    float a = dplus<K>(l_sxx, tid, 1, B.cf, B.hx, B.rhx);
    float b = dminus<K>(q_sxz, K, D, B.cf, B.hz, B.rhz);
    float cc = dminus<K>(l_sxz, tid, 1, B.cf, B.hx, B.rhx);
    float d = dplus<K>(q_szz, K, D, B.cf, B.hz, B.rhz);
    a = B.inx ? a : 0.0f;
    cc = B.inx ? cc : 0.0f;
    b = inz_m ? b : 0.0f;
    d = inz_m ? d : 0.0f;
    if (STRIP) {
        if (B.velc && m >= 0 && m < nzp) {
            int pzo = -1;
            if (m < B.wzt) pzo = m * ld4 + B.c4;
            else if (B.wzb > 0 && m >= nzp - B.wzb) pzo = (m - nzp + B.wzt + B.wzb) * ld4 + B.c4;
            const int pxo = B.jx4 >= 0 ? m * B.xrow4 + B.jx4 : -1;
            float pn;
            if (pxo >= 0) {
                a = cpml_r(a, bload(rPR, B.px0 - shR + pxo), B.axc, B.bxc, pn);
                bstore_l2(rPW, wr ? B.px0 - shW + pxo : ERTM_NOWRITE, pn);
                cc = cpml_r(cc, bload(rPR, B.px1 - shR + pxo), B.axc, B.bxc, pn);
                bstore_l2(rPW, wr ? B.px1 - shW + pxo : ERTM_NOWRITE, pn);
            }
            if (pzo >= 0) {
                const float az = B.AZ[m], bz = B.BZ[m];
                b = cpml_r(b, bload(rPR, B.pz0 - shR + pzo), az, bz, pn);
                bstore_l2(rPW, wr ? B.pz0 - shW + pzo : ERTM_NOWRITE, pn);
                d = cpml_r(d, bload(rPR, B.pz1 - shR + pzo), az, bz, pn);
                bstore_l2(rPW, wr ? B.pz1 - shW + pzo : ERTM_NOWRITE, pn);
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
    vx_new = vx_old + ux;
    vz_new = vz_old + uz;
}

// Stresses of row k from the new velocities (LDS rows of row k, queues with
// row k at index K), el_iter's arithmetic; returns the plain divcurl terms
// for imaging.  CPML as in el2_velocity, only where `own` holds (reads and
// writes).
template <int K, bool STRIP, int D>
__device__ __forceinline__ void el2_stress(const ElBlock& B, int k, bool own, const float* l_vx,
                                           const float* l_vz, const float* l_mu0,
                                           const float* l_mu1, const float* q_vx,
                                           const float* q_vz, float mu_k, float mu_k1,
                                           float lam_k, float sxx_old, float szz_old,
                                           float sxz_old, bool src, float amp, rsrc_t rPR,
                                           int shR, rsrc_t rPW, int shW, float& sxx, float& szz,
                                           float& sxz, float& ex_plain, float& ez_plain) {
    const int tid = B.tid, ld4 = B.ld4, nzp = B.nzp;
    const bool inz_k = k >= K && k < nzp - K;
    float ex = dminus<K>(l_vx, tid, 1, B.cf, B.hx, B.rhx);
    float ez = dminus<K>(q_vz, K, D, B.cf, B.hz, B.rhz);
    float e = dplus<K>(q_vx, K, D, B.cf, B.hz, B.rhz);
    ex_plain = ex;
    ez_plain = ez;
    float f = dplus<K>(l_vz, tid, 1, B.cf, B.hx, B.rhx);
    ex = B.inx ? ex : 0.0f;
    f = B.inx ? f : 0.0f;
    ez = inz_k ? ez : 0.0f;
    e = inz_k ? e : 0.0f;
    if (STRIP) {
        // stage A also computes rows beyond the grid (never read by an exact
        // value): no memories there
        if (own && k >= 0 && k < nzp) {
            int pzo = -1;
            if (k < B.wzt) pzo = k * ld4 + B.c4;
            else if (B.wzb > 0 && k >= nzp - B.wzb) pzo = (k - nzp + B.wzt + B.wzb) * ld4 + B.c4;
            const int pxo = B.jx4 >= 0 ? k * B.xrow4 + B.jx4 : -1;
            float pn;
            if (pxo >= 0) {
                ex = cpml_r(ex, bload(rPR, B.px2 - shR + pxo), B.axc, B.bxc, pn);
                bstore_l2(rPW, B.px2 - shW + pxo, pn);
                f = cpml_r(f, bload(rPR, B.px3 - shR + pxo), B.axc, B.bxc, pn);
                bstore_l2(rPW, B.px3 - shW + pxo, pn);
            }
            if (pzo >= 0) {
                const float az = B.AZ[k], bz = B.BZ[k];
                ez = cpml_r(ez, bload(rPR, B.pz2 - shR + pzo), az, bz, pn);
                bstore_l2(rPW, B.pz2 - shW + pzo, pn);
                e = cpml_r(e, bload(rPR, B.pz3 - shR + pzo), az, bz, pn);
                bstore_l2(rPW, B.pz3 - shW + pzo, pn);
            }
        }
    }
    const float two_mu = 2.0f * mu_k;
    const float l2m = lam_k + two_mu;
    float mx0, mx1;
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
    sxx = sxx_old + sxx_inc;
    szz = szz_old + szz_inc;
    const float dtmu = B.dt * muxz;
    const float ef = e + f;
    const float sxz_inc = dtmu * ef;
    sxz = sxz_old + sxz_inc;
    // point source as selects (no divergent branch inside the row loop)
    const bool at = src && k == B.src_iz;
    const bool shear = at && B.mech == ERTM_SRC_SHEAR;
    const bool expl = at && B.mech == ERTM_SRC_EXPLOSIVE;
    const float sxz_s = sxz + amp, sxx_s = sxx + amp, szz_s = szz + amp;
    sxz = shear ? sxz_s : sxz;
    sxx = expl ? sxx_s : sxx;
    szz = expl ? szz_s : szz;
    if (k == B.fs_row) {
        szz = 0.0f;
        sxz = 0.0f;
    }
}

// Source-side image field of a new state at row k (k_image's wave value):
// P = (sxx + szz) / 2 (stress separation) or Dx-vx + Dz-vz (divcurl), S =
// sxz or Dz-vx - Dx-vz.
template <int K, int IMG, int D>
__device__ __forceinline__ float el2_image_field(const ElBlock& B, float sxx, float szz, float sxz,
                                                 float ex_plain, float ez_plain, const float* q_vx,
                                                 const float* l_vz) {
    if (B.sep == ERTM_SEP_STRESS) {
        if (B.swave == ERTM_WAVE_S) return sxz;
        const float ss = sxx + szz;
        return 0.5f * ss;
    }
    if (img_may_s<IMG>() && B.swave == ERTM_WAVE_S) {
        const float dzvx = dminus<K>(q_vx, K, D, B.cf, B.hz, B.rhz);
        const float dxvz = dminus<K>(l_vz, B.tid, 1, B.cf, B.hx, B.rhx);
        return dzvx - dxvz;
    }
    return ex_plain + ez_plain;
}

// One row iteration j.  PA: 0 = stage A velocity only, 1 = stage A full.
// PB: 0 = stage B off (stage A's stresses still enter its queues), 1 =
// stage B velocity only, 2 = stage B full.  L = this iteration's LDS buffer
// (14 rows of NT floats: 7 for each stage).
template <int K, int NT, bool STRIP, int PA, int PB, int IMG, int D>
__device__ __forceinline__ void el2_iter(const El2Block& E, El2State<K>& X, int j, float* L) {
    constexpr int Q = 2 * K + 1;
    const ElBlock& B = E.G;
    const int tid = B.tid, ld4 = B.ld4;
    // ---- stage A inputs
    const int mA = j + D * K;
    const float vxA_old = X.n_vx, vzA_old = X.n_vz, lamA = X.n_lam;
    const float irA_m = D > 0 ? X.a_irc : X.n_ir;
    const float irA_m1 = D > 0 ? X.n_ir : X.a_irc;
    const float muA_k = D > 0 ? X.a_muc : X.n_mu;
    const float muA_k1 = D > 0 ? X.n_mu : X.a_muc;
#pragma unroll
    for (int q = 0; q < Q - 1; ++q) {
        X.a_sxx[q] = X.a_sxx[q + 1];
        X.a_szz[q] = X.a_szz[q + 1];
        X.a_sxz[q] = X.a_sxz[q + 1];
    }
    X.a_sxx[Q - 1] = X.n_sxx;
    X.a_szz[Q - 1] = X.n_szz;
    X.a_sxz[Q - 1] = X.n_sxz;
    el2_fetch_a<K, D>(B, X, j + D);
    // ---- stage B inputs (rows kB = j - 2K*D, mB = j - K*D)
    const int kB = j - 2 * K * D, mB = j - K * D;
    const float lamB = X.m_lam;
    const float irB_m = D > 0 ? X.b_irc : X.m_ir;
    const float irB_m1 = D > 0 ? X.m_ir : X.b_irc;
    const float muB_k = D > 0 ? X.b_muc : X.m_mu;
    const float muB_k1 = D > 0 ? X.m_mu : X.b_muc;
    if (PB > 0) el2_fetch_b<K, D>(B, X, kB + D);

    // ---- LDS rows, one barrier (buffers alternate between iterations)
    float* a_lsxx = L;
    float* a_lsxz = L + NT;
    float* a_lir = L + 2 * NT;
    float* a_lvx = L + 3 * NT;
    float* a_lvz = L + 4 * NT;
    float* a_lmu0 = L + 5 * NT;
    float* a_lmu1 = L + 6 * NT;
    float* b_lsxx = L + 7 * NT;
    float* b_lsxz = L + 8 * NT;
    float* b_lir = L + 9 * NT;
    float* b_lvx = L + 10 * NT;
    float* b_lvz = L + 11 * NT;
    float* b_lmu0 = L + 12 * NT;
    float* b_lmu1 = L + 13 * NT;
    a_lsxx[tid] = X.a_sxx[K];
    a_lsxz[tid] = X.a_sxz[K];
    a_lir[tid] = irA_m;
    if (PA > 0) {
        a_lvx[tid] = X.a_vx[K + 1];  // row j (the queue is pushed below)
        a_lvz[tid] = X.a_vz[K + 1];
        a_lmu0[tid] = muA_k;
        a_lmu1[tid] = muA_k1;
    }
    if (PB > 0) {
        b_lsxx[tid] = X.b_sxx[K + 1];  // first-step stress row mB (pushed below)
        b_lsxz[tid] = X.b_sxz[K + 1];
        b_lir[tid] = irB_m;
    }
    if (PB > 1) {
        b_lvx[tid] = X.b_vx[K + 1];  // second-step velocity row kB
        b_lvz[tid] = X.b_vz[K + 1];
        b_lmu0[tid] = muB_k;
        b_lmu1[tid] = muB_k1;
    }
    __syncthreads();

    // ---- stage A: velocity of row mA (first step), memories to the scratch
    {
        float vx, vz;
        el2_velocity<K, STRIP, D>(B, mA, a_lsxx, a_lsxz, a_lir, X.a_sxz, X.a_szz, irA_m, irA_m1,
                                  vxA_old, vzA_old, B.rS, 0, E.rP, E.psi0, B.velc, vx, vz);
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) {
            X.a_vx[q] = X.a_vx[q + 1];
            X.a_vz[q] = X.a_vz[q + 1];
        }
        X.a_vx[Q - 1] = vx;
        X.a_vz[Q - 1] = vz;
    }
    // ---- stage A: stresses of row j (first step)
    if (PA > 0) {
        float sxx, szz, sxz, exp_, ezp_;
        el2_stress<K, STRIP, D>(B, j, E.ownA, a_lvx, a_lvz, a_lmu0, a_lmu1, X.a_vx, X.a_vz, muA_k,
                                muA_k1, lamA, X.a_sxx[0], X.a_szz[0], X.a_sxz[0], E.srcA, E.amp1,
                                B.rS, 0, E.rP, E.psi0, sxx, szz, sxz, exp_, ezp_);
        if constexpr (IMG != ERTM_FIMG_NONE) {
            const int i = j - B.pad;
            const bool cell = B.jcol >= 0 && j >= B.z0 && j < B.z1 && i >= 0 && i < B.nzph;
            const float v = el2_image_field<K, IMG, D>(B, sxx, szz, sxz, exp_, ezp_, X.a_vx,
                                                       a_lvz);
            bstore_l2(B.rR, cell ? 4 * (i * B.nxph + B.jcol) : ERTM_NOWRITE, v);
        }
        // the first-step stresses of row j enter stage B's queues
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) {
            X.b_sxx[q] = X.b_sxx[q + 1];
            X.b_szz[q] = X.b_szz[q + 1];
            X.b_sxz[q] = X.b_sxz[q + 1];
        }
        X.b_sxx[Q - 1] = sxx;
        X.b_szz[Q - 1] = szz;
        X.b_sxz[Q - 1] = sxz;
    }
    X.a_irc = D > 0 ? irA_m1 : irA_m;
    X.a_muc = D > 0 ? muA_k1 : muA_k;
    if (PB == 0) return;

    // ---- stage B: velocity of row mB (second step).  Its old velocity is
    // stage A's first-step velocity of row mB (the oldest queue entry), with
    // a force_z source of the first step added as the stored value has it.
    const bool row_m = mB >= B.z0 && mB < B.z1;
    {
        float vzB_old = X.a_vz[0];
        if (E.srcA && mB == B.src_iz && B.mech == ERTM_SRC_FORCE_Z) vzB_old = vzB_old + E.amp1;
        float vx, vz;
        el2_velocity<K, STRIP, D>(B, mB, b_lsxx, b_lsxz, b_lir, X.b_sxz, X.b_szz, irB_m, irB_m1,
                                  X.a_vx[0], vzB_old, E.rP, E.psi0, B.rT, 0, E.own2 && row_m, vx,
                                  vz);
        float vzs = vz;
        if (E.srcB && mB == B.src_iz && B.mech == ERTM_SRC_FORCE_Z) vzs = vzs + E.amp2;
        bstore(B.rT, E.c4own2 + (row_m ? mB * ld4 + B.fVX : ERTM_BAD), vx);
        bstore(B.rT, E.c4own2 + (row_m ? mB * ld4 + B.fVZ : ERTM_BAD), vzs);
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) {
            X.b_vx[q] = X.b_vx[q + 1];
            X.b_vz[q] = X.b_vz[q + 1];
        }
        X.b_vx[Q - 1] = vx;
        X.b_vz[Q - 1] = vz;
    }
    // ---- stage B: stresses of row kB (second step)
    if (PB > 1) {
        float sxx, szz, sxz, exp_, ezp_;
        el2_stress<K, STRIP, D>(B, kB, E.own2, b_lvx, b_lvz, b_lmu0, b_lmu1, X.b_vx, X.b_vz,
                                muB_k, muB_k1, lamB, X.b_sxx[0], X.b_szz[0], X.b_sxz[0], E.srcB,
                                E.amp2, E.rP, E.psi0, B.rT, 0, sxx, szz, sxz, exp_, ezp_);
        const int rk = kB * ld4;
        bstore(B.rT, E.c4own2 + (rk + B.fXX), sxx);
        bstore(B.rT, E.c4own2 + (rk + B.fZZ), szz);
        bstore(B.rT, E.c4own2 + (rk + B.fXZ), sxz);
        if constexpr (IMG != ERTM_FIMG_NONE) {
            const int i = kB - B.pad;
            const bool cell = B.jcol >= 0 && i >= 0 && i < B.nzph;
            const float v = el2_image_field<K, IMG, D>(B, sxx, szz, sxz, exp_, ezp_, X.b_vx,
                                                       b_lvz);
            bstore_l2(E.rR2, cell ? 4 * (i * B.nxph + B.jcol) : ERTM_NOWRITE, v);
        }
    }
    X.b_irc = D > 0 ? irB_m1 : irB_m;
    X.b_muc = D > 0 ? muB_k1 : muB_k;
}

// One chunk in stream direction D: ZC + 6K row iterations (stage A runs 4K
// rows ahead of the chunk's first stress row, stage B starts 2K rows later).
template <int K, int NT, bool STRIP, int IMG, int D>
__device__ __forceinline__ void el2_loop(const El2Block& E, float* lds) {
    constexpr int Q = 2 * K + 1;
    const ElBlock& B = E.G;
    const int ld4 = B.ld4;
    El2State<K> X;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        X.a_vx[q] = 0.0f;
        X.a_vz[q] = 0.0f;
        X.b_sxx[q] = 0.0f;
        X.b_szz[q] = 0.0f;
        X.b_sxz[q] = 0.0f;
        X.b_vx[q] = 0.0f;
        X.b_vz[q] = 0.0f;
    }
    const int kbeg = D > 0 ? B.z0 - 4 * K : B.z1 - 1 + 4 * K;
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) {  // old stress rows kbeg + q*D -> slots 1..2K
        const int r = (kbeg + q * D) * ld4;
        X.a_sxx[q + 1] = bload(B.rS, B.c4 + (r + B.fXX));
        X.a_szz[q + 1] = bload(B.rS, B.c4 + (r + B.fZZ));
        X.a_sxz[q + 1] = bload(B.rS, B.c4 + (r + B.fXZ));
    }
    X.a_irc = bload(B.rIR, B.c4 + (kbeg + D * K + (D > 0 ? 0 : 1)) * ld4);
    X.a_muc = bload(B.rMU, B.c4 + (kbeg + (D > 0 ? 0 : 1)) * ld4);
    el2_fetch_a<K, D>(B, X, kbeg);
    // stage B's first row: kB0 = kbeg + 2K*D (at iteration 4K)
    const int kb0 = kbeg + 2 * K * D;
    X.b_irc = bload(B.rIR, B.c4 + (kb0 + D * K + (D > 0 ? 0 : 1)) * ld4);
    X.b_muc = bload(B.rMU, B.c4 + (kb0 + (D > 0 ? 0 : 1)) * ld4);
    el2_fetch_b<K, D>(B, X, kb0);
    int j = kbeg;
    int par = 0;
    for (int p = 0; p < 2 * K; ++p, j += D, par ^= 1)
        el2_iter<K, NT, STRIP, 0, 0, IMG, D>(E, X, j, lds + par * 14 * NT);
    for (int p = 0; p < 2 * K; ++p, j += D, par ^= 1)
        el2_iter<K, NT, STRIP, 1, 0, IMG, D>(E, X, j, lds + par * 14 * NT);
    for (int p = 0; p < 2 * K; ++p, j += D, par ^= 1)
        el2_iter<K, NT, STRIP, 1, 1, IMG, D>(E, X, j, lds + par * 14 * NT);
    for (int p = B.z0; p < B.z1; ++p, j += D, par ^= 1)
        el2_iter<K, NT, STRIP, 1, 2, IMG, D>(E, X, j, lds + par * 14 * NT);
}

template <int K, int NT, int IMG>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(ERTM_2STEP_WAVES)))
k_elastic_2step(const ErtmProp* __restrict__ props, const ErtmImage* __restrict__ imgs,
                LaunchArgs la, El2Args a2, int zc) {
    constexpr int TX = NT - 8 * K;
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
    El2Block E;
    ElBlock& B = E.G;
    const int x0 = band * TX;
    B.nzp = nzp;
    B.nxp = nxp;
    B.z0 = chunk * zc;
    B.z1 = min(B.z0 + zc, nzp);
    B.tid = threadIdx.x;
    B.c = x0 - 4 * K + B.tid;
    B.c4 = 4 * B.c;
    B.colv = B.c >= 0 && B.c < nxp;
    B.velc = B.colv && B.tid >= K && B.tid < NT - K;
    E.ownA = B.colv && B.tid >= 2 * K && B.tid < NT - 2 * K;
    E.own2 = B.colv && B.tid >= 4 * K && B.tid < NT - 4 * K;
    B.own = E.own2;
    E.c4own2 = E.own2 ? B.c4 : ERTM_BAD;
    B.c4own = E.c4own2;
    B.ld4 = 4 * P.ld;
    const int fs4 = (int)(4 * P.fstride);
    const int64_t sbytes = 4 * P.state_floats;
    const int cur = la_cur(la, ys);
    B.rS = make_rsrc(P.state[cur], sbytes);
    B.rT = make_rsrc(P.state[cur ^ 1], sbytes);
    // the scratch holds the CPML part of a state buffer: byte offsets in it
    // are state offsets minus psi0 (offsets below psi0 wrap out of range)
    E.psi0 = P.nfields * fs4;
    E.rP = make_rsrc(a2.psi1[ys], sbytes - (int64_t)E.psi0);
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
    B.xrow4 = 8 * P.wx;
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
    E.amp1 = la_amp(la, ys);
    E.amp2 = a2.amp2[ys];
    E.srcA = E.amp1 != 0.0f && B.c == P.src_ix;
    E.srcB = E.amp2 != 0.0f && B.c == P.src_ix;
    B.inx = B.c >= K && B.c < nxp - K;
    B.jx4 = -1;
    if (B.wx > 0 && B.colv) {
        if (B.c < B.wx) B.jx4 = B.c4;
        else if (B.c >= nxp - B.wx) B.jx4 = 4 * (B.c - nxp + 2 * B.wx);
    }
    B.axc = B.colv ? gp(P.ax)[B.c] : 0.0f;
    B.bxc = B.colv ? gp(P.bx)[B.c] : 1.0f;
    B.jcol = -1;
    if constexpr (IMG != ERTM_FIMG_NONE) {
        const ErtmImage& I = imgs[slot];
        const int t = la_time(la, ys);
        B.pad = P.pad;
        B.nzph = P.nz;
        B.nxph = P.nx;
        B.jcol = (E.own2 && B.c >= P.pad && B.c < P.pad + P.nx) ? B.c - P.pad : -1;
        const int64_t ncell = (int64_t)P.nz * P.nx;
        const int t0 = la_t0(la, ys, I);
        B.rR = make_rsrc(I.rbuf + (int64_t)(t - t0) * ncell, 4 * ncell);
        E.rR2 = make_rsrc(I.rbuf + (int64_t)(t + 1 - t0) * ncell, 4 * ncell);
        B.sep = I.separation;
        B.swave = I.swave;
    }

    __shared__ float lds[2 * 14 * NT];
    // CPML strips within the block's reach (columns x0-4K .. x0+TX+4K, rows
    // z0-4K .. z1+4K)?  Uniform per block.
    const int cx0 = x0 - 4 * K, cx1 = x0 + TX + 4 * K;
    const int rz0 = B.z0 - 4 * K, rz1 = B.z1 + 4 * K;
    const bool strip = (B.wx > 0 && (cx0 < B.wx || cx1 > nxp - B.wx)) ||
                       (B.wzt > 0 && rz0 < B.wzt) || (B.wzb > 0 && rz1 > nzp - B.wzb);
    const bool up = (chunk & 1) != 0;
    if (strip) {
        if (up) el2_loop<K, NT, true, IMG, -1>(E, lds);
        else el2_loop<K, NT, true, IMG, 1>(E, lds);
    } else {
        if (up) el2_loop<K, NT, false, IMG, -1>(E, lds);
        else el2_loop<K, NT, false, IMG, 1>(E, lds);
    }
}

template <int K, int NT>
static int blocks2_for(const ErtmProp& P, int zc) {
    constexpr int TX = NT - 8 * K;
    return ((P.nxp + TX - 1) / TX) * ((P.nzp + zc - 1) / zc);
}

template <int K, int NT, int IMG>
static void launch2_k(const ErtmProp* dprops, const ErtmImage* dimg, const ErtmProp* hprops,
                      const LaunchArgs& la, const El2Args& a2, int zc, hipStream_t s) {
    int nb = 0;
    for (int k = 0; k < la.n; ++k) nb = max(nb, blocks2_for<K, NT>(hprops[la_prop(la, k)], zc));
    hipLaunchKernelGGL((k_elastic_2step<K, NT, IMG>), dim3(nb, la.n), dim3(NT), 0, s, dprops,
                       dimg, la, a2, zc);
}

template <int K, int NT>
static bool launch2_img(const ErtmProp* dprops, const ErtmImage* dimg, const ErtmProp* hprops,
                        const LaunchArgs& la, const El2Args& a2, int zc, int img, hipStream_t s) {
    switch (img) {
        case ERTM_FIMG_NONE: launch2_k<K, NT, ERTM_FIMG_NONE>(dprops, dimg, hprops, la, a2, zc, s); return true;
        case ERTM_FIMG_EXTRACT: launch2_k<K, NT, ERTM_FIMG_EXTRACT>(dprops, dimg, hprops, la, a2, zc, s); return true;
        case ERTM_FIMG_EXTRACT_NOS: launch2_k<K, NT, ERTM_FIMG_EXTRACT_NOS>(dprops, dimg, hprops, la, a2, zc, s); return true;
        default: return false;
    }
}

// Two elastic steps (order 8) for the launch's propagators: variant 2 = 128
// lanes per block (the only shape built), zc rows per chunk; img = NONE or an
// extraction kind.  The host checks that every propagator is elastic at
// order 8 with a scratch for its CPML memories.
bool launch_step2_fused(const ErtmProp* dprops, const ErtmProp* hprops, const LaunchArgs& la,
                        const El2Args& a2, int variant, int zc, const ErtmImage* dimg, int img,
                        hipStream_t s) {
    if (zc < 8) return false;
    for (int k = 0; k < la.n; ++k) {
        const ErtmProp& P = hprops[la_prop(la, k)];
        if (P.engine != ERTM_ELASTIC || P.order != 8 || !a2.psi1[k]) return false;
        if (4 * P.state_floats + 128 * (int64_t)P.ld >= ((int64_t)1 << 30)) return false;
    }
    if (variant != 2) return false;
    return launch2_img<4, 128>(dprops, dimg, hprops, la, a2, zc, img, s);
}

}  // namespace ertm
