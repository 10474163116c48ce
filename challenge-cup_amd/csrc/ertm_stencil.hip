// Stencil kernels: one time step of the acoustic, elastic P-SV and
// viscoacoustic GSLS engines (state[cur] -> state[cur^1]).
//
// Each step is two launches: the velocity update (reads the old pressure /
// stress around each cell, writes the new velocities) and the pressure /
// stress update (reads the new velocities around each cell).  Ping-pong state
// buffers make every launch race-free.  All CPML memories live in compact
// strips (only where the reference's recursion is not an exact no-op).
//
// Reference arithmetic followed (float32, operation by operation):
//   acoustic      ertm/propagation.py:430-454
//   viscoacoustic ertm/propagation.py:686-723
//   elastic       ertm/propagation.py:840-883
//   CPML          ertm/numerics.py:221-227, :283-308
#include "ertm_kernels.h"

namespace ertm {

struct Cell {
    int slot, i, j;
    bool ok;
};

// Block -> (propagator, cell).  Block (64 x 4 threads) covers a 64 x 4 tile;
// blockIdx.y is the propagator inside the launch chunk.
__device__ __forceinline__ Cell cell_of(const ErtmProp* props, const LaunchArgs& la) {
    Cell c;
    c.slot = la_prop(la, blockIdx.y);
    const ErtmProp& P = props[c.slot];
    int nbx = (P.nxp + 63) >> 6;
    int bz = blockIdx.x / nbx;
    int bx = blockIdx.x - bz * nbx;
    c.j = bx * 64 + threadIdx.x;
    c.i = bz * 4 + threadIdx.y;
    c.ok = (c.i < P.nzp) && (c.j < P.nxp);
    return c;
}

// ---------------------------------------------------------------------------
// velocity update
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_velocity(const ErtmProp* __restrict__ props,
                                                  LaunchArgs la) {
    Cell c = cell_of(props, la);
    if (!c.ok) return;
    const ErtmProp& P = props[c.slot];
    const int i = c.i, j = c.j;
    const int cur = la_cur(la, blockIdx.y);
    const float* __restrict__ S = P.state[cur];
    float* __restrict__ T = P.state[cur ^ 1];
    const int64_t ld = P.ld;
    const int64_t o = (int64_t)i * ld + j;
    const float dt = P.dt;
    const bool inx = (j >= K) && (j < P.nxp - K);
    const bool inz = (i >= K) && (i < P.nzp - K);
    const int64_t px = psix_index(P, i, j);
    const int64_t pz = psiz_index(P, i, j);

    if (P.engine != ERTM_ELASTIC) {
        // acoustic / viscoacoustic: vx += (dt*bx) * D'x+ p ; vz += (dt*bz) * D'z+ p
        const float* p = S + (int64_t)ERTM_AC_P * P.fstride;
        float dx = inx ? dplus<K>(p, o, 1, P.coef, P.dx, P.rdx) : 0.0f;
        if (px >= 0) dx = cpml(dx, S[P.psix[0] + px], P.ax[j], P.bx[j], &T[P.psix[0] + px]);
        float dz = inz ? dplus<K>(p, o, ld, P.coef, P.dz, P.rdz) : 0.0f;
        if (pz >= 0) dz = cpml(dz, S[P.psiz[0] + pz], P.az[i], P.bz[i], &T[P.psiz[0] + pz]);
        float bxv = dt * avg_x(P.irho, P, i, j);
        float bzv = dt * avg_z(P.irho, P, i, j);
        const int64_t fx = (int64_t)ERTM_AC_VX * P.fstride, fz = (int64_t)ERTM_AC_VZ * P.fstride;
        float ux = bxv * dx;
        float uz = bzv * dz;
        T[fx + o] = S[fx + o] + ux;
        T[fz + o] = S[fz + o] + uz;
    } else {
        const float* sxx = S + (int64_t)ERTM_EL_SXX * P.fstride;
        const float* szz = S + (int64_t)ERTM_EL_SZZ * P.fstride;
        const float* sxz = S + (int64_t)ERTM_EL_SXZ * P.fstride;
        // vx += (dt*bx) * (D'x+ sxx + D'z- sxz)
        float a = inx ? dplus<K>(sxx, o, 1, P.coef, P.dx, P.rdx) : 0.0f;
        if (px >= 0) a = cpml(a, S[P.psix[0] + px], P.ax[j], P.bx[j], &T[P.psix[0] + px]);
        float b = inz ? dminus<K>(sxz, o, ld, P.coef, P.dz, P.rdz) : 0.0f;
        if (pz >= 0) b = cpml(b, S[P.psiz[0] + pz], P.az[i], P.bz[i], &T[P.psiz[0] + pz]);
        // vz += (dt*bz) * (D'x- sxz + D'z+ szz)
        float cc = inx ? dminus<K>(sxz, o, 1, P.coef, P.dx, P.rdx) : 0.0f;
        if (px >= 0) cc = cpml(cc, S[P.psix[1] + px], P.ax[j], P.bx[j], &T[P.psix[1] + px]);
        float d = inz ? dplus<K>(szz, o, ld, P.coef, P.dz, P.rdz) : 0.0f;
        if (pz >= 0) d = cpml(d, S[P.psiz[1] + pz], P.az[i], P.bz[i], &T[P.psiz[1] + pz]);
        float bxv = dt * avg_x(P.irho, P, i, j);
        float bzv = dt * avg_z(P.irho, P, i, j);
        const int64_t fx = (int64_t)ERTM_EL_VX * P.fstride, fz = (int64_t)ERTM_EL_VZ * P.fstride;
        float sx = a + b;
        float sz = cc + d;
        float ux = bxv * sx;
        float uz = bzv * sz;
        T[fx + o] = S[fx + o] + ux;
        T[fz + o] = S[fz + o] + uz;
    }
}

// ---------------------------------------------------------------------------
// pressure / stress update (+ point source, free surface)
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_stress(const ErtmProp* __restrict__ props,
                                                LaunchArgs la) {
    Cell c = cell_of(props, la);
    if (!c.ok) return;
    const ErtmProp& P = props[c.slot];
    const int i = c.i, j = c.j;
    const int cur = la_cur(la, blockIdx.y);
    const float* __restrict__ S = P.state[cur];
    float* __restrict__ T = P.state[cur ^ 1];
    const int64_t ld = P.ld;
    const int64_t o = (int64_t)i * ld + j;
    const float dt = P.dt;
    const bool inx = (j >= K) && (j < P.nxp - K);
    const bool inz = (i >= K) && (i < P.nzp - K);
    const int64_t px = psix_index(P, i, j);
    const int64_t pz = psiz_index(P, i, j);
    const float amp = la_amp(la, blockIdx.y);
    const bool at_src = (amp != 0.0f) && (i == P.src_iz) && (j == P.src_ix);

    if (P.engine == ERTM_ELASTIC) {
        const float* vx = T + (int64_t)ERTM_EL_VX * P.fstride;  // new velocities
        const float* vz = T + (int64_t)ERTM_EL_VZ * P.fstride;
        float ex = inx ? dminus<K>(vx, o, 1, P.coef, P.dx, P.rdx) : 0.0f;
        if (px >= 0) ex = cpml(ex, S[P.psix[2] + px], P.ax[j], P.bx[j], &T[P.psix[2] + px]);
        float ez = inz ? dminus<K>(vz, o, ld, P.coef, P.dz, P.rdz) : 0.0f;
        if (pz >= 0) ez = cpml(ez, S[P.psiz[2] + pz], P.az[i], P.bz[i], &T[P.psiz[2] + pz]);
        float e = inz ? dplus<K>(vx, o, ld, P.coef, P.dz, P.rdz) : 0.0f;
        if (pz >= 0) e = cpml(e, S[P.psiz[3] + pz], P.az[i], P.bz[i], &T[P.psiz[3] + pz]);
        float f = inx ? dplus<K>(vz, o, 1, P.coef, P.dx, P.rdx) : 0.0f;
        if (px >= 0) f = cpml(f, S[P.psix[3] + px], P.ax[j], P.bx[j], &T[P.psix[3] + px]);
        const float lam = P.lam[o];
        const float mu = P.mu[o];
        const float two_mu = 2.0f * mu;
        const float l2m = lam + two_mu;
        const float muxz = avg_xz(P.mu, P, i, j);
        const int64_t fxx = (int64_t)ERTM_EL_SXX * P.fstride;
        const int64_t fzz = (int64_t)ERTM_EL_SZZ * P.fstride;
        const int64_t fxz = (int64_t)ERTM_EL_SXZ * P.fstride;
        float t1 = l2m * ex, t2 = lam * ez;
        float sxx_rate = t1 + t2;
        float t3 = lam * ex, t4 = l2m * ez;
        float szz_rate = t3 + t4;
        float sxx_inc = dt * sxx_rate;
        float szz_inc = dt * szz_rate;
        float sxx = S[fxx + o] + sxx_inc;
        float szz = S[fzz + o] + szz_inc;
        float dtmu = dt * muxz;
        float ef = e + f;
        float sxz_inc = dtmu * ef;
        float sxz = S[fxz + o] + sxz_inc;
        if (at_src) {
            if (P.mechanism == ERTM_SRC_SHEAR) {
                sxz = sxz + amp;
            } else if (P.mechanism == ERTM_SRC_EXPLOSIVE) {
                sxx = sxx + amp;
                szz = szz + amp;
            }  // force_z is added to vz by k_point_source after this launch
        }
        if (P.free_surface && i == P.pad) {
            szz = 0.0f;
            sxz = 0.0f;
        }
        T[fxx + o] = sxx;
        T[fzz + o] = szz;
        T[fxz + o] = sxz;
        return;
    }

    const float* vx = T + (int64_t)ERTM_AC_VX * P.fstride;
    const float* vz = T + (int64_t)ERTM_AC_VZ * P.fstride;
    float dxv = inx ? dminus<K>(vx, o, 1, P.coef, P.dx, P.rdx) : 0.0f;
    if (px >= 0) dxv = cpml(dxv, S[P.psix[1] + px], P.ax[j], P.bx[j], &T[P.psix[1] + px]);
    float dzv = inz ? dminus<K>(vz, o, ld, P.coef, P.dz, P.rdz) : 0.0f;
    if (pz >= 0) dzv = cpml(dzv, S[P.psiz[1] + pz], P.az[i], P.bz[i], &T[P.psiz[1] + pz]);
    const float div = dxv + dzv;
    const float kap = P.kappa[o];
    const int64_t fp = (int64_t)ERTM_AC_P * P.fstride;

    if (P.engine == ERTM_ACOUSTIC) {
        float dk = dt * kap;
        float inc = dk * div;
        float p = S[fp + o] + inc;
        if (at_src) p = p + amp;
        if (P.free_surface && i == P.pad) p = 0.0f;
        T[fp + o] = p;
        return;
    }

    // viscoacoustic GSLS: eps += dt*div ; r_l = r_l*a_l + b_l*inc ; p = K_R*eps + sum r_l
    const float inc = dt * div;
    const int64_t fe = (int64_t)ERTM_VI_EPS * P.fstride;
    float eps = S[fe + o] + inc;
    const int q = P.qidx[o];
    float tot = 0.0f;
    for (int l = 0; l < P.nmech; ++l) {
        const int64_t fr = (int64_t)(ERTM_VI_R0 + l) * P.fstride;
        const float a = P.q_a[q * P.nmech + l];
        // b_l = f32(K_R * delta * f32(1 - a) * tau / dt) in float64, as
        // ertm/propagation.py:662-668 builds it.
        const double one_minus_a = (double)(1.0f - a);
        double bd = (double)kap * P.q_delta[q * P.nmech + l];
        bd = bd * one_minus_a;
        bd = bd * P.q_tau[q * P.nmech + l];
        bd = bd / P.dt64;
        const float bl = (float)bd;
        float r = S[fr + o] * a;
        float bi = bl * inc;
        r = r + bi;
        T[fr + o] = r;
        tot = (l == 0) ? r : (tot + r);
    }
    float ke = kap * eps;
    float p = ke + tot;
    if (at_src) {
        eps = eps + (float)((double)amp / (double)kap);
        p = p + amp;
    }
    T[fe + o] = eps;
    T[fp + o] = p;
}

// force_z source on vz after the stress update (ertm/propagation.py:870-871).
__global__ void k_point_source(const ErtmProp* __restrict__ props, LaunchArgs la) {
    const int slot = la_prop(la, blockIdx.x);
    const ErtmProp& P = props[slot];
    const float amp = la_amp(la, blockIdx.x);
    if (threadIdx.x != 0 || amp == 0.0f || P.engine != ERTM_ELASTIC ||
        P.mechanism != ERTM_SRC_FORCE_Z || P.src_iz < 0)
        return;
    float* T = P.state[la_cur(la, blockIdx.x) ^ 1] + (int64_t)ERTM_EL_VZ * P.fstride;
    const int64_t o = (int64_t)P.src_iz * P.ld + P.src_ix;
    T[o] = T[o] + amp;
}

template <int K>
static void launch_step_k(const ErtmProp* dprops, const LaunchArgs& la, int max_tiles,
                          bool any_force, hipStream_t s) {
    dim3 blk(64, 4);
    dim3 grd(max_tiles, la.n);
    hipLaunchKernelGGL(k_velocity<K>, grd, blk, 0, s, dprops, la);
    hipLaunchKernelGGL(k_stress<K>, grd, blk, 0, s, dprops, la);
    if (any_force) hipLaunchKernelGGL(k_point_source, dim3(la.n), dim3(64), 0, s, dprops, la);
}

void launch_step(const ErtmProp* dprops, const LaunchArgs& la, int order, int max_tiles,
                 bool any_force, hipStream_t s) {
    switch (order) {
        case 2: launch_step_k<1>(dprops, la, max_tiles, any_force, s); break;
        case 4: launch_step_k<2>(dprops, la, max_tiles, any_force, s); break;
        case 6: launch_step_k<3>(dprops, la, max_tiles, any_force, s); break;
        default: launch_step_k<4>(dprops, la, max_tiles, any_force, s); break;
    }
}

int tiles_of(const ErtmProp& P) { return ((P.nxp + 63) / 64) * ((P.nzp + 3) / 4); }

}  // namespace ertm
