// Internal interface between the C-ABI layer and the kernel translation units.
#pragma once

#include "ertm_device.h"

namespace ertm {

// Per-launch arguments: the launch covers propagators [base, base + n) of the
// batch table; propagator base+k is at time index t[k] with source amplitude
// amp[k]; states are read from state[cur] and written to state[cur ^ 1].
//
// Device clock (launches inside a captured time loop, replayed as a graph):
// with clk != nullptr the time index and amplitude come from device memory,
// so the same launch serves every step it is replayed for.  With
// q = clk[0] + qoff (clk[0] is the chunk's first step, advanced by k_tick
// after each replay; qoff the launch's step within the chunk), every
// propagator is at time clk[1] + clk[2] * q and its amplitude is
// ta[q * ntab + base + k] (0 without a table).
//
// Lanes (remap = 1, never with the device clock): member k of the launch is
// propagator slot[k] of the table, at parity cur ^ bit k of flip, and its
// replay slots start at time t0[k] instead of its image descriptor's t0.
struct LaunchArgs {
    int32_t base;
    int32_t n;
    int32_t cur;
    int32_t op;
    int32_t t[ERTM_MAX_LAUNCH];
    float amp[ERTM_MAX_LAUNCH];
    const int32_t* clk;
    const float* ta;
    int32_t ntab;
    int32_t qoff;
    int32_t remap;
    uint64_t flip;
    int32_t slot[ERTM_MAX_LAUNCH];
    int32_t t0[ERTM_MAX_LAUNCH];
};

// Table index of launch member k.
__host__ __device__ __forceinline__ int la_prop(const LaunchArgs& la, int k) {
    return la.remap ? la.slot[k] : la.base + k;
}

// State parity member k reads (it writes the other one).
__host__ __device__ __forceinline__ int la_cur(const LaunchArgs& la, int k) {
    return la.cur ^ (int)((la.flip >> k) & 1u);
}

// Time index of replay slot 0 for member k.  Both candidates are read and
// the values selected, and the result is made wave-uniform: selecting
// between the two addresses (a kernel argument and the image table) gives a
// flat load, whose result the compiler treats as per-lane, and the replay
// slot's buffer descriptor derived from it then went through a waterfall
// loop around every image store of the fused step.
__device__ __forceinline__ int la_t0(const LaunchArgs& la, int k, const ErtmImage& I) {
    const int from_lanes = la.t0[k];
    const int from_image = I.t0;
    return __builtin_amdgcn_readfirstlane(la.remap ? from_lanes : from_image);
}

__device__ __forceinline__ int la_time(const LaunchArgs& la, int k) {
    if (!la.clk) return la.t[k];
    return la.clk[1] + la.clk[2] * (la.clk[0] + la.qoff);
}

__device__ __forceinline__ float la_amp(const LaunchArgs& la, int k) {
    if (!la.clk) return la.amp[k];
    return la.ta ? la.ta[(int64_t)(la.clk[0] + la.qoff) * la.ntab + la.base + k] : 0.0f;
}

struct DerivCoef {
    float c[4];
    double c64[4];
};

int tiles_of(const ErtmProp& P);
void launch_step(const ErtmProp* dprops, const LaunchArgs& la, int order, int max_tiles,
                 bool any_force, hipStream_t s);
bool launch_step_fused(const ErtmProp* dprops, const ErtmProp* hprops, const LaunchArgs& la,
                       int order, int variant, int zc, const ErtmImage* dimg, int img,
                       hipStream_t s);
// Extra per-launch arguments of the two-step elastic kernel (ertm_fused2.hip):
// the second step's source amplitudes and, per propagator, a scratch for the
// CPML memories after the first step (the CPML part of a state buffer).
struct El2Args {
    float amp2[ERTM_MAX_LAUNCH];
    float* psi1[ERTM_MAX_LAUNCH];
};
// Two steps state[cur] -> state[cur ^ 1] (time t -> t + 2, la.t = t, la.amp
// = first step's amplitudes); img = ERTM_FIMG_NONE or an extraction kind
// (slots t - t0 and t + 1 - t0).  Returns false if it cannot launch.
bool launch_step2_fused(const ErtmProp* dprops, const ErtmProp* hprops, const LaunchArgs& la,
                        const El2Args& a2, int variant, int zc, const ErtmImage* dimg, int img,
                        hipStream_t s);
void launch_sample(const ErtmProp* dprops, const ErtmRecv* drecv, const LaunchArgs& la,
                   int max_samp, hipStream_t s);
void launch_inject(const ErtmProp* dprops, const ErtmRecv* drecv, const LaunchArgs& la,
                   int max_cells, hipStream_t s);
void launch_post_norm(const double* num, const double* ill, const double* rill, int64_t n,
                      double eps, int div_ill, int div_rill, double* out, hipStream_t s);
void launch_post_lap(const double* a, int nz, int nx, double dxx, double dzz, int to_complex,
                     double* out, hipStream_t s);
void launch_post_to_complex(const double* a, int64_t n, double* c, hipStream_t s);
void launch_post_response(double* c, int nz, int nx, double dx, double dz, double frac,
                          hipStream_t s);
void launch_post_finish(const double* src, int from_complex, int nz, int nx, int npml,
                        int mask_top, float* out, hipStream_t s);
void launch_prep_mute_taper(float* rec, int nt, int nrec, const double* t, const double* t_mute,
                            double taper, int w_end, hipStream_t s);
void launch_prep_rotate(float* vx, float* vz, int64_t n, double c, double s_, hipStream_t s);
void launch_resample(const float* in, int n_in, int nrec, const double* w, const int32_t* off,
                     int n_out, int taps, float* out, int out_rows, hipStream_t s);
// rscale: null, or the per-ring-slot factors of the viscoacoustic stabiliser
// (R rows of rs_n floats, one per batch-table entry)
void launch_image_flush(const ErtmProp* dprops, const ErtmImage* dimg, const LaunchArgs& la,
                        int max_cells, const float* rscale, int rs_n, hipStream_t s);
void launch_image(const ErtmProp* dprops, const ErtmImage* dimg, const LaunchArgs& la, int order,
                  int max_cells, hipStream_t s);
void launch_absmax(const float* x, int64_t n, float* out, hipStream_t s);
// out[k] = max |field slab| of launch member k (nzp * ld floats), one launch
void launch_absmax_batch(const ErtmProp* dprops, const LaunchArgs& la, int field, int64_t max_n,
                         float* out, hipStream_t s);
// device clock: advance clk[0] by `steps` (after a replayed chunk)
void launch_tick(int32_t* clk, int steps, hipStream_t s);
// checkpoint save inside a captured loop: state[cur] of every propagator of
// the launch goes to arenas[k] + slot * stride when its time t satisfies
// t >= origin and (t - origin) % interval == 0 (slot = (t - origin) / interval)
void launch_checkpoint(const ErtmProp* dprops, const LaunchArgs& la, float* const* darenas,
                       int64_t stride, int origin, int interval, int64_t max_floats,
                       hipStream_t s);
// n copies of n4[k] float4 groups, src[k] -> dst[k], in one launch (both
// 16-byte aligned); the host checks alignment and size before using it
struct CopyArgs {
    int32_t n;
    const float4* src[ERTM_MAX_LAUNCH];
    float4* dst[ERTM_MAX_LAUNCH];
    int64_t n4[ERTM_MAX_LAUNCH];
};
void launch_copy_batch(const CopyArgs& a, hipStream_t s);
void launch_derivative(const void* f, void* out, int n0, int n1, int axis, int plus, int order,
                       const DerivCoef& c, double h, int is_f64, hipStream_t s);
// Viscoacoustic stabiliser pieces (ertm/propagation.py:726-749).
void launch_stab_control(const ErtmProp* dprops, const LaunchArgs& la, const float* peaks,
                         double* ctl, int step_index, double cap, int lowpass_now, float* scales,
                         float* rscale, hipStream_t s);
void launch_stab_scale(const ErtmProp* dprops, const LaunchArgs& la, const float* scales,
                       int max_tiles, hipStream_t s);
// A group of the stabiliser's low-pass: members k < n are propagators prop[k]
// of the batch table, all of padded shape (nzp, nxp), filtered at cutoff
// ratio[k]; their p and eps fields in state[cur] are transforms 2k and 2k + 1
// of the batched plans.  Real buffer: 2n * nzp * nxp doubles; complex buffer:
// 2n * nzp * (nxp / 2 + 1) double2.
constexpr int ERTM_LP_GROUP = 16;
struct LpArgs {
    int32_t n, nzp, nxp, cur;
    int32_t ring_slot;  // >= 0: also write the filtered p into this ring slot
    int32_t prop[ERTM_LP_GROUP];
    double ratio[ERTM_LP_GROUP];
};
void launch_lowpass_prep(const ErtmProp* dprops, const LpArgs& a, double* real, hipStream_t s);
void launch_lowpass_response(const ErtmProp* dprops, const LpArgs& a, void* cplx, hipStream_t s);
void launch_lowpass_finish(const ErtmProp* dprops, const ErtmImage* dimg, const LpArgs& a,
                           const double* real, hipStream_t s);

}  // namespace ertm
