/*
 * ertm_hip.h -- C-ABI of the MI355X (gfx950) RTM engine, libertm_hip.so.
 *
 * The reference package is pure Python/NumPy, so it has no FFI of its own;
 * each entry point below replaces one group of NumPy operations in the
 * reference's hot path (citations are to the reference tree):
 *
 *   ertm_step            AcousticPropagator.step      ertm/propagation.py:430-454
 *                        ElasticPropagator.step       ertm/propagation.py:840-883
 *                        ViscoacousticPropagator.step ertm/propagation.py:686-723
 *                        (with CPML2D.derivative      ertm/numerics.py:283-308)
 *   ertm_stabilize       ViscoacousticPropagator._stabilize ertm/propagation.py:726-749
 *   ertm_sample          Propagator2D._sample_field + sample()  ertm/propagation.py:371-391,
 *                        :456-470, :751-762, :885-898 (bilinear_sample ertm/numerics.py:478-511)
 *   ertm_inject          inject()/_add_at_nodes/_spread_into  ertm/propagation.py:331-369,
 *                        :472-494, :764-774, :900-942 (np.add.at order kept)
 *   ertm_image           image_fields() + the imaging-condition accumulation of
 *                        RTMImager.migrate            ertm/propagation.py:496-497, :945-977;
 *                        ertm/imaging.py:263-331, :450-474
 *   ertm_copy_state      Propagator2D.state/load_state ertm/propagation.py:283-299
 *                        (checkpoint save/restore, ertm/imaging.py:113-162)
 *   ertm_run             the time loops of record_source_wavefield / replay_block /
 *                        the backward loop of migrate  ertm/imaging.py:113-162, :279-324
 *   ertm_derivative      dplus/dminus                  ertm/numerics.py:68-137
 *   ertm_absmax          max_amplitude                 ertm/propagation.py:503-505
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero ERTM_E* code on error;
 *     ertm_last_error() returns a thread-local message.  No C++ exception
 *     crosses the ABI.
 *   - All device memory named in the descriptors is allocated by the caller
 *     (PyTorch tensors on the Python side); the library only keeps small
 *     descriptor tables of its own.
 *   - Work is enqueued asynchronously on the given HIP stream (NULL = the
 *     default stream).  Nothing synchronises unless stated.
 *   - Arrays on the padded grid are row-major with row stride `ld` floats
 *     (x fastest); physical-grid arrays (images, accumulators) are dense
 *     (nz, nx).
 */
#ifndef ERTM_HIP_H
#define ERTM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ERTM_ABI_VERSION 6

enum { ERTM_OK = 0, ERTM_EINVAL = 1, ERTM_EHIP = 2, ERTM_ENOMEM = 3, ERTM_ESTATE = 4 };

enum { ERTM_ACOUSTIC = 0, ERTM_ELASTIC = 1, ERTM_VISCO = 2 };
enum { ERTM_SRC_EXPLOSIVE = 0, ERTM_SRC_SHEAR = 1, ERTM_SRC_FORCE_Z = 2 };

/* Field slab indices inside a state buffer. */
enum { ERTM_AC_P = 0, ERTM_AC_VX = 1, ERTM_AC_VZ = 2 };
enum { ERTM_EL_VX = 0, ERTM_EL_VZ = 1, ERTM_EL_SXX = 2, ERTM_EL_SZZ = 3, ERTM_EL_SXZ = 4 };
enum { ERTM_VI_P = 0, ERTM_VI_VX = 1, ERTM_VI_VZ = 2, ERTM_VI_EPS = 3, ERTM_VI_R0 = 4 };

#define ERTM_MAX_MECH 8

/*
 * One propagator: geometry, medium, CPML profiles, the two ping-pong state
 * buffers and the point source.  A step reads state[cur] and writes
 * state[cur ^ 1].
 *
 * State buffer layout (floats): `nfields` slabs of `fstride` floats, each an
 * (nzp, ld) array; then the CPML memory arrays at the offsets psix/psiz.
 * x-direction memories are (nzp, 2*wx) arrays covering columns [0, wx) and
 * [nxp-wx, nxp); z-direction memories are (wzt+wzb, ld) arrays covering rows
 * [0, wzt) and [nzp-wzb, nzp).  Outside those strips the reference's
 * recursion is an exact no-op (a = 0, b = 1), so nothing is stored there.
 *
 * Medium (padded grid, stride ld):
 *   acoustic  kappa = rho*vp*vp (f32 ops), irho = 1/rho
 *   elastic   lam, mu, irho
 *   visco     kappa = K_R (relaxed modulus), irho, qidx (Q class per cell),
 *             q_a[nq*L] = f32 exp(-dt/tau), q_delta/q_tau[nq*L] (f64)
 * Face averages (1/rho on x/z faces, mu on cell corners, lam+2mu) are
 * recomputed in the kernels with the reference's float32 operation order.
 */
typedef struct ErtmProp {
    int32_t engine;        /* ERTM_ACOUSTIC / ERTM_ELASTIC / ERTM_VISCO */
    int32_t nzp, nxp, ld;  /* padded shape and row stride (floats) */
    int32_t pad, nz, nx;   /* ghost width and physical shape */
    int32_t order;         /* 2, 4, 6 or 8 */
    int32_t free_surface;  /* 1: zero p (acoustic) or szz,sxz (elastic) on row `pad` */
    int32_t nmech;         /* viscoacoustic mechanisms L (<= ERTM_MAX_MECH) */
    int32_t wx, wzt, wzb;  /* CPML strip widths */
    int32_t nfields;
    float dx, dz, dt;      /* float32 values of the spacing and time step */
    float coef[4];         /* float32 stencil weights, zero beyond order/2 */
    double dt64;           /* the time step as given (float64), for the GSLS b_l */
    double rdx, rdz;       /* float64 reciprocals of the float32 dx, dz */
    int64_t fstride;       /* floats between field slabs */
    int64_t psix[4];       /* offsets of the x-memories (see ertm_step) */
    int64_t psiz[4];       /* offsets of the z-memories */
    int64_t state_floats;  /* floats in one state buffer */
    const float* kappa;
    const float* irho;
    const float* lam;
    const float* mu;
    const int32_t* qidx;
    const float* q_a;
    const double* q_delta;
    const double* q_tau;
    const float* ax;       /* CPML a, b per padded column (nxp) */
    const float* bx;
    const float* az;       /* CPML a, b per padded row (nzp) */
    const float* bz;
    float* state[2];
    int32_t src_iz, src_ix; /* padded source node, -1 = none */
    int32_t mechanism;      /* ERTM_SRC_* (elastic) */
    int32_t _pad0;
    /* optional medium palette (elastic and acoustic, npal > 0): one byte per
     * padded cell, row r at midx + r*ld, indexing the float32 table mtab:
     * elastic 1/rho, mu, lambda at [0, 256), [256, 512), [512, 768);
     * acoustic 1/rho, kappa at [0, 256), [256, 512).  Index 0 is the zero
     * medium (padding columns, and what rows outside the grid read), 1..npal
     * the distinct tuples of the model.  The single-launch step
     * kernels then read one byte of medium per cell instead of three floats;
     * the values, and so the results, are the same.  npal = 0: no palette. */
    const uint8_t* midx;
    const float* mtab;
    int32_t npal;
    int32_t _pad1;
} ErtmProp;

/*
 * Receiver tables of one propagator (built on the host in float64 exactly as
 * the reference computes fractional indices).
 *
 * Sampling: nsamp entries; entry e reads padded corner (i0[e], j0[e]) of
 * field `sfield[e]` (-1: acoustic "sxz", always 0; -2: elastic pressure
 * 0.5*(sxx+szz)) with weights tz[e], tx[e] (valid[e]==0 -> 0) and writes
 * record `srec[e]` at column `scol[e]`.
 *
 * Injection: for each target cell u (ncells), entries [cell_ptr[u],
 * cell_ptr[u+1]) are applied in order to field slab `cell_field[u]` at flat
 * padded offset `cell_off[u]`:  value = record[ecomp[e]][t][erec[e]] (f32),
 *   mode 0: c = (value * wa[e]) * wb[e]   (acoustic / elastic scatter)
 *   mode 1: c = value * wa[e]             (viscoacoustic _add_at_nodes)
 *   field = (float)((double)field + c)
 * which is np.add.at's order and rounding.
 */
typedef struct ErtmRecv {
    int32_t nsamp;
    int32_t ncells;
    int32_t nentries;
    int32_t inject_mode;
    int32_t nrec;
    int32_t ncomp;
    const int32_t* s_i0;
    const int32_t* s_j0;
    const int32_t* s_field;
    const int32_t* s_rec;
    const int32_t* s_col;
    const int32_t* s_valid;
    const double* s_tz;
    const double* s_tx;
    const int32_t* cell_ptr;
    const int32_t* cell_field;
    const int64_t* cell_off;
    const int32_t* e_comp;
    const int32_t* e_rec;
    const double* e_wa;
    const double* e_wb;
    float* records[8];       /* per component: (nt, nrec) float32 */
} ErtmRecv;

/* Image-field selectors. */
enum { ERTM_WAVE_P = 0, ERTM_WAVE_S = 1 };
enum { ERTM_SEP_STRESS = 0, ERTM_SEP_DIVCURL = 1 };
enum {
    ERTM_IMG_EXTRACT = 0,      /* write the source field into rbuf[t - t0] */
    ERTM_IMG_CORRELATE = 1,    /* accumulate zero_lag / normalized terms */
    ERTM_IMG_EXC_PEAK = 2,     /* excitation: update running/peak maps */
    ERTM_IMG_EXC_APPLY = 3     /* excitation: masked accumulation */
};

/*
 * Imaging buffers of one propagator (physical grid, dense nz*nx).
 *   rbuf      replay slots (K, nz, nx) float32, source-side field per step
 *   num[2]    cross-correlation numerators (float64), one per receiver wave
 *   ill       source illumination (float64); rill[2] receiver illumination
 *   peak_step int32, peak_amp / running float64 (excitation)
 */
/* ErtmImage.illum bits.  The reference accumulates both illuminations
 * whenever either is used (ertm/imaging.py:300-306); each is only read by its
 * own normalisation (:326-331), so only the ones used are accumulated. */
#define ERTM_ILLUM_SRC 1
#define ERTM_ILLUM_REC 2

typedef struct ErtmImage {
    int32_t separation;    /* ERTM_SEP_* (elastic) */
    int32_t swave;         /* source-side wave (ERTM_WAVE_*) */
    int32_t nrw;           /* receiver waves accumulated (1 or 2) */
    int32_t rwave[2];
    int32_t start_step;    /* cross term only for t >= start_step */
    int32_t illum;         /* ERTM_ILLUM_* bits: accumulate ill (source) and/or
                              rill (receiver) illumination every step */
    int32_t t0;            /* time index of replay slot 0 */
    int32_t nslots;        /* K */
    int32_t ring_len;      /* R: steps of receiver fields in rring (0 = none) */
    float* rbuf;
    double* num[2];
    double* ill;
    double* rill[2];
    int32_t* peak_step;
    double* peak_amp;
    double* running;
    /* optional (R, nrw, nz, ld) float32 ring, ld = the propagator's row
     * pitch (ErtmProp.ld), physical cell (i, j) at column j + pad: with it, a
     * fused batch stores the receiver fields of up to R backward steps and
     * folds them into the accumulators in one pass (same per-cell order:
     * identical results).  Rows on the state's pitch let each block write
     * the same cache-line spans as its state stores. */
    float* rring;
} ErtmImage;

typedef struct ertm_batch_s* ertm_batch_t;

/* ---- library ---------------------------------------------------------- */
int ertm_abi_version(void);
/* sizeof(ErtmProp), sizeof(ErtmRecv), sizeof(ErtmImage), sizeof(ErtmRun),
 * offsetof(ErtmProp, state), offsetof(ErtmRun, ctl), sizeof(ErtmSeg),
 * offsetof(ErtmSeg, wavelet): lets a binding check its structure layout
 * without a GPU. */
int ertm_struct_sizes(int64_t* out8);
const char* ertm_last_error(void);
int ertm_device_info(int32_t* device, int32_t* n_cu, char* name, size_t name_len);

/* ---- batches: n propagators advanced together --------------------------- */
/* Copies the descriptors to device memory.  recv/img may be NULL (or an
 * entry may have nsamp=ncells=0 / all-NULL buffers). */
int ertm_batch_create(const ErtmProp* props, const ErtmRecv* recv, const ErtmImage* img,
                      int32_t n, ertm_batch_t* out);
/* Replace the imaging descriptors.  Waits for the work already enqueued with
 * this batch (on any stream), then uploads on `stream` and waits for it. */
int ertm_batch_update_image(ertm_batch_t b, const ErtmImage* img, void* stream);
/* Releases the batch.  Its device tables return to the library's pool only
 * after the work already enqueued with it has run (events on its streams);
 * the call waits for that work only when the batch holds captured graphs or
 * FFT plans.  The caller's buffers named in the descriptors stay the
 * caller's to keep alive until that work has run. */
int ertm_batch_destroy(ertm_batch_t b);
/* Stencil kernel variant: 0 = two launches per step (velocity, then
 * pressure/stress; every engine and order); 1, 2, 3, 4 = single fused launch
 * per step (batches of one engine at order 8: elastic, acoustic, or
 * viscoacoustic with 1..4 mechanisms) with 256, 128, 64, 192 threads per
 * block streaming
 * chunks of zc rows (8 <= zc <= 4096), alternate chunks in opposite
 * directions.  Results are bit-identical across variants. */
int ertm_batch_set_variant(ertm_batch_t b, int32_t variant, int32_t zc);
/* Two time steps per launch in ertm_run's forward and replay passes
 * (elastic batches at order 8): variant 2 = 128 threads per block streaming
 * chunks of zc rows; 0 = off (the default: on MI355X it measured slower than
 * one step per launch, 756 against 619 us per step on 64 shots of 600 x
 * 1200, being VALU-bound at 3 waves per SIMD).  The state after
 * the first step of a pair is not stored; a pair never spans a checkpoint
 * save, receiver sampling or the stabiliser, which keep one step per launch.
 * The library allocates a scratch for the CPML memories between the two
 * steps (the CPML part of a state buffer per propagator) on first use.
 * Results are bit-identical to one step per launch. */
int ertm_batch_set_two_step(ertm_batch_t b, int32_t variant, int32_t zc);
/* Folds of the receiver-field ring beside the backward steps: no counterpart
 * in the reference (its migrate correlates every step, ertm/imaging.py:
 * 279-324).  With period > 0, ertm_run's eager backward passes with the ring
 * and without the stabiliser fold every `period` steps on a stream of the
 * library's own, while the next steps run on the caller's stream, when every
 * image's ring_len holds at least 2 * period steps (else they fold every
 * ring_len steps on the caller's stream, as with period = 0).  Events order
 * the slots' reuse and the folds themselves, and the caller's stream waits
 * for the last fold before ertm_run returns: the accumulators get the same
 * additions in the same order.  Creates the stream on first use. */
int ertm_batch_set_ring_flush(ertm_batch_t b, int32_t period);

/*
 * Per-launch time table: propagator i of the batch is at time index t[i] and
 * its point source adds amp[i] (0 = no source).  At most ERTM_MAX_LAUNCH
 * propagators are advanced per launch; larger batches are split internally.
 */
#define ERTM_MAX_LAUNCH 64
typedef struct ErtmTimes {
    int32_t t[ERTM_MAX_LAUNCH];
    float amp[ERTM_MAX_LAUNCH];
} ErtmTimes;

/*
 * One time step for every propagator of the batch: state[cur] -> state[cur^1].
 * amp (host, n entries, NULL = all 0) is added exactly as the reference adds
 * float32(amp) at the source node.
 */
int ertm_step(ertm_batch_t b, int32_t cur, const float* amp, void* stream);

/* Viscoacoustic compensation stabiliser after a step (state[cur] is the
 * post-step state).  step_index = the propagator's step counter before
 * increment; ctl (device, 4 doubles per propagator): [ref_amp, clipped,
 * lowpass_steps, unused]; cap = amplitude_cap; interval = lowpass_interval. */
int ertm_stabilize(ertm_batch_t b, int32_t cur, int32_t step_index, double cap,
                   int32_t interval, double* ctl, void* stream);

/* Sample every receiver of every propagator from state[cur] into record row
 * t[i] (host array, n entries). */
int ertm_sample(ertm_batch_t b, int32_t cur, const int32_t* t, void* stream);
/* Inject record row t[i] into state[cur] (before the backward step). */
int ertm_inject(ertm_batch_t b, int32_t cur, const int32_t* t, void* stream);
/* Imaging operation `op` (ERTM_IMG_*) at time index t[i] on state[cur]. */
int ertm_image(ertm_batch_t b, int32_t cur, const int32_t* t, int32_t op, void* stream);
/* Copy a whole state buffer (checkpoint save: dst = arena slot, src =
 * state[cur]; restore: the reverse).  Plain device-to-device copy. */
int ertm_copy_state(float* dst, const float* src, int64_t nfloats, void* stream);

/*
 * Time loops (one HIP launch sequence per step, no host synchronisation):
 *   phase 0 (forward):   for t in [t_begin, t_end): step(amp = w[t]);
 *                        sample(t) if do_sample; image(op_after, t) if >= 0;
 *                        checkpoint (below).
 *   phase 1 (replay):    first copy restore_arenas[i] + restore_slot *
 *                        restore_stride into state[cur0], then as phase 0.
 *   phase 2 (backward):  for t = t_end-1 .. t_begin: inject(t); step(0);
 *                        image(op_after, t).
 * Checkpoints (phases 0, 1): with ck_interval > 0, after step t with
 * t >= ck_origin and (t - ck_origin) % ck_interval == 0 the new state is
 * copied to save_arenas[i] + ((t - ck_origin) / ck_interval) * save_stride.
 * With stabilize = 1 (viscoacoustic compensation) ertm_stabilize runs after
 * every step with the step counter starting at step_index0.
 * `wavelets` is a host array of n pointers to host float32 wavelets of length
 * nw (unused in phase 2).  The parity of the final state goes to *cur_out.
 */
typedef struct ErtmRun {
    int32_t phase;
    int32_t t_begin, t_end;
    int32_t cur0;
    int32_t do_sample;
    int32_t op_after;          /* -1 none, else ERTM_IMG_* */
    int32_t ck_interval;       /* 0 = no checkpoints */
    int32_t ck_origin;
    int64_t save_stride;       /* floats between saved slots */
    float* const* save_arenas;
    int32_t restore_slot;
    int32_t nw;                /* wavelet length */
    int64_t restore_stride;
    float* const* restore_arenas;
    const float* const* wavelets;
    int32_t stabilize;         /* viscoacoustic compensation stabiliser */
    int32_t step_index0;
    int32_t lp_interval;
    int32_t graph_steps;       /* > 0: capture chunks of this many steps as one
                                  hipGraph and replay it (even; a multiple of the
                                  ring length when the ring images; not with the
                                  stabiliser); the rest of the range runs eagerly */
    double cap;
    double* ctl;               /* device, 4 doubles per propagator */
} ErtmRun;
int ertm_run(ertm_batch_t b, const ErtmRun* run, int32_t* cur_out, void* stream);

/*
 * Lanes: the time loops of several independent migrations, each on its own
 * schedule, advanced in lockstep.  No counterpart in the reference, which
 * migrates one shot per call (ertm/imaging.py:205-349); it is what lets a
 * server batch shots whose nt, checkpoint interval, replay segments and
 * phase all differ.
 *
 * A segment is one ertm_run-like pass of one propagator.  Segments with the
 * same `lane` run one after another in list order; the lanes run side by
 * side: each tick advances the current segment of every lane by one step,
 * and the step launches of a tick cover every lane in the same kind of pass
 * (forward, extracting replay, backward), so shots at different times,
 * phases and checkpoint blocks share one launch.  A propagator belongs to
 * one lane.  Per step a segment does what ertm_run does for its phase:
 *   phase 0: step(amp = wavelet[t]); image(op_after, t); checkpoint
 *   phase 1: as phase 0, after copying `restore` into the state before the
 *            segment's first step
 *   phase 2: inject(t); step(0); image(op_after, t), for t = t_end-1 down
 * op_after: -1, ERTM_IMG_EXTRACT (phases 0, 1) or ERTM_IMG_CORRELATE (phase
 * 2), with replay slot t - img_t0 of the propagator's image buffers (slots
 * [t_begin - img_t0, t_end - img_t0) must exist).  Checkpoints (phases 0, 1):
 * after step t with t >= ck_origin and (t - ck_origin) % ck_interval == 0 the
 * state goes to save_arena + ((t - ck_origin) / ck_interval) * save_stride.
 * cur (n entries, in/out) holds each propagator's state parity.  `done`
 * (in/out) counts the steps of a segment already run, so a call can resume
 * where an earlier one stopped.  The call returns when every segment is
 * done, after max_ticks ticks (> 0), or, with stop_at_lane_end, after the
 * tick in which a lane ran out of segments.  stats3 (optional): ticks run,
 * lane steps (one per lane per tick) and step-launch groups (the launches
 * of one pass kind in one tick; each covers up to ERTM_MAX_LAUNCH lanes per
 * HIP launch), so lane steps / groups = lanes sharing a launch.  Each
 * propagator's results equal those of the same passes run by ertm_run.
 * With a fused variant set on the batch, each launch picks its own chunk
 * length from the blocks its lanes' grids give (long chunks when the chip
 * is full, short ones when few blocks would leave each launch waiting on one
 * chunk's chain of rows).  No capture, stabiliser, sampling or two-step
 * launches.
 */
typedef struct ErtmSeg {
    int32_t lane;
    int32_t prop;              /* propagator of the batch */
    int32_t phase;             /* 0 forward, 1 replay, 2 backward */
    int32_t t_begin, t_end;
    int32_t op_after;          /* -1, ERTM_IMG_EXTRACT or ERTM_IMG_CORRELATE */
    int32_t img_t0;            /* time of replay slot 0 */
    int32_t ck_interval;       /* 0 = no checkpoints */
    int32_t ck_origin;
    int32_t done;              /* in/out: steps already run */
    int64_t save_stride;       /* floats between saved slots */
    float* save_arena;
    const float* restore;      /* phase 1: state to start from */
    const float* wavelet;      /* phases 0, 1: host amplitudes, >= t_end entries */
} ErtmSeg;
int ertm_run_lanes(ertm_batch_t b, ErtmSeg* segs, int32_t nseg, int32_t* cur, int32_t max_ticks,
                   int32_t stop_at_lane_end, int64_t* stats3, void* stream);

/* ---- utilities ----------------------------------------------------------- */
/* Staggered derivative of a dense (n0, n1) array (dplus/dminus): axis 0 or
 * 1, plus = 1 for the value at i+1/2; is_f64 selects float64 data (else
 * float32, where h is rounded to float32 first, as NumPy does). */
int ertm_derivative(const void* f, void* out, int32_t n0, int32_t n1, int32_t axis,
                    int32_t plus, int32_t order, double h, int32_t is_f64, void* stream);
/* max |x| over n floats into *out (device float). */
int ertm_absmax(const float* x, int64_t n, float* out, void* stream);
/* Block until all work on the stream is done. */
int ertm_sync(void* stream);

/* ---- image post-processing --------------------------------------------------
 * RTMImager normalisation and _postprocess (ertm/imaging.py:326-331, :476-502:
 * laplacian ertm/numerics.py:398-404, lowpass_wavenumber :424-442) on the
 * device: image = num [/ (ill + eps)] [/ (rill + eps)], 5-point Laplacian
 * with edge padding, radial wavenumber low-pass (hipFFT, float64), zero
 * npml cells at the left/right/bottom (and top if mask_top), cast to float32.
 * work: caller-owned device scratch of 4*nz*nx doubles. */
typedef struct ErtmPost {
    int32_t nz, nx;
    int32_t div_ill, div_rill;   /* normalise by ill / rill */
    int32_t laplacian, lowpass;  /* apply the Laplacian / the low-pass */
    int32_t npml, mask_top;
    double dx, dz, eps, fraction;
    const double* num;
    const double* ill;
    const double* rill;
    double* work;
    float* out;                  /* (nz, nx) float32 image */
} ErtmPost;
int ertm_postprocess(const ErtmPost* p, void* stream);

/* Allocation-free variant of ertm_postprocess.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  p->work is ignored (may be NULL);
 * the scratch is a static device array of the library holding
 * 2 * ERTM_NOALLOC_MAX_CELLS doubles (32 MiB) on every GPU the library runs
 * on, for the life of the process.  Limits: nz * nx <= ERTM_NOALLOC_MAX_CELLS
 * and lowpass = 0 (the wavenumber low-pass needs a hipFFT plan, which
 * allocates); either violation returns ERTM_EINVAL before any GPU work.
 * Calls no hipMalloc, hipFree or allocating library routine.  All calls on
 * one GPU share that GPU's scratch: calls that overlap on one GPU (on
 * different streams, or from several threads without ordering) race.  Same
 * kernels and results as ertm_postprocess. */
#define ERTM_NOALLOC_MAX_CELLS (1 << 21)
int ertm_postprocess_noalloc(const ErtmPost* p, void* stream);

/* ---- record preparation ------------------------------------------------------
 * On a device record (nt, nrec) float32, row-major: the direct-wave top mute
 * (mute_direct_wave, ertm/preprocess.py:121-169: weight 0.5 - 0.5 cos(pi r),
 * r rising linearly over t_mute -+ taper; t_mute NULL = no mute), then the
 * end taper (taper_record_ends, :172-189; w_end = 0: none).  t: nt float64
 * sample times, t_mute: nrec float64.  Float64 arithmetic, float32 after each
 * step, as the reference. */
int ertm_prepare_record(float* rec, int32_t nt, int32_t nrec, const double* t,
                        const double* t_mute, double taper, int32_t w_end, void* stream);
/* (vx, vz) rotation by an azimuth, cos/sin given (rotate_components,
 * ertm/preprocess.py:36-63), in place on n floats each. */
int ertm_rotate_components(float* vx, float* vz, int64_t n, double cos_a, double sin_a,
                           void* stream);
/* Kaiser-windowed sinc resampling of a record onto the solver's dt
 * (resample_sinc, ertm/numerics.py:337-392) fused with the length fit of
 * RTMImager._prepare_record (ertm/imaging.py:396-423).  in: (n_in, nrec)
 * float32 device record; w, off: (n_out, taps) float64 weights (normalised as
 * the reference normalises them) and int32 tap rows, on the device, computed
 * on the host; out: (out_rows, nrec) float32 device record.  Row k < n_out is
 * the float64 sum over the taps in order of w * in[clip(off, 0, n_in - 1)],
 * rounded to float32; later rows are zero. */
int ertm_resample_record(const float* in, int32_t n_in, int32_t nrec, const double* w,
                         const int32_t* off, int32_t n_out, int32_t taps, float* out,
                         int32_t out_rows, void* stream);

/* ---- measurement (no counterpart in the reference) ------------------------ */
/* Launch timing.  While enabled, every launch that ertm_step / ertm_run /
 * ertm_sample / ertm_inject / ertm_image make outside a captured time loop
 * (one launch per group of up to ERTM_MAX_LAUNCH propagators) is bracketed
 * by two HIP events recorded on its stream and tagged with its kind.
 * Enabling clears earlier records. */
enum {
    ERTM_TK_TWO_LAUNCH = 0,  /* k_velocity + k_stress (one pair = one launch) */
    ERTM_TK_PLAIN = 1,       /* fused step, no imaging */
    ERTM_TK_EXTRACT = 2,     /* fused step + source-field extraction (replay) */
    ERTM_TK_RING = 3,        /* fused step + receiver-field ring store (backward) */
    ERTM_TK_CORRELATE = 4,   /* fused step + correlation */
    ERTM_TK_FLUSH = 5,       /* ring flush into the accumulators */
    ERTM_TK_INJECT = 6,      /* receiver injection */
    ERTM_TK_COPY = 7,        /* checkpoint save / restore (k_copy_batch) */
    ERTM_TK_SAMPLE = 8,      /* receiver sampling */
    ERTM_TK_IMAGE = 9,       /* separate imaging pass */
    ERTM_TK_PLAIN2 = 10,     /* two-step launch, no imaging (cells: 2 per cell) */
    ERTM_TK_EXTRACT2 = 11,   /* two-step launch + extraction of both steps */
    ERTM_TK_COUNT = 12
};
int ertm_step_timing(int32_t enable);
/* Waits for the recorded events; per kind k (arrays of ERTM_TK_COUNT): summed
 * milliseconds, launches and physical cells covered (nz*nx per propagator;
 * for a flush, times the steps it folds). */
int ertm_timing_read(double* ms, int64_t* launches, int64_t* cells);
/* The step launches without imaging (ERTM_TK_PLAIN + ERTM_TK_TWO_LAUNCH):
 * summed ms, launches, cell-updates, and whether they were the fused kernel
 * (1), the two-launch pair (0), both (2) or none (-1). */
int ertm_step_timing_read(double* total_ms, int64_t* launches, int64_t* cell_updates,
                          int32_t* fused);

#ifdef __cplusplus
}
#endif
#endif /* ERTM_HIP_H */
