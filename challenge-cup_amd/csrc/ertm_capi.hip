// C-ABI of libertm_hip.so (declared in include/ertm_hip.h): validation,
// descriptor tables, launch chunking and the time loops.  No C++ exception
// crosses this boundary; every entry point returns an ERTM_* status and
// leaves a thread-local message for ertm_last_error().
#include <hipfft/hipfft.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <mutex>
#include <vector>

#include "ertm_kernels.h"

using namespace ertm;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(ERTM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

#define LAUNCHCHK(what)                                                               \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess)                                                         \
            return fail(ERTM_EHIP, "%s launch failed: %s", what, hipGetErrorString(e_)); \
    } while (0)

// Batched real-to-complex plans and buffers of the stabiliser's low-pass,
// one per (nzp, nxp, members of the group).
struct FftCache {
    hipfftHandle fwd = 0, inv = 0;
    double* real = nullptr;
    void* cplx = nullptr;
};

// Descriptor tables come from a process-wide pool of device blocks that are
// never freed: hipFree waits for the whole device, which would couple batches
// running on different streams.  Blocks are reused by size class (powers of
// two from 4 KiB).
std::mutex g_pool_m;
std::multimap<size_t, void*> g_pool;

size_t pool_class(size_t bytes) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

// Blocks released while launches that read them may still be queued: each
// returns to the pool once the events recorded after those launches (one per
// stream that used it) have completed.
struct Deferred {
    void* p;
    size_t cls;
    std::vector<hipEvent_t> ev;
};
std::vector<Deferred> g_deferred;

// Move deferred blocks whose events have completed into the pool (lock
// held).  The blocks still waiting move to a new list: compacting in place
// would move an entry onto itself, and a self-moved std::vector is left
// empty, which would free its block with its events forgotten.
void pool_reap() {
    std::vector<Deferred> waiting;
    for (Deferred& d : g_deferred) {
        bool done = true;
        for (hipEvent_t e : d.ev) done = done && hipEventQuery(e) == hipSuccess;
        if (done) {
            for (hipEvent_t e : d.ev) (void)hipEventDestroy(e);
            g_pool.emplace(d.cls, d.p);
        } else {
            waiting.push_back(std::move(d));
        }
    }
    g_deferred.swap(waiting);
}

void* pool_get(size_t bytes) {
    const size_t c = pool_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_pool_m);
        pool_reap();
        auto it = g_pool.find(c);
        if (it != g_pool.end()) {
            void* p = it->second;
            g_pool.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) return nullptr;
    return p;
}

void pool_put(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_m);
    g_pool.emplace(pool_class(bytes), p);
}

// pool_put once the work now queued on `streams` has run.  Returns false
// (and releases nothing) if an event cannot be recorded.
bool pool_put_after(void* p, size_t bytes, const std::vector<hipStream_t>& streams) {
    if (!p) return true;
    Deferred d{p, pool_class(bytes), {}};
    for (hipStream_t s : streams) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            for (hipEvent_t x : d.ev) (void)hipEventDestroy(x);
            return false;
        }
        d.ev.push_back(e);
        if (hipEventRecord(e, s) != hipSuccess) {
            for (hipEvent_t x : d.ev) (void)hipEventDestroy(x);
            return false;
        }
    }
    std::lock_guard<std::mutex> lk(g_pool_m);
    g_deferred.push_back(std::move(d));
    return true;
}

}  // namespace

struct ertm_batch_s {
    std::vector<ErtmProp> props;
    std::vector<ErtmRecv> recv;
    std::vector<ErtmImage> img;
    ErtmProp* dprops = nullptr;
    ErtmRecv* drecv = nullptr;
    ErtmImage* dimg = nullptr;
    int order = 8;
    int max_tiles = 0, max_samp = 0, max_cells = 0, max_img = 0;
    bool any_force = false;
    bool all_visco = false;
    bool all_elastic = false;
    bool all_acoustic = false;
    int variant = 0;
    int zc = 64;
    // two-step launches (ertm_batch_set_two_step): launch shape, 0 = off; and
    // the scratch for the CPML memories after the first step, one CPML part
    // of a state buffer per propagator
    int v2 = 0, zc2 = 0;
    void* psi1 = nullptr;
    size_t psi1_bytes = 0;
    std::vector<float*> psi1_ptr;
    float* scratch = nullptr;  // per propagator: peak, scale
    // per receiver-ring slot and propagator: the stabiliser's factor for the
    // ring entry (compensated viscoacoustic backward passes with the ring)
    float* rscale = nullptr;
    size_t rscale_bytes = 0;
    void* block = nullptr;     // pool block holding dprops, drecv, dimg, scratch
    size_t block_bytes = 0;
    std::vector<hipStream_t> streams;  // streams this batch has enqueued work on
    std::map<std::tuple<int, int, int>, FftCache> fft;
    // captured time loops (ertm_run with graph_steps > 0): device clock
    // {step, t0, dir}, amplitude table, checkpoint arena pointers, and the
    // instantiated graphs keyed by everything their launches bake in
    int32_t* dclk = nullptr;
    float* damp = nullptr;
    size_t damp_bytes = 0;
    float** dark = nullptr;
    std::map<std::vector<int64_t>, hipGraphExec_t> graphs;
    hipStream_t cap = nullptr;  // private stream the loops are captured on
    // ring folds beside the backward steps (ertm_batch_set_ring_flush): every
    // flush_period steps, on a private stream; fl_done[w % 2] marks fold w
    int flush_period = 0;
    hipStream_t fls = nullptr;
    hipEvent_t fl_step = nullptr, fl_done[2] = {nullptr, nullptr};
    int fl_count = 0;
    void note(hipStream_t s) {
        for (hipStream_t t : streams)
            if (t == s) return;
        streams.push_back(s);
    }
    ~ertm_batch_s() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (cap) (void)hipStreamDestroy(cap);
        if (fl_step) (void)hipEventDestroy(fl_step);
        for (hipEvent_t e : fl_done)
            if (e) (void)hipEventDestroy(e);
        if (fls) (void)hipStreamDestroy(fls);
        pool_put(dclk, 64);
        pool_put(rscale, rscale_bytes);
        pool_put(damp, damp_bytes);
        pool_put(dark, sizeof(float*) * props.size());
        pool_put(block, block_bytes);
        pool_put(psi1, psi1_bytes);
        for (auto& kv : fft) {
            if (kv.second.fwd) hipfftDestroy(kv.second.fwd);
            if (kv.second.inv) hipfftDestroy(kv.second.inv);
            if (kv.second.real) (void)hipFree(kv.second.real);
            if (kv.second.cplx) (void)hipFree(kv.second.cplx);
        }
    }
};

static int check_prop(const ErtmProp& P, int k) {
    if (P.engine < ERTM_ACOUSTIC || P.engine > ERTM_VISCO)
        return fail(ERTM_EINVAL, "prop %d: unknown engine %d", k, P.engine);
    if (P.order != 2 && P.order != 4 && P.order != 6 && P.order != 8)
        return fail(ERTM_EINVAL, "prop %d: order %d not in {2,4,6,8}", k, P.order);
    const int K = P.order / 2;
    if (P.nzp <= 2 * K || P.nxp <= 2 * K)
        return fail(ERTM_EINVAL, "prop %d: padded grid %dx%d too small for order %d", k, P.nzp,
                    P.nxp, P.order);
    if (P.ld < P.nxp) return fail(ERTM_EINVAL, "prop %d: ld %d < nxp %d", k, P.ld, P.nxp);
    if (P.pad < K || P.nz + 2 * P.pad != P.nzp || P.nx + 2 * P.pad != P.nxp)
        return fail(ERTM_EINVAL, "prop %d: inconsistent pad/nz/nx", k);
    if (P.fstride < (int64_t)P.nzp * P.ld) return fail(ERTM_EINVAL, "prop %d: fstride too small", k);
    if (P.wx < 0 || 2 * P.wx > P.nxp || P.wzt < 0 || P.wzb < 0 || P.wzt + P.wzb > P.nzp)
        return fail(ERTM_EINVAL, "prop %d: bad CPML widths", k);
    if (!P.state[0] || !P.state[1] || P.state[0] == P.state[1])
        return fail(ERTM_EINVAL, "prop %d: need two distinct state buffers", k);
    if (!P.irho || !P.ax || !P.bx || !P.az || !P.bz)
        return fail(ERTM_EINVAL, "prop %d: missing medium or CPML arrays", k);
    if (P.engine == ERTM_ELASTIC && (!P.lam || !P.mu))
        return fail(ERTM_EINVAL, "prop %d: elastic needs lam and mu", k);
    if (P.engine != ERTM_ELASTIC && !P.kappa)
        return fail(ERTM_EINVAL, "prop %d: needs kappa", k);
    if (P.engine == ERTM_VISCO &&
        (P.nmech < 1 || P.nmech > ERTM_MAX_MECH || !P.qidx || !P.q_a || !P.q_delta || !P.q_tau))
        return fail(ERTM_EINVAL, "prop %d: viscoacoustic tables missing or L out of range", k);
    const int nf = P.engine == ERTM_ELASTIC ? 5 : (P.engine == ERTM_ACOUSTIC ? 3 : 4 + P.nmech);
    if (P.nfields != nf) return fail(ERTM_EINVAL, "prop %d: nfields %d != %d", k, P.nfields, nf);
    const int nterms = P.engine == ERTM_ELASTIC ? 4 : 2;
    int64_t end = (int64_t)nf * P.fstride;
    for (int t = 0; t < nterms; ++t) {
        if (P.wx > 0) {
            if (P.psix[t] < end) return fail(ERTM_EINVAL, "prop %d: psix[%d] overlaps", k, t);
            end = P.psix[t] + (int64_t)P.nzp * 2 * P.wx;
        }
    }
    for (int t = 0; t < nterms; ++t) {
        if (P.wzt + P.wzb > 0) {
            if (P.psiz[t] < end) return fail(ERTM_EINVAL, "prop %d: psiz[%d] overlaps", k, t);
            end = P.psiz[t] + (int64_t)(P.wzt + P.wzb) * P.ld;
        }
    }
    if (end > P.state_floats) return fail(ERTM_EINVAL, "prop %d: state buffer too small", k);
    if (P.src_iz >= P.nzp || P.src_ix >= P.nxp || (P.src_iz < 0) != (P.src_ix < 0))
        return fail(ERTM_EINVAL, "prop %d: source node out of range", k);
    return ERTM_OK;
}

static int check_recv(const ErtmRecv& R, const ErtmProp& P, int k) {
    if (R.nsamp < 0 || R.ncells < 0 || R.ncomp < 0 || R.ncomp > 8)
        return fail(ERTM_EINVAL, "recv %d: bad counts", k);
    if (R.nsamp > 0 && (!R.s_i0 || !R.s_j0 || !R.s_field || !R.s_rec || !R.s_col || !R.s_valid ||
                        !R.s_tz || !R.s_tx))
        return fail(ERTM_EINVAL, "recv %d: missing sampling tables", k);
    if (R.ncells > 0 && (!R.cell_ptr || !R.cell_field || !R.cell_off || !R.e_comp || !R.e_rec ||
                         !R.e_wa || (R.inject_mode == 0 && !R.e_wb)))
        return fail(ERTM_EINVAL, "recv %d: missing injection tables", k);
    for (int c = 0; c < R.ncomp; ++c)
        if (!R.records[c]) return fail(ERTM_EINVAL, "recv %d: record %d missing", k, c);
    (void)P;
    return ERTM_OK;
}

extern "C" {

int ertm_abi_version(void) { return ERTM_ABI_VERSION; }

int ertm_struct_sizes(int64_t* out8) {
    if (!out8) return fail(ERTM_EINVAL, "struct_sizes: NULL");
    out8[0] = (int64_t)sizeof(ErtmProp);
    out8[1] = (int64_t)sizeof(ErtmRecv);
    out8[2] = (int64_t)sizeof(ErtmImage);
    out8[3] = (int64_t)sizeof(ErtmRun);
    out8[4] = (int64_t)offsetof(ErtmProp, state);
    out8[5] = (int64_t)offsetof(ErtmRun, ctl);
    out8[6] = (int64_t)sizeof(ErtmSeg);
    out8[7] = (int64_t)offsetof(ErtmSeg, wavelet);
    return ERTM_OK;
}

const char* ertm_last_error(void) { return g_err.c_str(); }

int ertm_device_info(int32_t* device, int32_t* n_cu, char* name, size_t name_len) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (device) *device = dev;
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (name && name_len) {
        snprintf(name, name_len, "%s", prop.gcnArchName);
    }
    return ERTM_OK;
}

int ertm_batch_create(const ErtmProp* props, const ErtmRecv* recv, const ErtmImage* img,
                      int32_t n, ertm_batch_t* out) {
    if (!out || !props || n <= 0) return fail(ERTM_EINVAL, "batch_create: bad arguments");
    *out = nullptr;
    auto* b = new ertm_batch_s();
    b->props.assign(props, props + n);
    b->recv.resize(n);
    b->img.resize(n);
    std::memset(b->recv.data(), 0, sizeof(ErtmRecv) * n);
    std::memset(b->img.data(), 0, sizeof(ErtmImage) * n);
    if (recv) b->recv.assign(recv, recv + n);
    if (img) b->img.assign(img, img + n);
    b->order = props[0].order;
    b->all_visco = true;
    b->all_elastic = true;
    b->all_acoustic = true;
    for (int k = 0; k < n; ++k) {
        int rc = check_prop(b->props[k], k);
        if (rc == ERTM_OK) rc = check_recv(b->recv[k], b->props[k], k);
        if (rc == ERTM_OK && b->props[k].order != b->order)
            rc = fail(ERTM_EINVAL, "batch mixes stencil orders");
        if (rc != ERTM_OK) {
            delete b;
            return rc;
        }
        const ErtmProp& P = b->props[k];
        b->max_tiles = std::max(b->max_tiles, tiles_of(P));
        b->max_samp = std::max(b->max_samp, b->recv[k].nsamp);
        b->max_cells = std::max(b->max_cells, b->recv[k].ncells);
        b->max_img = std::max(b->max_img, P.nz * P.nx);
        if (P.engine == ERTM_ELASTIC && P.mechanism == ERTM_SRC_FORCE_Z && P.src_iz >= 0)
            b->any_force = true;
        if (P.engine != ERTM_VISCO) b->all_visco = false;
        if (P.engine != ERTM_ELASTIC) b->all_elastic = false;
        if (P.engine != ERTM_ACOUSTIC) b->all_acoustic = false;
    }
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t sp = up(sizeof(ErtmProp) * n), sr = up(sizeof(ErtmRecv) * n);
    const size_t si = up(sizeof(ErtmImage) * n), ss = up(sizeof(float) * 2 * n + 64);
    b->block_bytes = sp + sr + si + ss;
    b->block = pool_get(b->block_bytes);
    if (!b->block) {
        delete b;
        return fail(ERTM_ENOMEM, "batch_create: device allocation failed");
    }
    char* base = (char*)b->block;
    b->dprops = (ErtmProp*)base;
    b->drecv = (ErtmRecv*)(base + sp);
    b->dimg = (ErtmImage*)(base + sp + sr);
    b->scratch = (float*)(base + sp + sr + si);
    hipError_t c1 = hipMemcpy(b->dprops, b->props.data(), sizeof(ErtmProp) * n, hipMemcpyHostToDevice);
    hipError_t c2 = hipMemcpy(b->drecv, b->recv.data(), sizeof(ErtmRecv) * n, hipMemcpyHostToDevice);
    hipError_t c3 = hipMemcpy(b->dimg, b->img.data(), sizeof(ErtmImage) * n, hipMemcpyHostToDevice);
    if (c1 != hipSuccess || c2 != hipSuccess || c3 != hipSuccess) {
        delete b;
        return fail(ERTM_EHIP, "batch_create: descriptor upload failed");
    }
    *out = b;
    return ERTM_OK;
}

int ertm_batch_update_image(ertm_batch_t b, const ErtmImage* img, void* stream) {
    if (!b || !img) return fail(ERTM_EINVAL, "batch_update_image: bad arguments");
    const int n = (int)b->props.size();
    hipStream_t s = (hipStream_t)stream;
    // the old table may be read by work already enqueued on the batch's streams
    for (hipStream_t t : b->streams) HIPCHK(hipStreamSynchronize(t));
    b->img.assign(img, img + n);
    HIPCHK(hipMemcpyAsync(b->dimg, b->img.data(), sizeof(ErtmImage) * n, hipMemcpyHostToDevice,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    return ERTM_OK;
}

int ertm_batch_destroy(ertm_batch_t b) {
    if (!b) return ERTM_OK;
    // A batch with captured graphs, FFT plans or a capture stream waits for
    // its streams; otherwise its device blocks (descriptor tables, clock,
    // tables, scratch) return to the pool once the launches already queued
    // on its streams have run, and the caller's streams go on undisturbed.
    bool wait = !b->graphs.empty() || !b->fft.empty() || b->cap;
    if (!wait) {
        void** blocks[] = {&b->block, (void**)&b->dclk, (void**)&b->damp, (void**)&b->dark,
                           &b->psi1, (void**)&b->rscale};
        const size_t sizes[] = {b->block_bytes, 64, b->damp_bytes, sizeof(float*) * b->props.size(),
                                b->psi1_bytes, b->rscale_bytes};
        for (int i = 0; i < 6 && !wait; ++i) {
            if (!*blocks[i]) continue;
            if (pool_put_after(*blocks[i], sizes[i], b->streams)) *blocks[i] = nullptr;
            else wait = true;
        }
    }
    if (wait)
        for (hipStream_t t : b->streams) (void)hipStreamSynchronize(t);
    delete b;
    return ERTM_OK;
}

static bool pressure_fused_ok(ertm_batch_t b);

int ertm_batch_set_variant(ertm_batch_t b, int32_t variant, int32_t zc) {
    if (!b) return fail(ERTM_EINVAL, "set_variant: NULL batch");
    if (variant < 0 || variant > 4) return fail(ERTM_EINVAL, "unknown kernel variant %d", variant);
    if (variant != 0 && !((b->all_elastic && b->order == 8) || pressure_fused_ok(b)))
        return fail(ERTM_EINVAL, "fused variants need an order-8 batch of one engine: elastic, "
                                 "acoustic, or viscoacoustic with 1..4 mechanisms");
    if (zc < 8 || zc > 4096) return fail(ERTM_EINVAL, "z chunk %d out of [8, 4096]", zc);
    b->variant = variant;
    b->zc = zc;
    return ERTM_OK;
}

int ertm_batch_set_ring_flush(ertm_batch_t b, int32_t period) {
    if (!b) return fail(ERTM_EINVAL, "set_ring_flush: NULL batch");
    if (period < 0) return fail(ERTM_EINVAL, "set_ring_flush: period %d < 0", period);
    if (period > 0 && !b->fls) {
        HIPCHK(hipStreamCreateWithFlags(&b->fls, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&b->fl_step, hipEventDisableTiming));
        for (hipEvent_t& e : b->fl_done) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    b->flush_period = period;
    return ERTM_OK;
}

int ertm_batch_set_two_step(ertm_batch_t b, int32_t variant, int32_t zc) {
    if (!b) return fail(ERTM_EINVAL, "set_two_step: NULL batch");
    if (variant == 0) {
        b->v2 = 0;
        return ERTM_OK;
    }
    if (variant != 2) return fail(ERTM_EINVAL, "two-step launch shape %d: only 2 is built", variant);
    if (!(b->all_elastic && b->order == 8))
        return fail(ERTM_EINVAL, "two-step launches need an elastic batch at order 8");
    if (zc < 8 || zc > 4096) return fail(ERTM_EINVAL, "z chunk %d out of [8, 4096]", zc);
    if (!b->psi1) {
        // one CPML part per propagator, each from a 256-byte boundary
        size_t total = 0;
        std::vector<size_t> off;
        for (const ErtmProp& P : b->props) {
            off.push_back(total);
            const int64_t floats = P.state_floats - (int64_t)P.nfields * P.fstride;
            total += ((size_t)std::max<int64_t>(floats, 1) * 4 + 255) / 256 * 256;
        }
        b->psi1 = pool_get(total);
        if (!b->psi1) return fail(ERTM_ENOMEM, "set_two_step: CPML scratch allocation failed");
        b->psi1_bytes = total;
        b->psi1_ptr.clear();
        for (size_t o : off) b->psi1_ptr.push_back((float*)((char*)b->psi1 + o));
    }
    b->v2 = variant;
    b->zc2 = zc;
    return ERTM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// launch helpers (chunks of ERTM_MAX_LAUNCH propagators)
// ---------------------------------------------------------------------------
// Launch timing: while enabled, every launch made outside a captured loop is
// bracketed by a pair of HIP events recorded on its stream and tagged with
// its kind (ERTM_TK_*); reading sums the elapsed times per kind.  Event
// pairs are pooled and reused across reads.
namespace {
struct LaunchTimer {
    std::mutex m;
    bool on = false;
    std::vector<hipEvent_t> ev;     // 2 per recorded launch
    std::vector<int8_t> kind;       // 1 per recorded launch
    size_t used = 0;                // events in use
    int64_t cells[ERTM_TK_COUNT] = {};  // physical cells covered, per kind
};
LaunchTimer g_timer;

int timer_event(hipEvent_t* out) {
    if (g_timer.used == g_timer.ev.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        g_timer.ev.push_back(e);
    }
    *out = g_timer.ev[g_timer.used++];
    return ERTM_OK;
}

// One timed launch: the constructor takes the timer's lock and records the
// start event if timing is on (and the launch is not being captured); done()
// records the end event under the given kind.
class Timed {
  public:
    Timed(hipStream_t s, bool captured) : s_(s), lk_(g_timer.m, std::defer_lock) {
        if (captured) return;
        lk_.lock();
        if (!g_timer.on) {
            lk_.unlock();
            return;
        }
        rc_ = timer_event(&e0_);
        if (rc_ == ERTM_OK) {
            const hipError_t e = hipEventRecord(e0_, s_);
            if (e != hipSuccess) rc_ = fail(ERTM_EHIP, "hipEventRecord: %s", hipGetErrorString(e));
        }
        if (rc_ != ERTM_OK) e0_ = nullptr;
    }
    int status() const { return rc_; }
    int done(int kind, int64_t cells) {
        if (!e0_) return rc_;
        hipEvent_t e1 = nullptr;
        int rc = timer_event(&e1);
        if (rc != ERTM_OK) return rc;
        HIPCHK(hipEventRecord(e1, s_));
        g_timer.kind.push_back((int8_t)kind);
        g_timer.cells[kind] += cells;
        e0_ = nullptr;
        return ERTM_OK;
    }

  private:
    hipStream_t s_;
    std::unique_lock<std::mutex> lk_;
    hipEvent_t e0_ = nullptr;
    int rc_ = ERTM_OK;
};
}  // namespace

// Can the fused elastic kernel do imaging op `op` inside the step?  Needs a
// fused PD = 1 variant, no force_z source (the stored vz then differs from the
// one the stencil sees), and every propagator's accumulators in one
// allocation starting at num[0] (as ertm/device.py allocates them).
// Does the fused pressure kernel (ertm_fused.hip: acoustic, or viscoacoustic
// with 1..4 mechanisms, order 8) cover this batch?
static bool pressure_fused_ok(ertm_batch_t b) {
    if (b->order != 8 || b->props.empty()) return false;
    if (b->all_acoustic) return true;
    if (!b->all_visco) return false;
    const int L = b->props[0].nmech;
    for (const ErtmProp& P : b->props)
        if (P.nmech != L) return false;
    return L >= 1 && L <= 4;
}

// The propagators a launch sequence covers: every one of the batch in table
// order (Members == nullptr), or the listed ones (lanes), each at its own
// parity pcur[p] and replay-slot origin pt0[p] (both indexed by propagator).
struct Members {
    const int* idx;
    int n;
    const int32_t* pcur;
    const int32_t* pt0;
};

static int count_of(ertm_batch_t b, const Members* m) {
    return m ? m->n : (int)b->props.size();
}

static int member_at(const Members* m, int i) { return m ? m->idx[i] : i; }

// Chunk [i0, i0 + cnt) of the members into la: table indices, parities and
// replay-slot origins (time and amplitude are the caller's to fill).
static void fill_members(LaunchArgs& la, const Members* m, int i0, int cnt, int cur) {
    la.n = cnt;
    la.cur = cur;
    if (!m) {
        la.base = i0;
        return;
    }
    la.base = 0;
    la.cur = 0;
    la.remap = 1;
    for (int k = 0; k < cnt; ++k) {
        const int p = m->idx[i0 + k];
        la.slot[k] = p;
        if (m->pcur[p]) la.flip |= (uint64_t)1 << k;
        la.t0[k] = m->pt0 ? m->pt0[p] : 0;
    }
}

static bool fused_image_ok(ertm_batch_t b, int op, const Members* m = nullptr) {
    if (op != ERTM_IMG_EXTRACT && op != ERTM_IMG_CORRELATE) return false;
    const int cnt = count_of(b, m);
    // the fused pressure step extracts (its correlation goes through the ring)
    if (pressure_fused_ok(b) && b->variant >= 1 && b->variant <= 4 && op == ERTM_IMG_EXTRACT) {
        for (int i = 0; i < cnt; ++i) {
            const int k = member_at(m, i);
            const ErtmProp& P = b->props[k];
            if (!b->img[k].rbuf || (int64_t)P.nz * P.nx * 4 >= ((int64_t)1 << 31) - 64) return false;
        }
        return true;
    }
    if (b->variant < 1 || b->variant > 4 || !b->all_elastic || b->order != 8 || b->any_force)
        return false;
    for (int i = 0; i < cnt; ++i) {
        const int k = member_at(m, i);
        const ErtmImage& I = b->img[k];
        const ErtmProp& P = b->props[k];
        if (!I.rbuf || I.nrw < (op == ERTM_IMG_CORRELATE ? 1 : 0) || I.nrw > 2) return false;
        if ((int64_t)P.nz * P.nx * 4 >= ((int64_t)1 << 31) - 64) return false;
        if (op != ERTM_IMG_CORRELATE) continue;
        const int64_t ncell = (int64_t)P.nz * P.nx;
        const int64_t span = 8 * ncell * (2 * I.nrw + 1);
        if (!I.num[0] || span >= ((int64_t)1 << 31)) return false;
        const char* a0 = (const char*)I.num[0];
        auto inside = [&](const double* p) {
            const int64_t d = (const char*)p - a0;
            return d >= 0 && d + 8 * ncell <= span;
        };
        if (I.nrw > 1 && !inside(I.num[1])) return false;
        if ((I.illum & ERTM_ILLUM_SRC) && !(I.ill && inside(I.ill))) return false;
        if (I.illum & ERTM_ILLUM_REC)
            for (int w = 0; w < I.nrw; ++w)
                if (!(I.rill[w] && inside(I.rill[w]))) return false;
    }
    return true;
}

// Internal imaging op: the fused step stores the receiver-side fields of the
// new state into ring slot t mod R; do_flush later folds R of them into the
// accumulators (same per-cell order as R correlation passes).
#define ERTM_IMG_RING_STORE 100

static bool ring_ok(ertm_batch_t b, const Members* m = nullptr) {
    if (b->variant < 1 || b->variant > 4 || !(b->all_elastic || pressure_fused_ok(b)) ||
        b->order != 8 || b->any_force)
        return false;
    const int cnt = count_of(b, m);
    for (int i = 0; i < cnt; ++i) {
        const int k = member_at(m, i);
        const ErtmImage& I = b->img[k];
        const ErtmProp& P = b->props[k];
        if (!I.rbuf || !I.rring || I.ring_len < 1 || I.nrw < 1 || I.nrw > 2 || !I.num[0])
            return false;
        if ((int64_t)P.nz * P.ld * 4 * I.nrw >= ((int64_t)1 << 31) - 64) return false;
        if (I.ring_len != b->img[member_at(m, 0)].ring_len) return false;
    }
    return true;
}

// Fused imaging kind for the kernel (ertm_fused.hip): extraction and ring
// stores use their no-S variants (fewer registers) unless some propagator of
// the launch images the divcurl S field.
static int img_kind(ertm_batch_t b, int op, const LaunchArgs& la) {
    bool s = false;
    for (int i = 0; i < la.n; ++i) {
        const ErtmImage& I = b->img[la_prop(la, i)];
        if (I.separation != ERTM_SEP_DIVCURL) continue;
        if (op == ERTM_IMG_EXTRACT) s = s || I.swave == ERTM_WAVE_S;
        else
            for (int w = 0; w < I.nrw; ++w) s = s || I.rwave[w] == ERTM_WAVE_S;
    }
    if (op == ERTM_IMG_EXTRACT) return s ? 1 : 4;
    if (op == ERTM_IMG_RING_STORE) return s ? 3 : 5;
    return op == ERTM_IMG_CORRELATE ? 2 : 0;
}

// Launches recorded into a captured time loop read time and amplitude from
// the batch's device clock (LaunchArgs::clk); qoff shifts the step they see.
// clk[0] holds the first step of the chunk being replayed; the launches of
// the chunk's step j see step clk[0] + j (q is j while recording).
struct DevClock {
    const int32_t* clk;
    const float* ta;
    int ntab;
    int q;
};

static void use_clock(LaunchArgs& la, const DevClock* dc, int qoff = 0) {
    if (!dc) return;
    la.clk = dc->clk;
    la.ta = dc->ta;
    la.ntab = dc->ntab;
    la.qoff = dc->q + qoff;
}

// Physical cells of the propagators of a launch.
namespace {
int64_t cells_of(ertm_batch_t b, const LaunchArgs& la) {
    int64_t c = 0;
    for (int k = 0; k < la.n; ++k) {
        const ErtmProp& P = b->props[la_prop(la, k)];
        c += (int64_t)P.nz * P.nx;
    }
    return c;
}

// One whole-state copy of propagator `prop` (for the timers' cell count).
struct CopyItem {
    const float* src;
    float* dst;
    int64_t floats;
    int prop;
};
}  // namespace

// The copies in one k_copy_batch launch per ERTM_MAX_LAUNCH of them when
// every one is 16-byte aligned and a whole number of float4; otherwise one
// async copy each.
static int copy_list(ertm_batch_t b, const std::vector<CopyItem>& items, hipStream_t s) {
    bool fast = true;
    for (const CopyItem& c : items)
        fast = fast && c.floats % 4 == 0 && ((uintptr_t)c.src | (uintptr_t)c.dst) % 16 == 0;
    if (!fast) {
        for (const CopyItem& c : items)
            HIPCHK(hipMemcpyAsync(c.dst, c.src, sizeof(float) * (size_t)c.floats,
                                  hipMemcpyDeviceToDevice, s));
        return ERTM_OK;
    }
    for (size_t base = 0; base < items.size(); base += ERTM_MAX_LAUNCH) {
        CopyArgs a{};
        a.n = (int)std::min<size_t>(ERTM_MAX_LAUNCH, items.size() - base);
        int64_t cells = 0;
        for (int k = 0; k < a.n; ++k) {
            const CopyItem& c = items[base + k];
            a.src[k] = (const float4*)c.src;
            a.dst[k] = (float4*)c.dst;
            a.n4[k] = c.floats / 4;
            cells += (int64_t)b->props[c.prop].nz * b->props[c.prop].nx;
        }
        Timed tm(s, false);
        if (tm.status() != ERTM_OK) return tm.status();
        launch_copy_batch(a, s);
        LAUNCHCHK("copy_batch");
        if (int rc = tm.done(ERTM_TK_COPY, cells)) return rc;
    }
    return ERTM_OK;
}

// Copy state[cur] of every propagator to (to_arena) or from its arena at
// arenas[k] + slot * stride.
static int copy_states(ertm_batch_t b, int cur, float* const* arenas, int64_t slot,
                       int64_t stride, bool to_arena, hipStream_t s) {
    std::vector<CopyItem> items;
    for (size_t k = 0; k < b->props.size(); ++k) {
        const ErtmProp& P = b->props[k];
        float* a = arenas[k] + slot * stride;
        items.push_back({to_arena ? P.state[cur] : a, to_arena ? a : P.state[cur], P.state_floats,
                         (int)k});
    }
    return copy_list(b, items, s);
}

// t_hi, amp and t arrays of the do_* helpers are indexed by propagator.
static int do_flush(ertm_batch_t b, const int32_t* t_hi, int count, hipStream_t s,
                    const DevClock* dc = nullptr, const Members* m = nullptr,
                    const float* rscale = nullptr) {
    const int n = count_of(b, m);
    for (int i0 = 0; i0 < n; i0 += ERTM_MAX_LAUNCH) {
        LaunchArgs la{};
        fill_members(la, m, i0, std::min(ERTM_MAX_LAUNCH, n - i0), 0);
        la.op = count;
        for (int k = 0; k < la.n; ++k) la.t[k] = t_hi[la_prop(la, k)];
        use_clock(la, dc, -(count - 1));  // the group's first (latest) step
        Timed tm(s, dc != nullptr);
        if (tm.status() != ERTM_OK) return tm.status();
        launch_image_flush(b->dprops, b->dimg, la, b->max_img, rscale, (int)b->props.size(), s);
        LAUNCHCHK("image_flush");
        // cells: the physical cells of the flushed steps
        if (int rc = tm.done(ERTM_TK_FLUSH, count * cells_of(b, la))) return rc;
    }
    return ERTM_OK;
}

// One ring fold on the batch's flush stream, beside the backward steps: it
// starts after the steps queued on s so far (they wrote its slots), and s
// waits for the previous fold before the steps that follow, which overwrite
// that fold's slots.  Folds run in order on their stream, so the
// accumulators see them in step order.
static int flush_beside(ertm_batch_t b, const int32_t* t_hi, int count, hipStream_t s,
                        const float* rscale) {
    b->note(b->fls);
    HIPCHK(hipEventRecord(b->fl_step, s));
    HIPCHK(hipStreamWaitEvent(b->fls, b->fl_step, 0));
    if (b->fl_count > 0) HIPCHK(hipStreamWaitEvent(s, b->fl_done[(b->fl_count - 1) & 1], 0));
    if (int rc = do_flush(b, t_hi, count, b->fls, nullptr, nullptr, rscale)) return rc;
    HIPCHK(hipEventRecord(b->fl_done[b->fl_count & 1], b->fls));
    ++b->fl_count;
    return ERTM_OK;
}

// Timing kind of a step launch: the fused kernel's imaging variant, or the
// two-launch pair.
static int step_kind(bool fused, int kind) {
    if (!fused) return ERTM_TK_TWO_LAUNCH;
    switch (kind) {
        case 0: return ERTM_TK_PLAIN;
        case 1: case 4: return ERTM_TK_EXTRACT;
        case 3: case 5: return ERTM_TK_RING;
        default: return ERTM_TK_CORRELATE;
    }
}

// One step; with img_op >= 0 and *imaged set on return, the step launches
// also did that imaging op (t = time index per propagator).
// shape: launch shape (variant, zc) instead of the batch's, or nullptr.
static int do_step(ertm_batch_t b, int cur, const float* amp, hipStream_t s,
                   int img_op = -1, const int32_t* t = nullptr, bool* imaged = nullptr,
                   const DevClock* dc = nullptr, const Members* m = nullptr,
                   const int* shape = nullptr) {
    const int variant = shape ? shape[0] : b->variant, zc = shape ? shape[1] : b->zc;
    const int n = count_of(b, m);
    const bool fuse_img = img_op >= 0 && t &&
                          (img_op == ERTM_IMG_RING_STORE ? ring_ok(b, m) : fused_image_ok(b, img_op, m));
    if (imaged) *imaged = fuse_img;
    for (int i0 = 0; i0 < n; i0 += ERTM_MAX_LAUNCH) {
        LaunchArgs la{};
        fill_members(la, m, i0, std::min(ERTM_MAX_LAUNCH, n - i0), cur);
        for (int k = 0; k < la.n; ++k) la.amp[k] = amp ? amp[la_prop(la, k)] : 0.0f;
        if (fuse_img)
            for (int k = 0; k < la.n; ++k) la.t[k] = t[la_prop(la, k)];
        use_clock(la, dc);
        const int kind = fuse_img ? img_kind(b, img_op, la) : 0;
        Timed tm(s, dc != nullptr);
        if (tm.status() != ERTM_OK) return tm.status();
        bool fused = variant != 0 &&
                     launch_step_fused(b->dprops, b->props.data(), la, b->order, variant, zc,
                                       b->dimg, kind, s);
        if (!fused) {
            if (fuse_img) return fail(ERTM_ESTATE, "step: fused imaging launch refused");
            launch_step(b->dprops, la, b->order, b->max_tiles, b->any_force, s);
        }
        LAUNCHCHK("step");
        if (int rc = tm.done(step_kind(fused, kind), cells_of(b, la))) return rc;
    }
    return ERTM_OK;
}

// Two steps t, t + 1 (amplitudes a1, a2): state[cur] -> state[cur ^ 1]; with
// img_op = ERTM_IMG_EXTRACT the replay slots of both steps are written too.
static int do_step2(ertm_batch_t b, int cur, const float* a1, const float* a2, hipStream_t s,
                    int img_op, const int32_t* t) {
    const int n = (int)b->props.size();
    for (int base = 0; base < n; base += ERTM_MAX_LAUNCH) {
        LaunchArgs la{};
        El2Args e2{};
        la.base = base;
        la.n = std::min(ERTM_MAX_LAUNCH, n - base);
        la.cur = cur;
        for (int k = 0; k < la.n; ++k) {
            la.amp[k] = a1[base + k];
            e2.amp2[k] = a2[base + k];
            e2.psi1[k] = b->psi1_ptr[base + k];
            la.t[k] = t[base + k];
        }
        const int kind = img_op == ERTM_IMG_EXTRACT ? img_kind(b, img_op, la) : 0;
        Timed tm(s, false);
        if (tm.status() != ERTM_OK) return tm.status();
        if (!launch_step2_fused(b->dprops, b->props.data(), la, e2, b->v2, b->zc2, b->dimg, kind,
                                s))
            return fail(ERTM_ESTATE, "two-step launch refused");
        LAUNCHCHK("step2");
        if (int rc = tm.done(kind ? ERTM_TK_EXTRACT2 : ERTM_TK_PLAIN2, 2 * cells_of(b, la)))
            return rc;
    }
    return ERTM_OK;
}

static int do_sample(ertm_batch_t b, int cur, const int32_t* t, hipStream_t s,
                     const DevClock* dc = nullptr, const Members* m = nullptr) {
    const int n = count_of(b, m);
    for (int i0 = 0; i0 < n; i0 += ERTM_MAX_LAUNCH) {
        LaunchArgs la{};
        fill_members(la, m, i0, std::min(ERTM_MAX_LAUNCH, n - i0), cur);
        for (int k = 0; k < la.n; ++k) la.t[k] = t[la_prop(la, k)];
        use_clock(la, dc);
        Timed tm(s, dc != nullptr);
        if (tm.status() != ERTM_OK) return tm.status();
        launch_sample(b->dprops, b->drecv, la, b->max_samp, s);
        LAUNCHCHK("sample");
        if (int rc = tm.done(ERTM_TK_SAMPLE, cells_of(b, la))) return rc;
    }
    return ERTM_OK;
}

static int do_inject(ertm_batch_t b, int cur, const int32_t* t, hipStream_t s,
                     const DevClock* dc = nullptr, const Members* m = nullptr) {
    const int n = count_of(b, m);
    for (int i0 = 0; i0 < n; i0 += ERTM_MAX_LAUNCH) {
        LaunchArgs la{};
        fill_members(la, m, i0, std::min(ERTM_MAX_LAUNCH, n - i0), cur);
        for (int k = 0; k < la.n; ++k) la.t[k] = t[la_prop(la, k)];
        use_clock(la, dc);
        Timed tm(s, dc != nullptr);
        if (tm.status() != ERTM_OK) return tm.status();
        launch_inject(b->dprops, b->drecv, la, b->max_cells, s);
        LAUNCHCHK("inject");
        if (int rc = tm.done(ERTM_TK_INJECT, cells_of(b, la))) return rc;
    }
    return ERTM_OK;
}

static int do_image(ertm_batch_t b, int cur, const int32_t* t, int op, hipStream_t s,
                    const DevClock* dc = nullptr, const Members* m = nullptr) {
    const int n = count_of(b, m);
    for (int i0 = 0; i0 < n; i0 += ERTM_MAX_LAUNCH) {
        LaunchArgs la{};
        fill_members(la, m, i0, std::min(ERTM_MAX_LAUNCH, n - i0), cur);
        la.op = op;
        for (int k = 0; k < la.n; ++k) la.t[k] = t[la_prop(la, k)];
        use_clock(la, dc);
        Timed tm(s, dc != nullptr);
        if (tm.status() != ERTM_OK) return tm.status();
        launch_image(b->dprops, b->dimg, la, b->order, b->max_img, s);
        LAUNCHCHK("image");
        if (int rc = tm.done(ERTM_TK_IMAGE, cells_of(b, la))) return rc;
    }
    return ERTM_OK;
}

// rscale, ring_slot: with the receiver-field ring, the slot's row of
// per-propagator factors (k_stab_control writes it) and the slot index the
// low-pass rewrites with the filtered pressure; nullptr and -1 without.
static int do_stabilize(ertm_batch_t b, int cur, int step_index, double cap, int interval,
                        double* ctl, hipStream_t s, float* rscale = nullptr,
                        int ring_slot = -1) {
    if (!b->all_visco) return fail(ERTM_EINVAL, "stabilize: batch is not viscoacoustic");
    if (!ctl) return fail(ERTM_EINVAL, "stabilize: ctl is NULL");
    const int n = (int)b->props.size();
    if (n > ERTM_MAX_LAUNCH) return fail(ERTM_EINVAL, "stabilize: batch larger than %d", ERTM_MAX_LAUNCH);
    float* peaks = b->scratch;
    float* scales = b->scratch + n;
    LaunchArgs la{};
    la.base = 0;
    la.n = n;
    la.cur = cur;
    int64_t max_n = 0;
    for (const ErtmProp& P : b->props) max_n = std::max<int64_t>(max_n, (int64_t)P.nzp * P.ld);
    launch_absmax_batch(b->dprops, la, ERTM_VI_P, max_n, peaks, s);
    LAUNCHCHK("absmax");
    const bool lowpass_now = step_index >= 30 && interval > 1 && step_index % interval == 0;
    launch_stab_control(b->dprops, la, peaks, ctl, step_index, cap, lowpass_now ? 1 : 0, scales,
                        rscale, s);
    LAUNCHCHK("stab_control");
    if (step_index < 30) return ERTM_OK;
    launch_stab_scale(b->dprops, la, scales, b->max_tiles, s);
    LAUNCHCHK("stab_scale");
    if (lowpass_now) {
        // members grouped by padded shape, in groups of at most ERTM_LP_GROUP;
        // each group is one batched D2Z / Z2D pair over its p and eps fields
        std::map<std::pair<int, int>, std::vector<int>> shapes;
        for (int k = 0; k < n; ++k) shapes[{b->props[k].nzp, b->props[k].nxp}].push_back(k);
        for (const auto& sh : shapes) {
            const int nzp = sh.first.first, nxp = sh.first.second, nh = nxp / 2 + 1;
            const std::vector<int>& ks = sh.second;
            for (size_t g0 = 0; g0 < ks.size(); g0 += ERTM_LP_GROUP) {
                const int m = (int)std::min<size_t>(ERTM_LP_GROUP, ks.size() - g0);
                FftCache& fc = b->fft[std::make_tuple(nzp, nxp, m)];
                if (!fc.inv) {
                    if ((int64_t)nzp * nxp > INT_MAX)
                        return fail(ERTM_EINVAL, "lowpass: padded grid %dx%d too large", nzp, nxp);
                    int dims[2] = {nzp, nxp}, remb[2] = {nzp, nxp}, cemb[2] = {nzp, nh};
                    const int rdist = nzp * nxp, cdist = nzp * nh;
                    if (!fc.real) HIPCHK(hipMalloc(&fc.real, sizeof(double) * 2 * (size_t)m * rdist));
                    if (!fc.cplx)
                        HIPCHK(hipMalloc(&fc.cplx, 2 * sizeof(double) * 2 * (size_t)m * cdist));
                    if (!fc.fwd && hipfftPlanMany(&fc.fwd, 2, dims, remb, 1, rdist, cemb, 1, cdist,
                                                  HIPFFT_D2Z, 2 * m) != HIPFFT_SUCCESS) {
                        fc.fwd = 0;
                        return fail(ERTM_EHIP, "hipfftPlanMany D2Z (%d,%d) x %d failed", nzp, nxp,
                                    2 * m);
                    }
                    if (hipfftPlanMany(&fc.inv, 2, dims, cemb, 1, cdist, remb, 1, rdist,
                                       HIPFFT_Z2D, 2 * m) != HIPFFT_SUCCESS) {
                        fc.inv = 0;
                        return fail(ERTM_EHIP, "hipfftPlanMany Z2D (%d,%d) x %d failed", nzp, nxp,
                                    2 * m);
                    }
                }
                if (hipfftSetStream(fc.fwd, s) != HIPFFT_SUCCESS ||
                    hipfftSetStream(fc.inv, s) != HIPFFT_SUCCESS)
                    return fail(ERTM_EHIP, "hipfftSetStream failed");
                LpArgs a{};
                a.n = m;
                a.nzp = nzp;
                a.nxp = nxp;
                a.cur = cur;
                a.ring_slot = ring_slot;
                for (int i = 0; i < m; ++i) {
                    const int k = ks[g0 + i];
                    const double time = step_index * b->props[k].dt64;
                    double ratio = 0.9 - 0.4 * std::min(time / 1.0, 1.0);
                    a.prop[i] = k;
                    a.ratio[i] = std::max(ratio, 0.35);
                }
                launch_lowpass_prep(b->dprops, a, fc.real, s);
                LAUNCHCHK("lowpass_prep");
                if (hipfftExecD2Z(fc.fwd, fc.real, (hipfftDoubleComplex*)fc.cplx) != HIPFFT_SUCCESS)
                    return fail(ERTM_EHIP, "hipfftExecD2Z failed");
                launch_lowpass_response(b->dprops, a, fc.cplx, s);
                LAUNCHCHK("lowpass_response");
                if (hipfftExecZ2D(fc.inv, (hipfftDoubleComplex*)fc.cplx, fc.real) != HIPFFT_SUCCESS)
                    return fail(ERTM_EHIP, "hipfftExecZ2D failed");
                launch_lowpass_finish(b->dprops, b->dimg, a, fc.real, s);
                LAUNCHCHK("lowpass_finish");
            }
        }
    }
    return ERTM_OK;
}

// hipFFT plans for image post-processing, one per (stream, shape).  A plan
// owns a work area that its transforms use on the GPU, so two streams never
// share one: their transforms may run at the same time.
namespace {
std::mutex g_post_m;
std::map<std::tuple<hipStream_t, int, int>, hipfftHandle> g_post_plans;
}  // namespace

// Post-processing of one image with the scratch at p->work: the normalised
// image in [0, n) doubles, then the Laplacian (n doubles, or 2n as the
// complex buffer of the low-pass).  Every scratch double a call reads was
// written earlier in the same call.
static int post_run(const ErtmPost* p, hipStream_t s) {
    const int nz = p->nz, nx = p->nx;
    const int64_t n = (int64_t)nz * nx;
    double* a = p->work;      // n doubles: normalised image
    double* c = p->work + n;  // n doubles (Laplacian) or 2n (complex buffer)
    launch_post_norm(p->num, p->ill, p->rill, n, p->eps, p->div_ill, p->div_rill, a, s);
    LAUNCHCHK("post_norm");
    const double* img = a;
    int from_complex = 0;
    if (p->laplacian) {
        launch_post_lap(a, nz, nx, p->dx * p->dx, p->dz * p->dz, p->lowpass ? 1 : 0, c, s);
        LAUNCHCHK("post_lap");
        img = c;
    }
    if (p->lowpass) {
        if (!p->laplacian) {
            launch_post_to_complex(a, n, c, s);
            LAUNCHCHK("post_to_complex");
        }
        std::lock_guard<std::mutex> lk(g_post_m);
        hipfftHandle& plan = g_post_plans[std::make_tuple(s, nz, nx)];
        if (!plan && hipfftPlan2d(&plan, nz, nx, HIPFFT_Z2Z) != HIPFFT_SUCCESS) {
            plan = 0;
            return fail(ERTM_EHIP, "hipfftPlan2d(%d,%d) failed", nz, nx);
        }
        if (hipfftSetStream(plan, s) != HIPFFT_SUCCESS) return fail(ERTM_EHIP, "hipfftSetStream failed");
        auto* z = (hipfftDoubleComplex*)c;
        if (hipfftExecZ2Z(plan, z, z, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
            return fail(ERTM_EHIP, "hipfftExecZ2Z forward failed");
        launch_post_response(c, nz, nx, p->dx, p->dz, p->fraction, s);
        LAUNCHCHK("post_response");
        if (hipfftExecZ2Z(plan, z, z, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
            return fail(ERTM_EHIP, "hipfftExecZ2Z backward failed");
        img = c;
        from_complex = 1;
    }
    launch_post_finish(img, from_complex, nz, nx, p->npml, p->mask_top, p->out, s);
    LAUNCHCHK("post_finish");
    return ERTM_OK;
}

static int post_check(const ErtmPost* p, const char* who) {
    if (!p || p->nz <= 0 || p->nx <= 0 || !p->num || !p->out)
        return fail(ERTM_EINVAL, "%s: bad arguments", who);
    if ((p->div_ill && !p->ill) || (p->div_rill && !p->rill))
        return fail(ERTM_EINVAL, "%s: normalisation array missing", who);
    return ERTM_OK;
}

// Static device scratch of ertm_postprocess_noalloc: the normalised image and
// the Laplacian, 2 * ERTM_NOALLOC_MAX_CELLS doubles (32 MiB).  Every GPU the
// library runs on holds its own copy until the process ends.  In host code
// the name `scratch` is only a host-side stand-in: kernels get the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[2 * (size_t)ERTM_NOALLOC_MAX_CELLS];

namespace {
std::mutex g_scratch_m;
std::vector<double*> g_scratch_dev;  // per device id; nullptr until looked up
}  // namespace

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static int scratch_of_device(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_m);
    if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        void* addr = nullptr;
        HIPCHK(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
        if (!addr) return fail(ERTM_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_dev[dev] = (double*)addr;
    }
    *out = g_scratch_dev[dev];
    return ERTM_OK;
}

extern "C" {

int ertm_postprocess(const ErtmPost* p, void* stream) {
    int rc = post_check(p, "postprocess");
    if (rc != ERTM_OK) return rc;
    if (!p->work) return fail(ERTM_EINVAL, "postprocess: work is NULL");
    return post_run(p, (hipStream_t)stream);
}

int ertm_postprocess_noalloc(const ErtmPost* p, void* stream) {
    int rc = post_check(p, "postprocess_noalloc");
    if (rc != ERTM_OK) return rc;
    const int64_t n = (int64_t)p->nz * p->nx;
    if (n > (int64_t)ERTM_NOALLOC_MAX_CELLS)
        return fail(ERTM_EINVAL,
                    "postprocess_noalloc: %d x %d image (%lld cells) exceeds the static scratch "
                    "of %d cells; use ertm_postprocess with a work buffer",
                    p->nz, p->nx, (long long)n, ERTM_NOALLOC_MAX_CELLS);
    if (p->lowpass)
        return fail(ERTM_EINVAL,
                    "postprocess_noalloc: the wavenumber low-pass needs a hipFFT plan, which "
                    "allocates; use ertm_postprocess");
    double* w = nullptr;
    if ((rc = scratch_of_device(&w)) != ERTM_OK) return rc;
    ErtmPost q = *p;
    q.work = w;
    return post_run(&q, (hipStream_t)stream);
}

int ertm_prepare_record(float* rec, int32_t nt, int32_t nrec, const double* t,
                        const double* t_mute, double taper, int32_t w_end, void* stream) {
    if (!rec || nt <= 0 || nrec <= 0 || (t_mute && !t) || w_end < 0)
        return fail(ERTM_EINVAL, "prepare_record: bad arguments");
    launch_prep_mute_taper(rec, nt, nrec, t, t_mute, taper, w_end, (hipStream_t)stream);
    LAUNCHCHK("prep_mute_taper");
    return ERTM_OK;
}

int ertm_rotate_components(float* vx, float* vz, int64_t n, double cos_a, double sin_a,
                           void* stream) {
    if (!vx || !vz || n <= 0) return fail(ERTM_EINVAL, "rotate_components: bad arguments");
    launch_prep_rotate(vx, vz, n, cos_a, sin_a, (hipStream_t)stream);
    LAUNCHCHK("prep_rotate");
    return ERTM_OK;
}

int ertm_resample_record(const float* in, int32_t n_in, int32_t nrec, const double* w,
                         const int32_t* off, int32_t n_out, int32_t taps, float* out,
                         int32_t out_rows, void* stream) {
    if (!in || !w || !off || !out || n_in < 1 || nrec < 1 || n_out < 0 || out_rows < 0 ||
        taps < 1 || taps > 64)
        return fail(ERTM_EINVAL, "resample_record: bad arguments");
    if ((int64_t)out_rows * nrec >= ((int64_t)1 << 40))
        return fail(ERTM_EINVAL, "resample_record: record too large");
    if (out_rows == 0) return ERTM_OK;
    launch_resample(in, n_in, nrec, w, off, n_out, taps, out, out_rows, (hipStream_t)stream);
    LAUNCHCHK("resample");
    return ERTM_OK;
}

int ertm_step_timing(int32_t enable) {
    std::lock_guard<std::mutex> lk(g_timer.m);
    if (enable) {
        g_timer.used = 0;
        g_timer.kind.clear();
        for (int64_t& c : g_timer.cells) c = 0;
    }
    g_timer.on = enable != 0;
    return ERTM_OK;
}

int ertm_timing_read(double* ms, int64_t* launches, int64_t* cells) {
    if (!ms || !launches || !cells) return fail(ERTM_EINVAL, "timing_read: bad arguments");
    std::lock_guard<std::mutex> lk(g_timer.m);
    for (int k = 0; k < ERTM_TK_COUNT; ++k) {
        ms[k] = 0.0;
        launches[k] = 0;
        cells[k] = g_timer.cells[k];
    }
    for (size_t i = 0; i < g_timer.kind.size(); ++i) {
        HIPCHK(hipEventSynchronize(g_timer.ev[2 * i + 1]));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, g_timer.ev[2 * i], g_timer.ev[2 * i + 1]));
        ms[g_timer.kind[i]] += t;
        launches[g_timer.kind[i]] += 1;
    }
    return ERTM_OK;
}

int ertm_step_timing_read(double* total_ms, int64_t* launches, int64_t* cell_updates,
                          int32_t* fused) {
    if (!total_ms || !launches || !cell_updates || !fused)
        return fail(ERTM_EINVAL, "step_timing_read: bad arguments");
    double ms[ERTM_TK_COUNT];
    int64_t n[ERTM_TK_COUNT], c[ERTM_TK_COUNT];
    if (int rc = ertm_timing_read(ms, n, c)) return rc;
    *total_ms = ms[ERTM_TK_PLAIN] + ms[ERTM_TK_TWO_LAUNCH];
    *launches = n[ERTM_TK_PLAIN] + n[ERTM_TK_TWO_LAUNCH];
    *cell_updates = c[ERTM_TK_PLAIN] + c[ERTM_TK_TWO_LAUNCH];
    *fused = n[ERTM_TK_PLAIN] > 0 ? (n[ERTM_TK_TWO_LAUNCH] > 0 ? 2 : 1)
                                  : (n[ERTM_TK_TWO_LAUNCH] > 0 ? 0 : -1);
    return ERTM_OK;
}

int ertm_step(ertm_batch_t b, int32_t cur, const float* amp, void* stream) {
    if (!b || (cur & ~1)) return fail(ERTM_EINVAL, "step: bad arguments");
    b->note((hipStream_t)stream);
    return do_step(b, cur, amp, (hipStream_t)stream);
}

int ertm_stabilize(ertm_batch_t b, int32_t cur, int32_t step_index, double cap, int32_t interval,
                   double* ctl, void* stream) {
    if (!b || (cur & ~1)) return fail(ERTM_EINVAL, "stabilize: bad arguments");
    b->note((hipStream_t)stream);
    return do_stabilize(b, cur, step_index, cap, interval, ctl, (hipStream_t)stream);
}

int ertm_sample(ertm_batch_t b, int32_t cur, const int32_t* t, void* stream) {
    if (!b || !t || (cur & ~1)) return fail(ERTM_EINVAL, "sample: bad arguments");
    b->note((hipStream_t)stream);
    return do_sample(b, cur, t, (hipStream_t)stream);
}

int ertm_inject(ertm_batch_t b, int32_t cur, const int32_t* t, void* stream) {
    if (!b || !t || (cur & ~1)) return fail(ERTM_EINVAL, "inject: bad arguments");
    b->note((hipStream_t)stream);
    return do_inject(b, cur, t, (hipStream_t)stream);
}

int ertm_image(ertm_batch_t b, int32_t cur, const int32_t* t, int32_t op, void* stream) {
    if (!b || !t || (cur & ~1) || op < ERTM_IMG_EXTRACT || op > ERTM_IMG_EXC_APPLY)
        return fail(ERTM_EINVAL, "image: bad arguments");
    b->note((hipStream_t)stream);
    return do_image(b, cur, t, op, (hipStream_t)stream);
}

int ertm_copy_state(float* dst, const float* src, int64_t nfloats, void* stream) {
    if (!dst || !src || nfloats < 0) return fail(ERTM_EINVAL, "copy_state: bad arguments");
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)nfloats, hipMemcpyDeviceToDevice,
                          (hipStream_t)stream));
    return ERTM_OK;
}

int ertm_run(ertm_batch_t b, const ErtmRun* r, int32_t* cur_out, void* stream) {
    if (!b || !r) return fail(ERTM_EINVAL, "run: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    b->note(s);
    const int n = (int)b->props.size();
    if (r->t_begin < 0 || r->t_end < r->t_begin) return fail(ERTM_EINVAL, "run: bad time range");
    if (r->phase < 0 || r->phase > 2) return fail(ERTM_EINVAL, "run: bad phase %d", r->phase);
    if (r->phase != 2 && r->t_end > r->t_begin) {
        if (!r->wavelets || r->nw < r->t_end)
            return fail(ERTM_EINVAL, "run: wavelets missing or shorter than t_end");
        for (int k = 0; k < n; ++k)
            if (!r->wavelets[k]) return fail(ERTM_EINVAL, "run: wavelet %d is NULL", k);
    }
    if (r->phase == 1 && (!r->restore_arenas || r->restore_slot < 0))
        return fail(ERTM_EINVAL, "run: replay needs restore arenas");
    if (r->phase != 2 && r->ck_interval > 0 && !r->save_arenas)
        return fail(ERTM_EINVAL, "run: checkpoint arenas missing");
    if (r->stabilize && !b->all_visco) return fail(ERTM_EINVAL, "run: stabilize needs visco");
    for (int k = 0; k < n; ++k) {
        const int64_t sf = b->props[k].state_floats;
        if (r->phase == 1 && r->restore_stride < sf)
            return fail(ERTM_EINVAL, "run: restore_stride %lld below state size %lld of prop %d",
                        (long long)r->restore_stride, (long long)sf, k);
        if (r->phase != 2 && r->ck_interval > 0 && r->save_stride < sf)
            return fail(ERTM_EINVAL, "run: save_stride %lld below state size %lld of prop %d",
                        (long long)r->save_stride, (long long)sf, k);
    }
    if (r->op_after > ERTM_IMG_EXC_APPLY) return fail(ERTM_EINVAL, "run: bad image op");
    if ((r->op_after == ERTM_IMG_EXTRACT || r->op_after == ERTM_IMG_CORRELATE) &&
        r->t_end > r->t_begin) {
        // every imaged step t writes or reads replay slot t - t0
        for (int k = 0; k < (int)b->img.size(); ++k) {
            const ErtmImage& I = b->img[k];
            if (I.rbuf && (r->t_begin < I.t0 || (int64_t)r->t_end - I.t0 > I.nslots))
                return fail(ERTM_EINVAL,
                            "run: steps [%d, %d) of prop %d outside its replay slots [%d, %d)",
                            r->t_begin, r->t_end, k, I.t0, I.t0 + I.nslots);
        }
    }
    if (r->graph_steps < 0) return fail(ERTM_EINVAL, "run: graph_steps must be >= 0");
    std::vector<float> amp(n, 0.0f);
    std::vector<int32_t> tt(n, 0);
    int cur = r->cur0 & 1;
    int step_index = r->step_index0;
    int rc = ERTM_OK;
    if (r->phase == 1 && (rc = copy_states(b, cur, r->restore_arenas, r->restore_slot,
                                           r->restore_stride, false, s)) != ERTM_OK)
        return rc;
    const int nsteps = r->t_end - r->t_begin;
    // backward correlation through the receiver-field ring.  With the
    // viscoacoustic stabiliser the ring holds the step's pressure before the
    // stabiliser; its factor per slot (the clip scale, or 1 when the low-pass
    // has rewritten the slot) makes the flushed value the stabilised one.
    const bool use_ring = r->phase == 2 && r->op_after == ERTM_IMG_CORRELATE && ring_ok(b);
    const int ring_len = use_ring ? b->img[0].ring_len : 0;
    // folds beside the steps (eager loops without the stabiliser, a ring of
    // at least two fold periods), else every ring_len steps on s
    const bool beside = use_ring && !r->stabilize && r->graph_steps == 0 && b->fls &&
                        b->flush_period > 0 && ring_len >= 2 * b->flush_period;
    const int fold = beside ? b->flush_period : ring_len;
    b->fl_count = 0;
    float* rscale = nullptr;
    if (use_ring && r->stabilize) {
        const size_t bytes = sizeof(float) * (size_t)ring_len * n;
        if (b->rscale_bytes < bytes) {
            // flushes already queued with this batch may still read the old rows
            if (b->rscale)
                for (hipStream_t t : b->streams) HIPCHK(hipStreamSynchronize(t));
            pool_put(b->rscale, b->rscale_bytes);
            b->rscale = nullptr;
            b->rscale_bytes = 0;
            if (!(b->rscale = (float*)pool_get(bytes)))
                return fail(ERTM_EHIP, "run: ring scale allocation failed");
            b->rscale_bytes = bytes;
        }
        rscale = b->rscale;
    }
    const int t0 = r->phase == 2 ? r->t_end - 1 : r->t_begin, dir = r->phase == 2 ? -1 : 1;
    const int fop = use_ring ? ERTM_IMG_RING_STORE
                             : ((!r->stabilize && r->op_after >= 0) ? r->op_after : -1);
    std::vector<int32_t> t_hi(n, 0);
    int pending = 0;
    int64_t max_sf = 0;
    for (int k = 0; k < n; ++k) max_sf = std::max<int64_t>(max_sf, b->props[k].state_floats);

    // One time step.  dc == nullptr: eager launches with host time and
    // amplitude; dc set: launches recorded into a captured loop, which read
    // both from the device clock, save checkpoints through k_checkpoint and
    // end by advancing the clock.
    auto one_step = [&](int q, const DevClock* dc, hipStream_t s) -> int {
        const int t = t0 + dir * q;
        for (int k = 0; k < n; ++k) {
            tt[k] = t;
            amp[k] = r->phase == 2 ? 0.0f : r->wavelets[k][t];
        }
        int rc2;
        if (r->phase == 2 && (rc2 = do_inject(b, cur, tt.data(), s, dc)) != ERTM_OK) return rc2;
        // imaging of the new state inside the step when nothing runs between
        // them that changes the state (the viscoacoustic stabiliser)
        bool imaged = false;
        if ((rc2 = do_step(b, cur, amp.data(), s, fop, tt.data(), &imaged, dc)) != ERTM_OK)
            return rc2;
        if (use_ring && !imaged) return fail(ERTM_ESTATE, "run: ring imaging launch refused");
        cur ^= 1;
        if (r->stabilize) {
            const int rslot = use_ring ? t % ring_len : -1;
            if ((rc2 = do_stabilize(b, cur, step_index, r->cap, r->lp_interval, r->ctl, s,
                                    rscale ? rscale + (size_t)rslot * n : nullptr, rslot)) !=
                ERTM_OK)
                return rc2;
        }
        // the flush comes after the stabiliser, whose factors it reads
        if (use_ring) {
            if (pending == 0) t_hi.assign(tt.begin(), tt.end());
            if (++pending == fold || (!dc && q == nsteps - 1)) {
                rc2 = beside ? flush_beside(b, t_hi.data(), pending, s, rscale)
                             : do_flush(b, t_hi.data(), pending, s, dc, nullptr, rscale);
                if (rc2 != ERTM_OK) return rc2;
                pending = 0;
            }
        }
        ++step_index;
        if (r->phase != 2) {
            if (r->do_sample && (rc2 = do_sample(b, cur, tt.data(), s, dc)) != ERTM_OK) return rc2;
            if (r->ck_interval > 0 && dc) {
                for (int base = 0; base < n; base += ERTM_MAX_LAUNCH) {
                    LaunchArgs la{};
                    la.base = base;
                    la.n = std::min(ERTM_MAX_LAUNCH, n - base);
                    la.cur = cur;
                    use_clock(la, dc);
                    launch_checkpoint(b->dprops, la, b->dark, r->save_stride, r->ck_origin,
                                      r->ck_interval, max_sf, s);
                    LAUNCHCHK("checkpoint");
                }
            } else if (r->ck_interval > 0 && t >= r->ck_origin &&
                       (t - r->ck_origin) % r->ck_interval == 0) {
                const int64_t slot = (t - r->ck_origin) / r->ck_interval;
                if ((rc2 = copy_states(b, cur, r->save_arenas, slot, r->save_stride, true, s)) !=
                    ERTM_OK)
                    return rc2;
            }
        }
        if (r->op_after >= 0 && !imaged) {
            if ((rc2 = do_image(b, cur, tt.data(), r->op_after, s, dc)) != ERTM_OK) return rc2;
        }
        return ERTM_OK;
    };

    // Captured loop: chunks of S steps replayed as one graph.  S is even (the
    // chunk ends in the parity it began with) and a multiple of the ring
    // length (every chunk flushes its own ring); the rest runs eagerly.
    const int S = r->graph_steps;
    bool graph = S > 0 && S % 2 == 0 && nsteps >= S && !r->stabilize &&
                 (!use_ring || S % ring_len == 0);
    if (graph && r->phase != 2 && r->ck_interval > 0) {
        graph = r->save_stride % 4 == 0;  // k_checkpoint copies float4
        for (int k = 0; k < n && graph; ++k)
            graph = b->props[k].state_floats % 4 == 0 &&
                    ((uintptr_t)b->props[k].state[0] | (uintptr_t)b->props[k].state[1] |
                     (uintptr_t)r->save_arenas[k]) % 16 == 0;
    }
    int q = 0;
    if (graph) {
        const int nchunks = nsteps / S;
        if (!b->dclk && !(b->dclk = (int32_t*)pool_get(64)))
            return fail(ERTM_EHIP, "run: device clock allocation failed");
        if (!b->dark && !(b->dark = (float**)pool_get(sizeof(float*) * n)))
            return fail(ERTM_EHIP, "run: arena table allocation failed");
        const int32_t clk_h[4] = {0, t0, dir, 0};
        HIPCHK(hipMemcpyAsync(b->dclk, clk_h, sizeof(clk_h), hipMemcpyHostToDevice, s));
        std::vector<float> ta;
        if (r->phase != 2) {
            const size_t bytes = sizeof(float) * (size_t)nchunks * S * n;
            if (bytes > b->damp_bytes) {
                // the amplitude table is baked into the graphs: drop them,
                // after the launches already enqueued with this batch that
                // may still read the old table have finished
                for (hipStream_t t : b->streams) HIPCHK(hipStreamSynchronize(t));
                for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
                b->graphs.clear();
                pool_put(b->damp, b->damp_bytes);
                b->damp = (float*)pool_get(bytes);
                b->damp_bytes = b->damp ? pool_class(bytes) : 0;
                if (!b->damp) return fail(ERTM_EHIP, "run: amplitude table allocation failed");
            }
            ta.resize((size_t)nchunks * S * n);
            for (int qq = 0; qq < nchunks * S; ++qq)
                for (int k = 0; k < n; ++k) ta[(size_t)qq * n + k] = r->wavelets[k][t0 + dir * qq];
            HIPCHK(hipMemcpyAsync(b->damp, ta.data(), sizeof(float) * ta.size(),
                                  hipMemcpyHostToDevice, s));
        }
        if (r->phase != 2 && r->ck_interval > 0)
            HIPCHK(hipMemcpyAsync(b->dark, r->save_arenas, sizeof(float*) * n,
                                  hipMemcpyHostToDevice, s));
        const DevClock dc{b->dclk, r->phase == 2 ? nullptr : b->damp, n, 0};
        std::vector<int64_t> key = {r->phase, fop, use_ring, ring_len, r->do_sample, r->op_after,
                                    r->ck_interval, r->ck_origin, r->save_stride, S, cur,
                                    b->variant, b->zc, n};
        for (const ErtmImage& I : b->img)
            key.push_back((int64_t)I.separation | (int64_t)I.swave << 2 | (int64_t)I.nrw << 4 |
                          (int64_t)I.rwave[0] << 8 | (int64_t)I.rwave[1] << 10 |
                          (int64_t)I.illum << 12 | (int64_t)(I.rring != nullptr) << 16 |
                          (int64_t)(I.rbuf != nullptr) << 17 | (int64_t)I.ring_len << 20);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            if (b->graphs.size() >= 32) {  // keep the cache small
                for (hipStream_t t : b->streams) HIPCHK(hipStreamSynchronize(t));
                for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
                b->graphs.clear();
            }
            // recorded on a private stream (the caller's may be the null
            // stream, which cannot be captured), replayed on the caller's
            if (!b->cap) HIPCHK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
            // one capture at a time in the process
            static std::mutex capture_m;
            std::lock_guard<std::mutex> cap_lk(capture_m);
            const int cur_save = cur, si_save = step_index;
            HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
            for (int qq = 0; qq < S && rc == ERTM_OK; ++qq) {
                DevClock dq = dc;
                dq.q = qq;
                rc = one_step(qq, &dq, b->cap);
            }
            if (rc == ERTM_OK) {  // the next replay starts S steps later
                launch_tick(b->dclk, S, b->cap);
                if (hipGetLastError() != hipSuccess) rc = fail(ERTM_EHIP, "run: tick launch failed");
            }
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(b->cap, &g);
            if (rc != ERTM_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return fail(ERTM_EHIP, "run: capture failed: %s", hipGetErrorString(ec));
            hipGraphExec_t ex = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess)
                return fail(ERTM_EHIP, "run: graph instantiation failed: %s", hipGetErrorString(ei));
            it = b->graphs.emplace(key, ex).first;
            cur = cur_save;  // capture recorded the steps; nothing has run yet
            step_index = si_save;
            pending = 0;
        }
        for (int c = 0; c < nchunks; ++c) HIPCHK(hipGraphLaunch(it->second, s));
        q = nchunks * S;
        step_index += q;  // S is even: the parity is back where it started
    }
    // Two steps per launch where the pass allows it: forward or replay
    // phase, no sampling or stabiliser, imaging none or the fused
    // extraction; the pair never ends between a step and its checkpoint.
    const bool pairs = b->v2 != 0 && r->phase != 2 && !r->do_sample && !r->stabilize &&
                       (r->op_after < 0 ||
                        (r->op_after == ERTM_IMG_EXTRACT && fused_image_ok(b, ERTM_IMG_EXTRACT)));
    auto ck_at = [&](int t) {
        return r->ck_interval > 0 && t >= r->ck_origin && (t - r->ck_origin) % r->ck_interval == 0;
    };
    std::vector<float> amp2(n, 0.0f);
    while (q < nsteps) {
        const int t = t0 + dir * q;
        if (!pairs || q + 1 >= nsteps || ck_at(t)) {
            if ((rc = one_step(q, nullptr, s)) != ERTM_OK) return rc;
            ++q;
            continue;
        }
        for (int k = 0; k < n; ++k) {
            tt[k] = t;
            amp[k] = r->wavelets[k][t];
            amp2[k] = r->wavelets[k][t + 1];
        }
        if ((rc = do_step2(b, cur, amp.data(), amp2.data(), s, r->op_after, tt.data())) != ERTM_OK)
            return rc;
        cur ^= 1;
        step_index += 2;
        if (ck_at(t + 1)) {
            const int64_t slot = (t + 1 - r->ck_origin) / r->ck_interval;
            if ((rc = copy_states(b, cur, r->save_arenas, slot, r->save_stride, true, s)) != ERTM_OK)
                return rc;
        }
        q += 2;
    }
    // what s runs next (a replay overwriting the slots, reads of the
    // accumulators) comes after the last fold
    if (beside && b->fl_count > 0) HIPCHK(hipStreamWaitEvent(s, b->fl_done[(b->fl_count - 1) & 1], 0));
    if (cur_out) *cur_out = cur;
    return ERTM_OK;
}

// Lanes (ertm_run_lanes): every lane's current segment advances one step per
// tick; per tick the launches group the lanes by pass kind.

// Compute units of the current device (cached per device).
static int cu_count() {
    static std::mutex m;
    static std::map<int, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lk(m);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
        n = 256;
    cache[dev] = n;
    return n;
}

// Launch shape of one lane launch group.  A fused block streams the rows of
// its chunk one after another, so a launch of few blocks lasts as long as
// one chunk's chain of row loads: lanes take the longest chunks that still
// give 16 blocks per CU (the batch tuner's rule), over the blocks their
// grids really have; 8-row chunks while they give 4 blocks per CU; below
// that the two-launch kernels (variant 0), whose 64 x 4 tiles have no chain
// (a 208 x 529 elastic step: 38 us fused with 8-row chunks, 30 us as two
// launches and an imaging pass).  Every shape gives bit-identical results.
static void lane_shape(ertm_batch_t b, const Members& m, int* shape) {
    shape[0] = b->variant;
    shape[1] = b->zc;
    if (b->variant == 0) return;
    auto blocks = [&](int threads, int zc) {
        const int tx = threads - 16;
        int64_t nb = 0;
        for (int i = 0; i < m.n; ++i) {
            const ErtmProp& P = b->props[m.idx[i]];
            nb += (int64_t)((P.nxp + tx - 1) / tx) * ((P.nzp + zc - 1) / zc);
        }
        return nb;
    };
    const int64_t target = 16 * (int64_t)cu_count();
    if (blocks(192, 48) >= target) {
        shape[0] = 4, shape[1] = 48;
        return;
    }
    for (int zc : {48, 40, 32, 24, 16, 12}) {
        if (blocks(128, zc) >= target) {
            shape[0] = 2, shape[1] = zc;
            return;
        }
    }
    if (blocks(128, 8) >= target / 4) {
        shape[0] = 2, shape[1] = 8;
        return;
    }
    shape[0] = 0;
}
static int check_seg(ertm_batch_t b, const ErtmSeg& g, int i) {
    const int n = (int)b->props.size();
    if (g.prop < 0 || g.prop >= n) return fail(ERTM_EINVAL, "lanes: segment %d: bad propagator %d", i, g.prop);
    if (g.phase < 0 || g.phase > 2) return fail(ERTM_EINVAL, "lanes: segment %d: bad phase %d", i, g.phase);
    if (g.t_begin < 0 || g.t_end < g.t_begin)
        return fail(ERTM_EINVAL, "lanes: segment %d: bad time range", i);
    if (g.done < 0 || g.done > g.t_end - g.t_begin)
        return fail(ERTM_EINVAL, "lanes: segment %d: done %d outside its range", i, g.done);
    const ErtmProp& P = b->props[g.prop];
    const ErtmImage& I = b->img[g.prop];
    if (g.phase != 2 && !g.wavelet) return fail(ERTM_EINVAL, "lanes: segment %d: wavelet missing", i);
    if (g.phase == 1 && !g.restore) return fail(ERTM_EINVAL, "lanes: segment %d: replay without restore", i);
    if (g.ck_interval < 0 || (g.phase == 2 && g.ck_interval != 0))
        return fail(ERTM_EINVAL, "lanes: segment %d: bad checkpoint interval", i);
    if (g.ck_interval > 0 && (!g.save_arena || g.save_stride < P.state_floats))
        return fail(ERTM_EINVAL, "lanes: segment %d: checkpoint arena missing or stride below the state", i);
    const int want = g.phase == 2 ? ERTM_IMG_CORRELATE : ERTM_IMG_EXTRACT;
    if (g.op_after != -1 && g.op_after != want)
        return fail(ERTM_EINVAL, "lanes: segment %d: image op %d not allowed in phase %d", i, g.op_after,
                    g.phase);
    if (g.op_after >= 0) {
        if (!I.rbuf || g.img_t0 > g.t_begin || (int64_t)g.t_end - g.img_t0 > I.nslots)
            return fail(ERTM_EINVAL, "lanes: segment %d: replay slots [%d, %d) outside the %d slots",
                        i, g.t_begin - g.img_t0, g.t_end - g.img_t0, I.nslots);
        if (g.op_after == ERTM_IMG_CORRELATE && (!I.num[0] || I.nrw < 1 || I.nrw > 2))
            return fail(ERTM_EINVAL, "lanes: segment %d: no accumulators", i);
    }
    return ERTM_OK;
}

int ertm_run_lanes(ertm_batch_t b, ErtmSeg* segs, int32_t nseg, int32_t* cur, int32_t max_ticks,
                   int32_t stop_at_lane_end, int64_t* stats3, void* stream) {
    if (!b || nseg < 0 || (nseg > 0 && !segs) || !cur) return fail(ERTM_EINVAL, "lanes: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    b->note(s);
    const int n = (int)b->props.size();
    for (int k = 0; k < n; ++k)
        if (cur[k] != 0 && cur[k] != 1) return fail(ERTM_EINVAL, "lanes: parity of prop %d not 0/1", k);
    // lanes in order of first appearance; each propagator in one lane
    std::vector<int> lane_of(n, INT32_MIN);
    std::vector<std::vector<int>> lanes;
    std::map<int, size_t> lane_ix;
    std::vector<int> ring_members;
    for (int i = 0; i < nseg; ++i) {
        const ErtmSeg& g = segs[i];
        if (int rc = check_seg(b, g, i)) return rc;
        if (lane_of[g.prop] != INT32_MIN && lane_of[g.prop] != g.lane)
            return fail(ERTM_EINVAL, "lanes: propagator %d is in lanes %d and %d", g.prop,
                        lane_of[g.prop], g.lane);
        lane_of[g.prop] = g.lane;
        auto it = lane_ix.find(g.lane);
        if (it == lane_ix.end()) {
            it = lane_ix.emplace(g.lane, lanes.size()).first;
            lanes.emplace_back();
        }
        lanes[it->second].push_back(i);
        if (g.phase == 2 && g.op_after == ERTM_IMG_CORRELATE &&
            std::find(ring_members.begin(), ring_members.end(), g.prop) == ring_members.end())
            ring_members.push_back(g.prop);
    }
    // the ring is used when every backward lane has one, in the ticks whose
    // backward launch is fused
    const Members all_back{ring_members.data(), (int)ring_members.size(), cur, nullptr};
    const bool use_ring = !ring_members.empty() && ring_ok(b, &all_back);
    const int ring_len = use_ring ? b->img[ring_members[0]].ring_len : 0;
    auto len_of = [&](int i) { return segs[i].t_end - segs[i].t_begin; };
    std::vector<size_t> pos(lanes.size(), 0);
    auto skip_done = [&](size_t l) {
        while (pos[l] < lanes[l].size() && segs[lanes[l][pos[l]]].done >= len_of(lanes[l][pos[l]]))
            ++pos[l];
    };
    for (size_t l = 0; l < lanes.size(); ++l) skip_done(l);

    std::vector<int32_t> tt(n, 0), t0v(n, 0), pend(n, 0), thi(n, 0);
    std::vector<float> amp(n, 0.0f);
    int rc = ERTM_OK;
    auto flush = [&](std::vector<int>& due) -> int {
        // one launch per distinct count of pending steps
        std::sort(due.begin(), due.end(), [&](int a, int c) { return pend[a] < pend[c]; });
        for (size_t i = 0; i < due.size();) {
            size_t j = i;
            while (j < due.size() && pend[due[j]] == pend[due[i]]) ++j;
            std::vector<int> grp(due.begin() + i, due.begin() + j);
            std::sort(grp.begin(), grp.end());
            const Members m{grp.data(), (int)grp.size(), cur, t0v.data()};
            if (int rc2 = do_flush(b, thi.data(), pend[grp[0]], s, nullptr, &m)) return rc2;
            for (int p : grp) pend[p] = 0;
            i = j;
        }
        due.clear();
        return ERTM_OK;
    };
    int ticks = 0;
    int64_t lane_steps = 0, step_groups = 0;
    bool stop = false;
    std::vector<int> act, inj, fplain, fext, bimg, bplain, due;
    std::vector<CopyItem> copies;
    while (!stop && (max_ticks <= 0 || ticks < max_ticks)) {
        act.clear();
        for (size_t l = 0; l < lanes.size(); ++l)
            if (pos[l] < lanes[l].size()) act.push_back(lanes[l][pos[l]]);
        if (act.empty()) break;
        inj.clear(), fplain.clear(), fext.clear(), bimg.clear(), bplain.clear(), copies.clear();
        for (int i : act) {
            const ErtmSeg& g = segs[i];
            const int p = g.prop;
            const int t = g.phase == 2 ? g.t_end - 1 - g.done : g.t_begin + g.done;
            tt[p] = t;
            amp[p] = g.phase == 2 ? 0.0f : g.wavelet[t];
            t0v[p] = g.img_t0;
            if (g.phase == 1 && g.done == 0)
                copies.push_back({g.restore, b->props[p].state[cur[p]], b->props[p].state_floats, p});
            if (g.phase == 2) {
                inj.push_back(p);
                (g.op_after >= 0 ? bimg : bplain).push_back(p);
            } else {
                (g.op_after >= 0 ? fext : fplain).push_back(p);
            }
        }
        for (auto* v : {&inj, &fplain, &fext, &bimg, &bplain}) std::sort(v->begin(), v->end());
        if (!copies.empty() && (rc = copy_list(b, copies, s)) != ERTM_OK) return rc;
        auto mem = [&](const std::vector<int>& v) {
            return Members{v.data(), (int)v.size(), cur, t0v.data()};
        };
        if (!inj.empty()) {
            const Members m = mem(inj);
            if ((rc = do_inject(b, 0, tt.data(), s, nullptr, &m)) != ERTM_OK) return rc;
        }
        bool ext_done = false, back_done = false;
        int shape[2];
        if (!fplain.empty()) {
            const Members m = mem(fplain);
            lane_shape(b, m, shape);
            if ((rc = do_step(b, 0, amp.data(), s, -1, nullptr, nullptr, nullptr, &m, shape)) != ERTM_OK)
                return rc;
        }
        // extracting replays: fused extraction, or a plain step and k_image
        if (!fext.empty()) {
            const Members m = mem(fext);
            lane_shape(b, m, shape);
            if ((rc = do_step(b, 0, amp.data(), s, shape[0] ? ERTM_IMG_EXTRACT : -1, tt.data(),
                              &ext_done, nullptr, &m, shape)) != ERTM_OK)
                return rc;
        }
        // backward passes: ring stores (folded later by k_image_flush), fused
        // correlation, or a plain step and k_image.  Per cell the ring folds
        // steps in the order direct correlation adds them, so a propagator may
        // switch between the two after folding its pending ring steps.
        bool ring_tick = false;
        if (!bimg.empty()) {
            const Members m = mem(bimg);
            lane_shape(b, m, shape);
            ring_tick = use_ring && shape[0] != 0;
            if (use_ring && !ring_tick) {
                for (int p : bimg)
                    if (pend[p] > 0) due.push_back(p);
                if (!due.empty() && (rc = flush(due)) != ERTM_OK) return rc;
            }
            const int op = ring_tick ? ERTM_IMG_RING_STORE : (shape[0] ? ERTM_IMG_CORRELATE : -1);
            if ((rc = do_step(b, 0, amp.data(), s, op, tt.data(), &back_done, nullptr, &m, shape)) !=
                ERTM_OK)
                return rc;
            if (ring_tick && !back_done) return fail(ERTM_ESTATE, "lanes: ring imaging launch refused");
        }
        if (!bplain.empty()) {
            const Members m = mem(bplain);
            lane_shape(b, m, shape);
            if ((rc = do_step(b, 0, amp.data(), s, -1, nullptr, nullptr, nullptr, &m, shape)) != ERTM_OK)
                return rc;
        }
        for (int i : act) cur[segs[i].prop] ^= 1;
        lane_steps += (int64_t)act.size();
        for (auto* v : {&fplain, &fext, &bimg, &bplain}) step_groups += v->empty() ? 0 : 1;
        if (ring_tick) {
            for (int i : act) {
                const ErtmSeg& g = segs[i];
                if (g.phase != 2 || g.op_after < 0) continue;
                if (pend[g.prop] == 0) thi[g.prop] = tt[g.prop];
                if (++pend[g.prop] == ring_len || g.done + 1 == len_of(i)) due.push_back(g.prop);
            }
            if (!due.empty() && (rc = flush(due)) != ERTM_OK) return rc;
        }
        copies.clear();
        for (int i : act) {
            const ErtmSeg& g = segs[i];
            const int t = tt[g.prop];
            if (g.phase == 2 || g.ck_interval <= 0 || t < g.ck_origin ||
                (t - g.ck_origin) % g.ck_interval != 0)
                continue;
            const ErtmProp& P = b->props[g.prop];
            copies.push_back({P.state[cur[g.prop]],
                              g.save_arena + (int64_t)((t - g.ck_origin) / g.ck_interval) * g.save_stride,
                              P.state_floats, g.prop});
        }
        if (!copies.empty() && (rc = copy_list(b, copies, s)) != ERTM_OK) return rc;
        if (!fext.empty() && !ext_done) {
            const Members m = mem(fext);
            if ((rc = do_image(b, 0, tt.data(), ERTM_IMG_EXTRACT, s, nullptr, &m)) != ERTM_OK) return rc;
        }
        if (!bimg.empty() && !back_done) {
            const Members m = mem(bimg);
            if ((rc = do_image(b, 0, tt.data(), ERTM_IMG_CORRELATE, s, nullptr, &m)) != ERTM_OK) return rc;
        }
        for (size_t l = 0; l < lanes.size(); ++l) {
            if (pos[l] >= lanes[l].size()) continue;
            ++segs[lanes[l][pos[l]]].done;
            skip_done(l);
            if (pos[l] >= lanes[l].size() && stop_at_lane_end) stop = true;
        }
        ++ticks;
    }
    // a call that stops inside a backward segment folds its ring steps now
    for (int p : ring_members)
        if (pend[p] > 0) due.push_back(p);
    if (!due.empty() && (rc = flush(due)) != ERTM_OK) return rc;
    if (stats3) {
        stats3[0] = ticks;
        stats3[1] = lane_steps;
        stats3[2] = step_groups;
    }
    return ERTM_OK;
}

int ertm_derivative(const void* f, void* out, int32_t n0, int32_t n1, int32_t axis, int32_t plus,
                    int32_t order, double h, int32_t is_f64, void* stream) {
    if (!f || !out || n0 <= 0 || n1 <= 0 || (axis != 0 && axis != 1))
        return fail(ERTM_EINVAL, "derivative: bad arguments");
    static const double C[4][4] = {{1.0, 0, 0, 0},
                                   {9.0 / 8.0, -1.0 / 24.0, 0, 0},
                                   {75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0},
                                   {1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0}};
    if (order != 2 && order != 4 && order != 6 && order != 8)
        return fail(ERTM_EINVAL, "derivative: order %d", order);
    const int K = order / 2;
    if ((axis == 0 ? n0 : n1) <= 2 * K) return fail(ERTM_EINVAL, "derivative: axis too short");
    if (!(h > 0.0)) return fail(ERTM_EINVAL, "derivative: spacing must be positive");
    DerivCoef c;
    for (int k = 0; k < 4; ++k) {
        c.c64[k] = C[K - 1][k];
        c.c[k] = (float)C[K - 1][k];
    }
    const double hh = is_f64 ? h : (double)(float)h;
    launch_derivative(f, out, n0, n1, axis, plus, order, c, hh, is_f64, (hipStream_t)stream);
    LAUNCHCHK("derivative");
    return ERTM_OK;
}

int ertm_absmax(const float* x, int64_t n, float* out, void* stream) {
    if (!x || !out || n <= 0) return fail(ERTM_EINVAL, "absmax: bad arguments");
    launch_absmax(x, n, out, (hipStream_t)stream);
    LAUNCHCHK("absmax");
    return ERTM_OK;
}

int ertm_sync(void* stream) {
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return ERTM_OK;
}

}  // extern "C"
