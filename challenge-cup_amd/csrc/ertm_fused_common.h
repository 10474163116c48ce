// Device helpers shared by the fused step kernels (ertm_fused.hip: one step
// per launch; ertm_fused2.hip: two steps per launch): XCD-aware block order,
// buffer-resource loads and stores, imaging kinds, and the per-block
// constants of the elastic step.
#pragma once

#include "ertm_kernels.h"

namespace ertm {

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int xcd = bid & 7, idx = bid >> 3;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Buffer resources: a 128-bit descriptor per array with the byte size as
// range.  Buffer loads outside [0, size) return 0 and such stores are
// dropped, which covers rows above/below the grid without guards; columns
// beyond the grid only ever feed lanes whose results are discarded.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                             0x00020000);
}
// The b32 builtins move raw 32-bit patterns: bit-cast, never convert.
// State stores carry the non-temporal hint (cache-policy bit 1, "nt"): what a
// step writes is read back only by the next launch, after far more data than
// the caches hold has passed, so its lines should not displace the medium
// and the halo rows that blocks of the same launch re-read (measured: 630 us
// against 649 us per 64-shot elastic step).  Image stores (replay slots,
// receiver ring) keep the default policy: their rows are not aligned to
// cache lines, and a line that two blocks write in part is merged in L2
// before it goes to HBM (non-temporal, the replay and backward passes of a
// 64-shot migration took 6.50 s instead of 6.31 s).  The CPML memories'
// x-strip rows are 2 x 80-byte pieces, so they keep the default policy too.
#define ERTM_CPOL_NT 2
__device__ __forceinline__ float bload(rsrc_t r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, off, 0,
                                          ERTM_CPOL_NT);
}
__device__ __forceinline__ void bstore_l2(rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, off, 0, 0);
}

__device__ __forceinline__ double bload64(rsrc_t r, int off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void bstore64(rsrc_t r, int off, double v) {
    using u64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u64, v), r, off, 0, 0);
}

// Imaging inside the step (IMG template argument): 0 = none; 1 = extract the
// source-side field of the new state into the replay slot; 2 = correlate it
// with the receiver-side fields of the new state (ertm/imaging.py:279-324).
// Same values and float64 accumulation order as k_image: divcurl P = Dx-vx +
// Dz-vz and S = Dz-vx - Dx-vz from the plain derivatives (no CPML memory),
// stress P = (sxx + szz)/2 and S = sxz of the stored new stresses.
#define ERTM_FIMG_NONE 0
#define ERTM_FIMG_EXTRACT 1
#define ERTM_FIMG_CORRELATE 2
#define ERTM_FIMG_RING 3  // store the receiver-side fields into ring slot t mod R
// Variants of EXTRACT and RING that never need the divcurl S field (Dz-vx -
// Dx-vz, two derivatives the step does not compute otherwise): P-wave or
// stress-separation imaging.  Dropping that code path lowers register use.
#define ERTM_FIMG_EXTRACT_NOS 4
#define ERTM_FIMG_RING_NOS 5

// The imaging operation of an IMG value, and whether it may need the S field.
template <int IMG>
constexpr int img_op() {
    return IMG == ERTM_FIMG_EXTRACT_NOS ? ERTM_FIMG_EXTRACT
         : IMG == ERTM_FIMG_RING_NOS ? ERTM_FIMG_RING : IMG;
}
template <int IMG>
constexpr bool img_may_s() {
    return IMG != ERTM_FIMG_NONE && IMG != ERTM_FIMG_EXTRACT_NOS && IMG != ERTM_FIMG_RING_NOS;
}

// Per-block constants of the fused elastic step.
struct ElBlock {
    rsrc_t rS, rT, rIR, rLAM, rMU;
    rsrc_t rMI;         // medium palette indices (one byte per cell; ErtmProp.midx)
    const float* ptab;  // the palette in LDS: 1/rho, mu, lambda at 0, 256, 512
    gcfloat* AZ;
    gcfloat* BZ;
    int nzp, nxp, z0, z1, c, c4, c4own, tid, ld4;
    int fVX, fVZ, fXX, fZZ, fXZ;
    int px0, px1, px2, px3, pz0, pz1, pz2, pz3;
    int wx, wzt, wzb, xrow4, jx4;
    int src_iz, mech, fs_row;
    bool colv, velc, own, inx, src_col;
    float dt, hx, hz, amp, axc, bxc;
    double rhx, rhz;
    float cf[4];
    // imaging (IMG != 0): replay slot and accumulators of this propagator
    rsrc_t rR, rA;
    int pad, nzph, nxph, jcol;          // jcol: physical column of the lane, -1 outside
    int o_num0, o_num1, o_ill, o_rill0, o_rill1;  // byte offsets in rA
    int illum, nrw, rw0, rw1, swave, sep;
    bool cross, need_s;
};

// Byte offset that every buffer range check rejects: stores to it are dropped.
#define ERTM_NOWRITE (-4)

// The row loop.  STRIP = false is the interior version, used by blocks that
// touch no CPML strip: its body is straight-line (every lane issues the same
// loads and stores; stores that must not happen go out of range), so the
// number of memory operations per row is fixed and the compiler's wait for
// the prefetched loads leaves the previous row's stores in flight (vmcnt
// also counts stores on this architecture).  STRIP = true adds the CPML
// memory reads/writes, inside per-lane branches.
//
// Addressing: a lane's byte offset is its column (c4) plus a row part that
// is uniform across the block (row * ld4 + field slab offset), so only c4 and
// c4own live in vector registers.  One descriptor covers a whole state
// buffer (fields + CPML strips): reads of rows above the grid wrap to huge
// offsets and return 0, reads between slabs return finite values that only
// feed results that are discarded.  A store that must not happen gets
// ERTM_BAD added (per lane through c4own, per row through the row part),
// which puts it beyond every buffer's range (the launcher keeps state
// buffers below 1 GiB); the hardware drops it.
#define ERTM_BAD 0x40000000

}  // namespace ertm
