// traverse.h - the per-ray brickmap DDA on the device.
//
// One lane owns one ray.  The arithmetic is the reference CPU cast
// (src/VoxelRT/CpuRenderer.cpp:135-224) per lane: FP32, IEEE 1/d, explicit
// fmaf for sideDist and the next position, x86 min and floor2i semantics,
// +0.001 bias, 2-level LOD from the 64-bit masks (brick level 8/16/32 voxels,
// voxel level 1/2/4).  Storage is the popcount-compacted GPU layout
// (GpuRenderer.cpp:27-33, VoxelMap.glsl:45-64) with the sector's mask and
// base slot packed in one 16-byte record, so a lookup is one load.
//
// Two instantiations of the step compute identical traversals:
//   EXACT = true   floor2i with x86 cvtps2dq semantics (NaN / out of range ->
//                  INT_MIN) and x86 minps (a < b ? a : b) in the loop;
//   EXACT = false  plain floor + convert and v_min3_f32, valid for "tame" rays
//                  (finite 1/d and tStart, |o| < 2^20, 0.25 <= |d|^2 <= 4):
//                  their positions stay far inside int32 range for any
//                  iteration cap (a step moves at most 32*sqrt(3)/|d| + 0.001
//                  along d, i.e. < 56 voxels, and 65535 * 56 + 2^20 < 2^31),
//                  and their side distances are never NaN, so floor2i never
//                  takes its INT_MIN branch and the two minimums agree up to
//                  the sign of a zero, which cannot change tmin + 0.001.
// The reported t is recomputed with the exact x86 minimum in both cases.
//
// A third template parameter, ORIGIN, specializes the step for the rays' integer
// world origin W (uniform over a launch or a batch segment):
//   ORIGIN_ANY      any W: voxel = W + floor(cur), rel = aligned voxel - W;
//   ORIGIN_ALIGNED  W a multiple of 32 on every axis (a sector corner): the
//                   largest cell a step aligns to is 32 voxels (mask 31), so
//                   aligning W + f and subtracting W gives the same bits as
//                   aligning f itself; the step skips the three subtractions;
//   ORIGIN_ZERO     W = 0: voxel = floor(cur), so the additions go as well.
// Every form computes the same bits as ORIGIN_ANY.
//
// ONE_EXIT (the fast loop): the lookup is made before the bounds test, so a step
// leaves the loop at one place for both ends of a ray.  A lane that has left the
// world reads whatever record its index names (an index past the records reads 0:
// the structured buffer load checks the range) and does not descend, so it never
// reads an occupancy mask.  The loop keeps no hit flag: ray_trace tells the exits
// apart after the loop (by the step count, or by one lookup at any origin).  Each
// exit and flag costs lane-mask bookkeeping on the scalar unit, which one CU's four
// SIMDs share: the one-exit step issues 18 scalar ALU instructions and 5 branches (one of
// them skips the sector load when no lane needs it).
#pragma once

#include <hip/hip_runtime.h>

#include "vrt_internal.h"
#include "vrt_math.h"

namespace vrt {

struct RayState {
    float o[3], d[3], inv[3], ts[3], side[3], cur[3];
    int32_t W[3];
    int32_t dm[3];  // all ones where d >= 0: the cell corner to step to is the far one
    uint32_t iters, descents;
    // the sector record of the last lookup and its index (~0u before the first): a step
    // whose position is still in that sector reuses it instead of loading it again
    uint32_t rec_si;
    uint32_t rec[4];
};

enum StepResult : int { STEP_CONTINUE = 0, STEP_HIT = 1, STEP_MISS = 2 };

__device__ __forceinline__ void ray_setup(RayState& r, float ox, float oy, float oz, float dx, float dy, float dz,
                                          int32_t wx, int32_t wy, int32_t wz) {
    r.o[0] = ox; r.o[1] = oy; r.o[2] = oz;
    r.d[0] = dx; r.d[1] = dy; r.d[2] = dz;
    r.W[0] = wx; r.W[1] = wy; r.W[2] = wz;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.inv[a] = 1.0f / r.d[a];
        r.ts[a] = ((r.d[a] < 0.0f ? 0.0f : 1.0f) - r.o[a]) * r.inv[a];
        r.side[a] = 0.0f;
        r.cur[a] = r.o[a];
        r.dm[a] = r.d[a] < 0.0f ? 0 : -1;
        // a per-lane word, so the cell alignment below is one bitfield insert per axis
        asm volatile("" : "+v"(r.dm[a]));
    }
    r.iters = 0;
    r.descents = 0;
    // no index reaches ~0u (worlds hold at most 2^27 sectors), and an index past the
    // records reads 0, like this initial record
    r.rec_si = ~0u;
    r.rec[0] = r.rec[1] = r.rec[2] = r.rec[3] = 0;
}

// True when the fast step is exact for this ray (see the file comment).
__device__ __forceinline__ bool ray_is_tame(const RayState& r) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        ok = ok && fabsf(r.inv[a]) < INFINITY && fabsf(r.ts[a]) < INFINITY && fabsf(r.o[a]) < 1048576.0f;
    }
    float len2 = r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2];
    return ok && len2 >= 0.25f && len2 <= 4.0f;
}

// (mask & a) | (~mask & b): V_BFI_B32.
__device__ __forceinline__ uint32_t bitfield_insert(uint32_t mask, uint32_t a, uint32_t b) {
    return (mask & a) | (~mask & b);
}

// Structured buffer load (BUFFER_LOAD_DWORDX4 ... IDXEN): record vindex of the
// resource's array, the byte address index * stride computed by the load itself, and
// 0 for an index at or past the resource's record count (range-checked in hardware).
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
__device__ i32x4 vrt_struct_buffer_load_v4i32(i32x4 rsrc, int32_t vindex, int32_t voffset, int32_t soffset,
                                              int32_t aux) __asm("llvm.amdgcn.struct.buffer.load.v4i32");

// Buffer resource of the sector records: base address, 16-byte stride, one record per
// sector.  Word 3 is the gfx9 untyped-buffer configuration.
__device__ __forceinline__ i32x4 sector_records_rsrc(const WorldView& w) {
    uint64_t a = (uint64_t)w.sec;
    i32x4 r;
    r.x = (int32_t)(uint32_t)a;
    r.y = (int32_t)(((uint32_t)(a >> 32) & 0xFFFFu) | (16u << 16));
    r.z = (int32_t)(1u << (2 * w.shift_xz + w.shift_y));
    r.w = 0x00020000;
    return r;
}

// (a << sh) | b in one V_LSHL_OR_B32.
__device__ __forceinline__ uint32_t lshl_or(uint32_t a, uint32_t sh, uint32_t b) {
    uint32_t r;
    asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(sh), "v"(b));
    return r;
}

// (M & a) | (~M & b) in one V_BFI_B32 for a constant field mask M (written out,
// because the compiler splits a constant-mask insert into ANDs once it knows which
// result bits are used).
template <uint32_t M>
__device__ __forceinline__ uint32_t insert_bits(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "i"(M), "v"(a), "v"(b));
    return r;
}

// ~((M & a) | (~M & b)) in one V_BITOP3_B32: the truth table is indexed by
// (M bit) * 4 + (a bit) * 2 + (b bit), so 0x35 is "not (M ? a : b)".
template <uint32_t M>
__device__ __forceinline__ uint32_t insert_bits_not(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x35" : "=v"(r) : "i"(M), "v"(a), "v"(b));
    return r;
}

// min(x, 1) in one V_MIN_U32 (the compiler would compare and select).
__device__ __forceinline__ uint32_t min_1(uint32_t x) {
    uint32_t r;
    asm("v_min_u32 %0, %1, 1" : "=v"(r) : "v"(x));
    return r;
}

// floor(x) as int32 in one V_CVT_FLR_I32_F32; x finite and |x| < 2^31.
__device__ __forceinline__ int32_t floor_i32_in_range(float x) {
    int32_t f;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(f) : "v"(x));
    return f;
}

// One iteration of RayCast (CpuRenderer.cpp:185-202) with GetStepPos (:135-171).
// On STEP_HIT, v is the hit voxel and *hit_slot its brick slot.
//
// Integer work is arranged for few VALU instructions per iteration:
//   mask bit test: below = m << (63 - bi) moves bit bi to bit 63 and drops the
//     bits above it, so the test is a sign test and popcount(below) counts the
//     bricks at or below bi, i.e. slot + 1 (the occupancy load subtracts the
//     one as a constant offset);
//   index fields: only bits 0..5 of the in-mask indices are meaningful; the
//     64-bit shifts read only those bits (the & 63 costs nothing), and the
//     fields are merged with bitfield inserts;
//   cell mask: cm = 2^s - 1 with s = level + (group empty) + (mask empty), the
//     reference's LOD of +1 / +2 (a 4x4x4 mask that is 0 has every group empty).
//     With ce = 2^(level + 1 + (mask empty)) - 1, the mask of an empty group, this
//     is cm = ce >> min(group bits, 1); the sector record holds ce for its mask;
//   alignment: (v & ~cm) | (dm & cm), one bitfield insert per axis.
// Occupancy indices slot * 8 + cell are 32-bit: worlds hold at most 2^28 slots.
enum OriginKind : int { ORIGIN_ANY = 0, ORIGIN_ALIGNED = 1, ORIGIN_ZERO = 2 };

// The specialization that a world origin allows (uniform over the caller's rays).
__device__ __forceinline__ int origin_kind(int32_t wx, int32_t wy, int32_t wz) {
    if ((wx | wy | wz) == 0) return ORIGIN_ZERO;
    return ((wx | wy | wz) & 31) == 0 ? ORIGIN_ALIGNED : ORIGIN_ANY;
}

template <bool EXACT, bool CUBIC = false, int ORIGIN = ORIGIN_ANY, bool ONE_EXIT = false>
__device__ __forceinline__ bool ray_advance(const WorldView& w, RayState& r, int32_t v[3], bool& hit,
                                            uint32_t* hit_slot) {
    int32_t f[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        f[a] = EXACT ? floor2i(r.cur[a]) : floor_i32_in_range(r.cur[a]);
        v[a] = ORIGIN == ORIGIN_ZERO ? f[a] : (int32_t)((uint32_t)r.W[a] + (uint32_t)f[a]);
    }
    // GetInboundMask (:114-117); one compare when the world is a cube
    bool inb = CUBIC ? (uint32_t)(v[0] | v[1] | v[2]) < w.size_xz
                     : ((uint32_t)(v[0] | v[2]) < w.size_xz && (uint32_t)v[1] < w.size_y);
    if (!ONE_EXIT) {
        if (!inb) {
            hit = false;
            return false;
        }
        r.iters++;
        // a per-lane count: left to itself the compiler derives it from a wave-uniform loop
        // counter with three copies per step
        asm volatile("" : "+v"(r.iters));
    }
    uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
    uint32_t si = lshl_or(uy >> 5, 2 * w.shift_xz, lshl_or(uz >> 5, w.shift_xz, ux >> 5));
    // one 16-byte load for the whole record: a descent then waits on one load (the
    // occupancy mask) instead of two (a structured buffer load: the hardware scales the
    // index, one instruction fewer than a 64-bit address).  A lane still in the sector of
    // its previous step keeps that record and issues no load: divergent waves (bounce and
    // incoherent rays) then send L1 fewer distinct lines per step, and a wave whose lanes
    // all stay skips the load and its wait.  The compare and the copies out of the kept
    // record cost 5 VALU instructions per step; measured, the coherent primary launch
    // takes them without slowing, incoherent sets cast 9-17% faster.
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    if (si != r.rec_si) {
        const u32x4 ld = (u32x4)vrt_struct_buffer_load_v4i32(sector_records_rsrc(w), (int32_t)si, 0, 0, 0);
        r.rec[0] = ld.x;
        r.rec[1] = ld.y;
        r.rec[2] = ld.z;
        r.rec[3] = ld.w;
        r.rec_si = si;
    }
    u32x4 sec;
    sec.x = r.rec[0];
    sec.y = r.rec[1];
    sec.z = r.rec[2];
    sec.w = r.rec[3];
    uint64_t m = (uint64_t)sec.x | (uint64_t)sec.y << 32;
    // brick within the sector: bits 3..4 of x, z, y at bits 0..1, 2..3, 4..5
    // nbi = ~bi: bits 0..5 give the shift 63 - bi, and ~nbi & 0x2A the 2x2x2 group's shift
    uint32_t nbi = insert_bits_not<0x30u>(uy << 1, insert_bits<0x0Cu>(uz >> 1, ux >> 3));
    uint64_t below = m << (nbi & 63u);
    bool descend = (int32_t)(below >> 32) < 0;
    if (ONE_EXIT) descend = descend && inb;
    // the cell mask of a step whose 2x2x2 group is empty: 15 below a sector mask, 31
    // when that mask is 0 (from the record); 1 below a voxel mask, 3 when it is 0
    uint32_t ce = sec.w;
    bool solid = false;
    uint32_t slot1 = 0;
    if (descend) {
        r.descents++;
        slot1 = sec.z + (uint32_t)__popcll(below);
        uint32_t cell = __builtin_amdgcn_ubfe(ux, 2, 1) | __builtin_amdgcn_ubfe(uz, 2, 1) << 1 |
                        __builtin_amdgcn_ubfe(uy, 2, 1) << 2;
        m = w.occ[(size_t)(slot1 * 8u + cell) - 8u];
        // voxel within the brick's 4x4x4 cell: bits 0..1 of x, z, y
        nbi = insert_bits_not<0x30u>(uy << 4, insert_bits<0x0Cu>(uz << 2, ux));
        solid = (int32_t)((m << (nbi & 63u)) >> 32) < 0;
        ce = m == 0 ? 3u : 1u;
    }
    // the hit test after the join: its complement is then a scalar mask operation
    // rather than a second vector compare of the sector mask (ONE_EXIT: the only exit;
    // the lookup count then counts the steps that go on, and ray_trace adds the hit's)
    if (ONE_EXIT) {
        if (!inb || solid) return false;
        r.iters++;
        asm volatile("" : "+v"(r.iters));
    } else if (solid) {
        *hit_slot = slot1 - 1;
        hit = true;
        return false;
    }
    // the 2x2x2 group holding bi: (half >> (bi & 0xA)) & 0x00330033 on the 32-bit
    // half selected by bit 5 of bi
    uint32_t grp = (uint32_t)(m >> (~nbi & 0x2Au)) & 0x00330033u;
    uint32_t cm = ce >> min_1(grp);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // d < 0: v & ~cm (near corner); d >= 0: v | cm (far corner); rel = that corner - W
        int32_t rel;
        if (ORIGIN != ORIGIN_ANY) {
            rel = (int32_t)bitfield_insert(~cm, (uint32_t)f[a], (uint32_t)r.dm[a]);
        } else {
            v[a] = (int32_t)bitfield_insert(~cm, (uint32_t)v[a], (uint32_t)r.dm[a]);
            rel = (int32_t)((uint32_t)v[a] - (uint32_t)r.W[a]);
        }
        r.side[a] = fmaf((float)rel, r.inv[a], r.ts[a]);
    }
    float tmin;
    if (EXACT)
        tmin = min_x86(min_x86(r.side[0], r.side[1]), r.side[2]);
    else
        tmin = fminf(fminf(r.side[0], r.side[1]), r.side[2]);
    tmin += 0.001f;
#pragma unroll
    for (int a = 0; a < 3; a++) r.cur[a] = fmaf(tmin, r.d[a], r.o[a]);
    return true;
}

// ray_advance as a result code.
template <bool EXACT>
__device__ __forceinline__ int ray_step(const WorldView& w, RayState& r, int32_t v[3], uint32_t* hit_slot) {
    bool hit;
    if (ray_advance<EXACT>(w, r, v, hit, hit_slot)) return STEP_CONTINUE;
    return hit ? STEP_HIT : STEP_MISS;
}

// Brick slot of voxel v, which lies in an allocated brick (a hit voxel): the
// lookup of a descending step, repeated once after the loop so the loop does
// not carry the slot.
// The record is read by the range-checked structured load: an index past the records
// reads 0 (slot = 0), so no v can make this lookup leave the world's buffers.
__device__ __forceinline__ uint32_t brick_slot_at(const WorldView& w, const int32_t v[3]) {
    uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
    uint32_t si = (ux >> 5) | (uz >> 5) << w.shift_xz | (uy >> 5) << (2 * w.shift_xz);
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 sec = (u32x4)vrt_struct_buffer_load_v4i32(sector_records_rsrc(w), (int32_t)si, 0, 0, 0);
    uint64_t m = (uint64_t)sec.x | (uint64_t)sec.y << 32;
    uint32_t bi = __builtin_amdgcn_ubfe(ux, 3, 2) | __builtin_amdgcn_ubfe(uz, 3, 2) << 2 |
                  __builtin_amdgcn_ubfe(uy, 3, 2) << 4;
    return sec.z + (uint32_t)__popcll(m & ((1ull << bi) - 1ull));
}

// Whether voxel v (any coordinates) is solid, with its brick slot when it is.  A
// parked ray needs no hit flag: a hit leaves v on the solid voxel, a miss outside
// the world, and the cap on a corner of the empty cell it was stepping over.
__device__ __forceinline__ bool solid_voxel_at(const WorldView& w, const int32_t v[3], uint32_t& slot) {
    uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
    if (!((ux | uz) < w.size_xz && uy < w.size_y)) return false;
    uint32_t si = (ux >> 5) | (uz >> 5) << w.shift_xz | (uy >> 5) << (2 * w.shift_xz);
    uint4 sec = w.sec[si];
    uint64_t m = (uint64_t)sec.x | (uint64_t)sec.y << 32;
    uint32_t bi = __builtin_amdgcn_ubfe(ux, 3, 2) | __builtin_amdgcn_ubfe(uz, 3, 2) << 2 |
                  __builtin_amdgcn_ubfe(uy, 3, 2) << 4;
    if (!((m >> bi) & 1ull)) return false;
    slot = sec.z + (uint32_t)__popcll(m & ((1ull << bi) - 1ull));
    uint32_t cell = __builtin_amdgcn_ubfe(ux, 2, 1) | __builtin_amdgcn_ubfe(uz, 2, 1) << 1 |
                    __builtin_amdgcn_ubfe(uy, 2, 1) << 2;
    uint32_t vi = (ux & 3u) | (uz & 3u) << 2 | (uy & 3u) << 4;
    return (w.occ[(size_t)slot * 8u + cell] >> vi) & 1ull;
}

// RayCast's loop (CpuRenderer.cpp:185-202): steps until a hit, a miss or
// max_iters steps; true on a hit, with v the hit voxel and slot its brick slot.
// first: the steps the ray has already made (r.iters), for a ray resumed from a
// saved position and side distances; the loop then makes at most max_iters - first.
template <bool EXACT, bool CUBIC = false, int ORIGIN = ORIGIN_ANY>
__device__ __forceinline__ bool ray_trace(const WorldView& w, RayState& r, uint32_t max_iters, int32_t v[3],
                                          uint32_t& slot, uint32_t first = 0) {
    bool hit = false;
    if (EXACT) {
        for (uint32_t it = first; it < max_iters; it++)
            if (!ray_advance<EXACT, CUBIC, ORIGIN>(w, r, v, hit, &slot)) break;
        return hit;
    }
    // The fast loop has one exit and no hit flag (see ONE_EXIT above): a hit leaves v on
    // the solid voxel, a miss outside the world, and the cap on a voxel of the empty cell
    // it was stepping over (every entry point rejects a cap of 0, so v is set by at least
    // one step); r.iters counts the steps that went on, max_iters only after the cap.
    uint32_t unused;
    for (uint32_t it = first; it < max_iters; it++)
        if (!ray_advance<EXACT, CUBIC, ORIGIN, true>(w, r, v, hit, &unused)) break;
    if (ORIGIN == ORIGIN_ANY) {
        // here the step overwrites v with the cell corner, and a test of the step count
        // and v costs the loop 7 VALU instructions in exit copies: look the voxel up
        bool h = solid_voxel_at(w, v, slot);
        r.iters += h ? 1u : 0u;
        return h;
    }
    // a hit is a ray stopped inside the world before the cap
    uint32_t steps = r.iters;
    bool h = steps < max_iters && (CUBIC ? (uint32_t)(v[0] | v[1] | v[2]) < w.size_xz
                                         : ((uint32_t)(v[0] | v[2]) < w.size_xz && (uint32_t)v[1] < w.size_y));
    if (h) slot = brick_slot_at(w, v);
    r.iters = steps + (h ? 1u : 0u);
    return h;
}

template <bool CUBIC>
__device__ __forceinline__ bool ray_trace_tame(const WorldView& w, RayState& r, uint32_t max_iters, int32_t v[3],
                                               uint32_t& slot, int origin, uint32_t first) {
    if (origin == ORIGIN_ZERO) return ray_trace<false, CUBIC, ORIGIN_ZERO>(w, r, max_iters, v, slot, first);
    if (origin == ORIGIN_ALIGNED) return ray_trace<false, CUBIC, ORIGIN_ALIGNED>(w, r, max_iters, v, slot, first);
    return ray_trace<false, CUBIC>(w, r, max_iters, v, slot, first);
}

// ray_trace with the step picked per wave: the fast step when every ray of the
// wave is tame (wave_tame), with the single-compare bounds test in cubic worlds
// and specialized for the rays' world origin (origin: an OriginKind, uniform over
// the wave).
__device__ __forceinline__ bool ray_trace_wave(const WorldView& w, RayState& r, uint32_t max_iters, int32_t v[3],
                                               uint32_t& slot, bool wave_tame, int origin = ORIGIN_ANY,
                                               uint32_t first = 0) {
    if (!wave_tame) return ray_trace<true>(w, r, max_iters, v, slot, first);
    if (w.size_y == w.size_xz) return ray_trace_tame<true>(w, r, max_iters, v, slot, origin, first);
    return ray_trace_tame<false>(w, r, max_iters, v, slot, origin, first);
}

// Hit record fields (vrt_hit in vrt.h) of a finished ray.
__device__ __forceinline__ void ray_hit_record(const WorldView& w, const RayState& r, const int32_t v[3],
                                               uint32_t slot, float& t, int4& rec) {
    float tt = min_x86(min_x86(r.side[0], r.side[1]), r.side[2]);
    bool X = r.side[0] == tt, Y = r.side[1] == tt, Z = !X && !Y;
    uint32_t code = (X ? (signbit(r.d[0]) ? 1u : 3u) : 0u) | (Y ? (signbit(r.d[1]) ? 1u : 3u) : 0u) << 2 |
                    (Z ? (signbit(r.d[2]) ? 1u : 3u) : 0u) << 4;
    uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
    uint32_t mat = w.vox[(size_t)slot * 512 + ((ux & 7u) | (uz & 7u) << 3 | (uy & 7u) << 6)];
    t = tt;
    rec = make_int4(v[0], v[1], v[2], (int)(mat | code << 8 | (r.iters & 0xFFFFu) << 16));
}

__device__ __forceinline__ void ray_miss_record(const RayState& r, float& t, int4& rec) {
    t = -1.0f;
    rec = make_int4(0, 0, 0, (int)((r.iters & 0xFFFFu) << 16));
}

}  // namespace vrt
