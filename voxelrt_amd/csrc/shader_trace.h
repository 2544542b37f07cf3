// shader_trace.h - the GLSL traversal ("shader method") per lane.
//
// src/VoxelRT/Shaders/VoxelTraversal.glsl differs from the CPU cast that
// traverse.h reproduces: a clipped start (clipRayToAABB, :133-145), a
// sector-of-sectors LOD level (:110-114), the optional ray-cell interaction
// table (:116-121), a +5 ulp bias (:172), a 256-iteration cap, and a coarse
// variant for shadow and bounce rays (rayCastCoarse, :206-243) that accepts
// any solid voxel of a non-empty 4x4x4 cell from its 31st iteration on.  The
// FP contract is the one stated for vrt_cast_shader in vrt.h.
#pragma once

#include <hip/hip_runtime.h>

#include "traverse.h"
#include "vrt_internal.h"
#include "vrt_math.h"

namespace vrt {

__device__ __forceinline__ bool fine_bit(uint64_t m, uint32_t idx) { return (m >> (idx & 63u)) & 1u; }

// getStepPos (:92-131).  Returns true when v is solid (or, for a coarse step, when v
// was moved to the first solid voxel of its non-empty 4x4x4 cell); otherwise moves v
// to the far corner, in ray direction, of the largest empty aligned cell.
template <bool LUT, bool TAME>
__device__ __forceinline__ bool shader_step(const ShaderView& sv, int32_t v[3], const float d[3],
                                            const int32_t dm[3], uint32_t oct, bool coarse) {
    const WorldView& w = sv.w;
    uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
    uint32_t sx = ux >> 5, sy = uy >> 5, sz = uz >> 5;
    uint32_t si = sx | sz << w.shift_xz | sy << (2 * w.shift_xz);
    uint4 sec = w.sec[si];
    uint64_t m = (uint64_t)sec.x | (uint64_t)sec.y << 32;
    uint32_t idx = __builtin_amdgcn_ubfe(ux, 3, 2) | __builtin_amdgcn_ubfe(uz, 3, 2) << 2 |
                   __builtin_amdgcn_ubfe(uy, 3, 2) << 4;
    int32_t scale = 8;
    if (fine_bit(m, idx)) {
        uint32_t slot = sec.z + (uint32_t)__popcll(m & ((1ull << idx) - 1ull));
        uint32_t cell = __builtin_amdgcn_ubfe(ux, 2, 1) | __builtin_amdgcn_ubfe(uz, 2, 1) << 1 |
                        __builtin_amdgcn_ubfe(uy, 2, 1) << 2;
        m = w.occ[(size_t)slot * 8 + cell];
        idx = (ux & 3u) | (uz & 3u) << 2 | (uy & 3u) << 4;
        scale = 1;
        if (fine_bit(m, idx)) return true;
    } else if (m == 0) {
        uint32_t gmx = (1u << sv.gshift_xz) - 1u, gmy = (1u << sv.gshift_y) - 1u;
        uint32_t g = ((sx >> 2) & gmx) | ((sz >> 2) & gmx) << sv.gshift_xz | ((sy >> 2) & gmy) << (2 * sv.gshift_xz);
        m = sv.groups[g];
        idx = (sx & 3u) | (sz & 3u) << 2 | (sy & 3u) << 4;
        scale = 32;
    }
    if constexpr (LUT) m &= sv.lut[idx + oct * 64];
    int32_t iso;
    if (m == 0) {
        iso = 4;
    } else {
        uint32_t half = idx < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
        iso = ((half >> (idx & 0xAu)) & 0x00330033u) == 0 ? 2 : 1;
    }
    int32_t lod = iso * scale;
    if (coarse && lod < 4) {
        // findAnyOccupiedPos (:40-52): lowest set bit, low word first
        uint32_t bit = (uint32_t)m != 0 ? (uint32_t)__builtin_ctz((uint32_t)m)
                                        : 32u + (uint32_t)__builtin_ctz((uint32_t)(m >> 32));
        v[0] = (v[0] & ~3) | (int32_t)(bit & 3u);
        v[1] = (v[1] & ~3) | (int32_t)((bit >> 4) & 3u);
        v[2] = (v[2] & ~3) | (int32_t)((bit >> 2) & 3u);
        return true;
    }
    // d < 0: v & ~cm, else v | cm (the fast loop: one bitfield insert of dm, all ones unless
    // d < 0)
    uint32_t cm = (uint32_t)lod - 1u;
#pragma unroll
    for (int a = 0; a < 3; a++)
        v[a] = TAME ? (int32_t)bitfield_insert(~cm, (uint32_t)v[a], (uint32_t)dm[a])
                    : (d[a] < 0.0f ? (int32_t)((uint32_t)v[a] & ~cm) : (int32_t)((uint32_t)v[a] | cm));
    return false;
}

// clipRayToAABB (:133-145) against [1 - W, size - W - 1].
__device__ __forceinline__ void clip_ray(const WorldView& w, const float o[3], const float d[3], const int32_t W[3],
                                         float out[3]) {
    float size[3] = {(float)w.size_xz, (float)w.size_y, (float)w.size_xz};
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float inv = 1.0f / d[a];
        float bmin = (float)(int32_t)(1u - (uint32_t)W[a]);
        float bmax = (size[a] - (float)W[a]) - 1.0f;
        float t1 = (bmin - o[a]) * inv, t2 = (bmax - o[a]) * inv;
        lo[a] = fminf(t1, t2);
        hi[a] = fmaxf(t1, t2);
    }
    float tmin = fmaxf(lo[0], fmaxf(lo[1], lo[2]));
    float tmax = fminf(hi[0], fminf(hi[1], hi[2]));
    bool clip = tmin > 0.0f && tmin < tmax;
#pragma unroll
    for (int a = 0; a < 3; a++) out[a] = clip ? fmaf(d[a], tmin, o[a]) : o[a];
}

// Result of one shader-method trace.
struct ShaderHit {
    float t;        // biased tmin at the hit, -1 on a miss
    float cur[3];   // hit position (relative to W)
    int32_t v[3];   // hit voxel (absolute)
    uint32_t code;  // normal code (vrt_hit info bits 8-13)
    uint32_t iters; // loop index at the hit or miss, the cap when exhausted
    uint32_t mat;   // material id, 0 on a miss
};

// The loop of rayCast / rayCastCoarse.  TAME (every active ray of the wave has finite
// 1/d and tStart, |origin| < 2^20 and 0.25 <= |d|^2 <= 4): positions stay far inside int32
// range for any cap (a step moves less than 128 sqrt(3) / 0.5 voxels) and tmin is never NaN,
// so floor is one V_CVT_FLR_I32_F32 instead of floor2i's range test.
template <bool COARSE, bool LUT, bool TAME>
__device__ __forceinline__ bool shader_loop(const ShaderView& sv, const float d[3], const float org[3],
                                            const float inv[3], const float ts[3], const int32_t W[3],
                                            const int32_t dm[3], uint32_t oct, uint32_t cap, int32_t v[3],
                                            float side[3], float cur[3], float& tmin, uint32_t& it) {
    const WorldView& w = sv.w;
    bool hit = false;
    for (; it < cap; it++) {
#pragma unroll
        for (int a = 0; a < 3; a++)
            side[a] = fmaf((float)(int32_t)((uint32_t)v[a] - (uint32_t)W[a]), inv[a], ts[a]);
        tmin = fminf(fminf(side[0], side[1]), side[2]);
        if (COARSE)
            tmin = tmin + 0.001f;
        else
            tmin = __uint_as_float(__float_as_uint(tmin) + 5u);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            cur[a] = fmaf(tmin, d[a], org[a]);
            v[a] = (int32_t)((uint32_t)W[a] + (uint32_t)(TAME ? floor_i32_in_range(cur[a]) : floor2i(cur[a])));
        }
        if (!((uint32_t)(v[0] | v[2]) < w.size_xz && (uint32_t)v[1] < w.size_y)) break;
        if (shader_step<LUT, TAME>(sv, v, d, dm, oct, COARSE && it > 30)) {
            hit = true;
            break;
        }
    }
    return hit;
}

// rayCast (COARSE = false, :162-204) or rayCastCoarse (COARSE = true, :206-243).  FAST
// takes the tame loop when every ray of the wave is tame.
template <bool COARSE, bool LUT, bool FAST = !COARSE>
__device__ __forceinline__ bool shader_trace(const ShaderView& sv, const float o[3], const float d[3],
                                             const int32_t W[3], uint32_t cap, ShaderHit& out) {
    const WorldView& w = sv.w;
    float start[3], org[3], inv[3], ts[3], side[3] = {0.0f, 0.0f, 0.0f}, cur[3] = {0.0f, 0.0f, 0.0f};
    clip_ray(w, o, d, W, start);
    int32_t v[3], dm[3];
    bool tame = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        org[a] = COARSE ? start[a] : o[a];
        inv[a] = 1.0f / d[a];
        ts[a] = ((d[a] < 0.0f ? 0.0f : 1.0f) - org[a]) * inv[a];
        v[a] = (int32_t)((uint32_t)W[a] + (uint32_t)floor2i(start[a]));
        dm[a] = d[a] < 0.0f ? 0 : -1;
        tame = tame && fabsf(inv[a]) < INFINITY && fabsf(ts[a]) < INFINITY && fabsf(org[a]) < 1048576.0f;
    }
    float len2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    tame = tame && len2 >= 0.25f && len2 <= 4.0f;
    const uint32_t oct = (d[0] < 0.0f ? 0u : 1u) + (d[1] < 0.0f ? 0u : 2u) + (d[2] < 0.0f ? 0u : 4u);
    float tmin = 0.0f;
    uint32_t it = 0;
    // the fast loop by default for rayCast only: inside the renderer rayCastCoarse's short
    // traces (cap 96, shadow and bounce rays) measured no gain from it, and its second loop
    // copy slowed the renderer's kernel, which holds three traces; the batch kernel, which
    // holds one, takes it for both (coarse C2 camera rays: 14.8 against 13.8 Grays/s)
    bool hit = FAST && __all(tame)
                   ? shader_loop<COARSE, LUT, true>(sv, d, org, inv, ts, W, dm, oct, cap, v, side, cur, tmin, it)
                   : shader_loop<COARSE, LUT, false>(sv, d, org, inv, ts, W, dm, oct, cap, v, side, cur, tmin, it);
    // the record is built in locals and stored once, so the result stays in registers
    uint32_t code = 0, mat = 0;
    float t = -1.0f;
    int32_t hv[3] = {0, 0, 0};
    float hc[3] = {0.0f, 0.0f, 0.0f};
    if (hit) {
#pragma unroll
        for (int a = 0; a < 3; a++)
            if (tmin >= side[a]) code |= (d[a] < 0.0f ? 1u : (d[a] > 0.0f ? 3u : 0u)) << (2 * a);
        uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
        uint32_t si = (ux >> 5) | (uz >> 5) << w.shift_xz | (uy >> 5) << (2 * w.shift_xz);
        uint4 sec = w.sec[si];
        uint64_t m = (uint64_t)sec.x | (uint64_t)sec.y << 32;
        uint32_t bi = __builtin_amdgcn_ubfe(ux, 3, 2) | __builtin_amdgcn_ubfe(uz, 3, 2) << 2 |
                      __builtin_amdgcn_ubfe(uy, 3, 2) << 4;
        uint32_t slot = sec.z + (uint32_t)__popcll(m & ((1ull << bi) - 1ull));
        mat = w.vox[(size_t)slot * 512 + ((ux & 7u) | (uz & 7u) << 3 | (uy & 7u) << 6)];
        t = tmin;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            hc[a] = cur[a];
            hv[a] = v[a];
        }
    }
    out.iters = it;
    out.t = t;
    out.code = code;
    out.mat = mat;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        out.cur[a] = hc[a];
        out.v[a] = hv[a];
    }
    return hit;
}

}  // namespace vrt
