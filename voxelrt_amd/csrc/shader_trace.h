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

#include "vrt_internal.h"
#include "vrt_math.h"

namespace vrt {

__device__ __forceinline__ bool fine_bit(uint64_t m, uint32_t idx) { return (m >> (idx & 63u)) & 1u; }

// getStepPos (:92-131).  Returns true when v is solid (or, for a coarse step, when v
// was moved to the first solid voxel of its non-empty 4x4x4 cell); otherwise moves v
// to the far corner, in ray direction, of the largest empty aligned cell.
template <bool LUT>
__device__ __forceinline__ bool shader_step(const ShaderView& sv, int32_t v[3], const float d[3], uint32_t oct,
                                            bool coarse) {
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
    int32_t cm = lod - 1;
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = d[a] < 0.0f ? (v[a] & ~cm) : (v[a] | cm);
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

// rayCast (COARSE = false, :162-204) or rayCastCoarse (COARSE = true, :206-243).
template <bool COARSE, bool LUT>
__device__ __forceinline__ bool shader_trace(const ShaderView& sv, const float o[3], const float d[3],
                                             const int32_t W[3], uint32_t cap, ShaderHit& out) {
    const WorldView& w = sv.w;
    float start[3], org[3], inv[3], ts[3], side[3] = {0.0f, 0.0f, 0.0f}, cur[3] = {0.0f, 0.0f, 0.0f};
    clip_ray(w, o, d, W, start);
    int32_t v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        org[a] = COARSE ? start[a] : o[a];
        inv[a] = 1.0f / d[a];
        ts[a] = ((d[a] < 0.0f ? 0.0f : 1.0f) - org[a]) * inv[a];
        v[a] = (int32_t)((uint32_t)W[a] + (uint32_t)floor2i(start[a]));
    }
    const uint32_t oct = (d[0] < 0.0f ? 0u : 1u) + (d[1] < 0.0f ? 0u : 2u) + (d[2] < 0.0f ? 0u : 4u);
    float tmin = 0.0f;
    bool hit = false;
    uint32_t it = 0;
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
            v[a] = (int32_t)((uint32_t)W[a] + (uint32_t)floor2i(cur[a]));
        }
        if (!((uint32_t)(v[0] | v[2]) < w.size_xz && (uint32_t)v[1] < w.size_y)) break;
        if (shader_step<LUT>(sv, v, d, oct, COARSE && it > 30)) {
            hit = true;
            break;
        }
    }
    out.iters = it;
    if (!hit) {
        out.t = -1.0f;
        out.cur[0] = out.cur[1] = out.cur[2] = 0.0f;
        out.v[0] = out.v[1] = out.v[2] = 0;
        out.code = 0;
        out.mat = 0;
        return false;
    }
    uint32_t code = 0;
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
    out.mat = w.vox[(size_t)slot * 512 + ((ux & 7u) | (uz & 7u) << 3 | (uy & 7u) << 6)];
    out.t = tmin;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        out.cur[a] = cur[a];
        out.v[a] = v[a];
    }
    out.code = code;
    return true;
}

}  // namespace vrt
