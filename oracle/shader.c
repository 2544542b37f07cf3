/*
 * shader.c - scalar restatement of the GLSL traversal (the "shader method").
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 * Follows src/VoxelRT/Shaders/VoxelTraversal.glsl per ray:
 *   isFineLod / shr64       :14-20
 *   getIsotropicLod         :22-32
 *   alignToCellBoundaries   :33-39
 *   findAnyOccupiedPos      :40-52
 *   getStepPos              :92-131
 *   clipRayToAABB           :133-145
 *   rayCast                 :162-204
 *   rayCastCoarse           :206-243
 * with the sector-of-sectors masks of src/VoxelRT/GpuRenderer.cpp:133-149 and
 * the ray-cell interaction LUT of GpuRenderer.cpp:192-210.  The FP contract is
 * stated in oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t lin3(uint32_t x, uint32_t y, uint32_t z, uint32_t sxz, uint32_t sy) {
    uint32_t mxz = (1u << sxz) - 1u, my = (1u << sy) - 1u;
    return (x & mxz) | (z & mxz) << sxz | (y & my) << (2 * sxz);
}
static inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t floor2i(float x) {
    float f = floorf(x);
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}
static inline int fine_bit(uint64_t m, uint32_t idx) {
    uint32_t half = idx < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
    return (half >> (idx & 31u)) & 1u;
}

int oracle_world_build_groups(oracle_world* w) {
    w->gshift_xz = w->shift_xz > 2 ? w->shift_xz - 2 : 0;
    w->gshift_y = w->shift_y > 2 ? w->shift_y - 2 : 0;
    size_t ngroups = (size_t)1 << (2 * w->gshift_xz + w->gshift_y);
    w->sector_groups = (uint64_t*)calloc(ngroups, sizeof(uint64_t));
    if (!w->sector_groups) return 1;
    uint32_t nxz = 1u << w->shift_xz, ny = 1u << w->shift_y;
    for (uint32_t sy = 0; sy < ny; sy++)
        for (uint32_t sz = 0; sz < nxz; sz++)
            for (uint32_t sx = 0; sx < nxz; sx++) {
                if (!w->sector_masks[lin3(sx, sy, sz, w->shift_xz, w->shift_y)]) continue;
                uint32_t g = lin3(sx >> 2, sy >> 2, sz >> 2, w->gshift_xz, w->gshift_y);
                w->sector_groups[g] |= 1ull << lin3(sx & 3u, sy & 3u, sz & 3u, 2, 2);
            }
    return 0;
}

/* GenerateRayCellInteractionMaskLUT (GpuRenderer.cpp:192-210): for direction octant
 * oct (bit a set = positive along axis a) and origin cell o of a 4x4x4 mask, the
 * cells o + j * dir for all j in the mask's extent that stay inside it. */
void oracle_ray_cell_lut(uint64_t table[512]) {
    for (uint32_t oct = 0; oct < 8; oct++) {
        int dir[3] = {(oct & 1) ? 1 : -1, (oct & 2) ? 1 : -1, (oct & 4) ? 1 : -1};
        for (uint32_t oi = 0; oi < 64; oi++) {
            int org[3] = {(int)(oi & 3), (int)(oi >> 4 & 3), (int)(oi >> 2 & 3)}; /* x, y, z */
            uint64_t mask = 0;
            for (uint32_t j = 0; j < 64; j++) {
                int step[3] = {(int)(j & 3), (int)(j >> 4 & 3), (int)(j >> 2 & 3)};
                int p[3];
                int inside = 1;
                for (int a = 0; a < 3; a++) {
                    p[a] = org[a] + step[a] * dir[a];
                    inside = inside && p[a] >= 0 && p[a] < 4;
                }
                if (inside) mask |= 1ull << lin3((uint32_t)p[0], (uint32_t)p[1], (uint32_t)p[2], 2, 2);
            }
            table[oi + oct * 64] = mask;
        }
    }
}

enum { SH_CONTINUE = 0, SH_HIT = 1 };

/* getStepPos (:92-131). */
static int shader_step(const oracle_world* w, int32_t v[3], const float d[3], int coarse, const uint64_t* lut) {
    uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
    uint32_t sx = ux >> 5, sy = uy >> 5, sz = uz >> 5;
    uint32_t si = lin3(sx, sy, sz, w->shift_xz, w->shift_y);
    uint64_t m = w->sector_masks[si];
    uint32_t idx = lin3(ux >> 3, uy >> 3, uz >> 3, 2, 2);
    int32_t scale = 8;
    if (fine_bit(m, idx)) {
        m = w->occupancy[(size_t)si * 512 + (size_t)idx * 8 + lin3(ux >> 2, uy >> 2, uz >> 2, 1, 1)];
        idx = lin3(ux, uy, uz, 2, 2);
        scale = 1;
        if (fine_bit(m, idx)) return SH_HIT;
    } else if (m == 0) {
        m = w->sector_groups[lin3(sx >> 2, sy >> 2, sz >> 2, w->gshift_xz, w->gshift_y)];
        idx = lin3(sx, sy, sz, 2, 2);
        scale = 32;
    }
    if (lut) {
        uint32_t oct = (d[0] < 0 ? 0u : 1u) + (d[1] < 0 ? 0u : 2u) + (d[2] < 0 ? 0u : 4u);
        m &= lut[idx + oct * 64];
    }
    int32_t iso;
    if (m == 0) {
        iso = 4;
    } else {
        uint32_t half = idx < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
        iso = ((half >> (idx & 0xAu)) & 0x00330033u) == 0 ? 2 : 1;
    }
    int32_t lod = iso * scale;
    if (coarse && lod < 4) {
        /* findAnyOccupiedPos (:40-52): lowest set bit, low word first */
        uint32_t bit = (uint32_t)m != 0 ? (uint32_t)__builtin_ctz((uint32_t)m)
                                        : 32u + (uint32_t)__builtin_ctz((uint32_t)(m >> 32));
        v[0] = (v[0] & ~3) | (int32_t)(bit & 3u);
        v[1] = (v[1] & ~3) | (int32_t)(bit >> 4 & 3u);
        v[2] = (v[2] & ~3) | (int32_t)(bit >> 2 & 3u);
        return SH_HIT;
    }
    int32_t cm = lod - 1;
    for (int a = 0; a < 3; a++) v[a] = d[a] < 0 ? (v[a] & ~cm) : (v[a] | cm);
    return SH_CONTINUE;
}

/* clipRayToAABB (:133-145) against [1 - W, size - W - 1]. */
static void clip_ray(const oracle_world* w, const float o[3], const float d[3], const int32_t W[3], float out[3]) {
    float size[3] = {(float)w->size_xz, (float)w->size_y, (float)w->size_xz};
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        float inv = 1.0f / d[a];
        float bmin = (float)wrap_sub(1, W[a]);
        float bmax = (size[a] - (float)W[a]) - 1.0f;
        float t1 = (bmin - o[a]) * inv, t2 = (bmax - o[a]) * inv;
        lo[a] = fminf(t1, t2);
        hi[a] = fmaxf(t1, t2);
    }
    float tmin = fmaxf(lo[0], fmaxf(lo[1], lo[2]));
    float tmax = fminf(hi[0], fminf(hi[1], hi[2]));
    int clip = tmin > 0.0f && tmin < tmax;
    for (int a = 0; a < 3; a++) out[a] = clip ? fmaf(d[a], tmin, o[a]) : o[a];
}

void oracle_shader_ray(const oracle_world* w, const float o[3], const float d[3], const int32_t W[3],
                       uint32_t method, const uint64_t* lut, uint32_t cap, float* t_out, oracle_hit* rec,
                       float pos[3]) {
    float start[3], org[3], inv[3], ts[3], side[3], cur[3];
    clip_ray(w, o, d, W, start);
    for (int a = 0; a < 3; a++) {
        org[a] = method == 1 ? start[a] : o[a];
        inv[a] = 1.0f / d[a];
        ts[a] = ((d[a] < 0 ? 0.0f : 1.0f) - org[a]) * inv[a];
    }
    int32_t v[3];
    for (int a = 0; a < 3; a++) v[a] = wrap_add(W[a], floor2i(start[a]));
    for (uint32_t i = 0; i < cap; i++) {
        for (int a = 0; a < 3; a++) side[a] = fmaf((float)wrap_sub(v[a], W[a]), inv[a], ts[a]);
        float tmin = fminf(fminf(side[0], side[1]), side[2]);
        if (method == 1) {
            tmin = tmin + 0.001f;
        } else {
            uint32_t b;
            memcpy(&b, &tmin, 4);
            b += 5u;
            memcpy(&tmin, &b, 4);
        }
        for (int a = 0; a < 3; a++) cur[a] = fmaf(tmin, d[a], org[a]);
        int coarse = method == 1 && i > 30;
        for (int a = 0; a < 3; a++) v[a] = wrap_add(W[a], floor2i(cur[a]));
        if (!((uint32_t)(v[0] | v[2]) < w->size_xz && (uint32_t)v[1] < w->size_y)) {
            *t_out = -1.0f;
            rec->x = rec->y = rec->z = 0;
            rec->info = (i & 0xFFFFu) << 16;
            if (pos) pos[0] = pos[1] = pos[2] = 0.0f;
            return;
        }
        if (shader_step(w, v, d, coarse, lut) == SH_HIT) {
            uint32_t code = 0;
            for (int a = 0; a < 3; a++)
                if (tmin >= side[a]) code |= (d[a] < 0 ? 1u : (d[a] > 0 ? 3u : 0u)) << (2 * a);
            *t_out = tmin;
            rec->x = v[0];
            rec->y = v[1];
            rec->z = v[2];
            rec->info = oracle_world_get_voxel(w, v[0], v[1], v[2]) | code << 8 | (i & 0xFFFFu) << 16;
            if (pos) memcpy(pos, cur, sizeof(cur));
            return;
        }
    }
    *t_out = -1.0f;
    rec->x = rec->y = rec->z = 0;
    rec->info = (cap & 0xFFFFu) << 16;
    if (pos) pos[0] = pos[1] = pos[2] = 0.0f;
}

void oracle_cast_shader(const oracle_world* w, const float* const o[3], const float* const d[3],
                        const int32_t world_origin[3], uint32_t begin, uint32_t end, uint32_t method,
                        uint32_t lut, uint32_t max_iters, float* t, oracle_hit* rec, float* const pos[3]) {
    uint64_t table[512];
    oracle_ray_cell_lut(table);
    uint32_t cap = max_iters ? max_iters : (method == 1 ? 96u : 256u);
    for (uint32_t i = begin; i < end; i++) {
        float oi[3] = {o[0][i], o[1][i], o[2][i]};
        float di[3] = {d[0][i], d[1][i], d[2][i]};
        float p[3];
        oracle_shader_ray(w, oi, di, world_origin, method, lut ? table : NULL, cap, &t[i], &rec[i], pos ? p : NULL);
        if (pos) {
            pos[0][i] = p[0];
            pos[1][i] = p[1];
            pos[2][i] = p[2];
        }
    }
}
