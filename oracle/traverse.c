/*
 * traverse.c - scalar restatement of the reference CPU traversal.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 * Follows, per ray, the vector code of src/VoxelRT/CpuRenderer.cpp:
 *   world mirror      FlatVoxelStorage, SyncBuffers, UpdateOccupancy  (:20-84)
 *   bounds test       GetInboundMask                                 (:114-117)
 *   material fetch    GetVoxelMaterial                               (:120-132)
 *   one LOD step      GetStepPos                                     (:135-171)
 *   the cast loop     RayCast                                        (:172-224)
 * with the indexers of src/VoxelRT/VoxelMap.h:53-102 and the lane semantics of
 * src/LibGlimpsw/SwRast/SIMD_AVX2.h (floor2i :115, min :129, ucmp_lt :178,
 * masked gathers returning 0 :33).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- indexers (VoxelMap.h:60-102): x | z << s | y << 2s with wrap masks ---- */

static inline uint32_t lin3(int32_t x, int32_t y, int32_t z, uint32_t sxz, uint32_t sy) {
    uint32_t mxz = (1u << sxz) - 1u, my = (1u << sy) - 1u;
    return ((uint32_t)x & mxz) | ((uint32_t)z & mxz) << sxz | ((uint32_t)y & my) << (2 * sxz);
}
static inline uint32_t mask_index(int32_t x, int32_t y, int32_t z) { return lin3(x, y, z, 2, 2); }
static inline uint32_t brick_index(int32_t x, int32_t y, int32_t z) { return lin3(x, y, z, 3, 3); }
static inline uint32_t cell_index(int32_t x, int32_t y, int32_t z) { return lin3(x, y, z, 1, 1); }
static inline uint32_t sector_index(const oracle_world* w, int32_t x, int32_t y, int32_t z) {
    return lin3(x >> 5, y >> 5, z >> 5, w->shift_xz, w->shift_y);
}

/* ---- world ---- */

/* Occupancy of one brick: 8 cells of 4x4x4, bit vx + 4 vz + 16 vy (CpuRenderer.cpp:63-83). */
static void build_occupancy(const uint8_t* brick, uint64_t* cells) {
    for (uint32_t cy = 0; cy < 8; cy += 4)
        for (uint32_t cz = 0; cz < 8; cz += 4)
            for (uint32_t cx = 0; cx < 8; cx += 4) {
                uint64_t mask = 0;
                for (uint32_t vy = 0; vy < 4; vy++)
                    for (uint32_t vz = 0; vz < 4; vz++)
                        for (uint32_t vx = 0; vx < 4; vx++) {
                            int occ = brick[brick_index(cx + vx, cy + vy, cz + vz)] != 0;
                            mask |= (uint64_t)occ << (vx + vz * 4 + vy * 16);
                        }
                cells[cell_index(cx / 4, cy / 4, cz / 4)] = mask;
            }
}

int oracle_world_create(uint32_t log2_size_xz, uint32_t log2_size_y, uint32_t nsec_in,
                        const uint32_t* sector_idx, const uint64_t* alloc_mask,
                        const uint8_t* bricks, oracle_world** out) {
    if (!out || log2_size_xz < 5 || log2_size_y < 5 || log2_size_xz > 14 || log2_size_y > 14) return 1;
    oracle_world* w = (oracle_world*)calloc(1, sizeof(oracle_world));
    if (!w) return 2;
    w->shift_xz = log2_size_xz - 5;
    w->shift_y = log2_size_y - 5;
    w->size_xz = 1u << log2_size_xz;
    w->size_y = 1u << log2_size_y;
    w->nsec = 1u << (2 * w->shift_xz + w->shift_y);
    /* calloc: pages of empty sectors are never touched, as with the reference's zero-filled buffer */
    w->sector_masks = (uint64_t*)calloc(w->nsec, sizeof(uint64_t));
    w->storage = (uint8_t*)calloc((size_t)w->nsec, 32768);
    w->occupancy = (uint64_t*)calloc((size_t)w->nsec * 512, sizeof(uint64_t));
    if (!w->sector_masks || !w->storage || !w->occupancy) {
        oracle_world_destroy(w);
        return 2;
    }
    const uint8_t* src = bricks;
    for (uint32_t s = 0; s < nsec_in; s++) {
        uint32_t si = sector_idx[s];
        if (si >= w->nsec) {
            oracle_world_destroy(w);
            return 1;
        }
        uint64_t m = alloc_mask[s];
        w->sector_masks[si] = m;
        for (uint32_t b = 0; b < 64; b++) {
            if (!((m >> b) & 1)) continue;
            uint8_t* dst = w->storage + (size_t)si * 32768 + (size_t)b * 512;
            memcpy(dst, src, 512);
            build_occupancy(dst, w->occupancy + (size_t)si * 512 + (size_t)b * 8);
            src += 512;
        }
    }
    if (oracle_world_build_groups(w)) {
        oracle_world_destroy(w);
        return 2;
    }
    *out = w;
    return 0;
}

void oracle_world_destroy(oracle_world* w) {
    if (!w) return;
    free(w->sector_masks);
    free(w->storage);
    free(w->occupancy);
    free(w->sector_groups);
    free(w);
}

void oracle_world_set_palette(oracle_world* w, const uint64_t* palette) {
    memcpy(w->palette, palette, sizeof(w->palette));
}

uint8_t oracle_world_get_voxel(const oracle_world* w, int32_t x, int32_t y, int32_t z) {
    if (!((uint32_t)(x | z) < w->size_xz && (uint32_t)y < w->size_y)) return 0;
    uint32_t si = sector_index(w, x, y, z);
    return w->storage[(size_t)si * 32768 + (size_t)mask_index(x >> 3, y >> 3, z >> 3) * 512 +
                      brick_index(x, y, z)];
}

/* ---- lane semantics ---- */

/* _mm256_cvtps_epi32(_mm256_floor_ps(x)): NaN and out-of-range give INT_MIN. */
static inline int32_t floor2i(float x) {
    float f = floorf(x);
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}
/* _mm256_min_ps(a, b) = a < b ? a : b (second operand on ties and NaN). */
static inline float min_x86(float a, float b) { return a < b ? a : b; }
static inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

/* GetStepPos (CpuRenderer.cpp:135-171) for one active lane.  Returns 1 when the
 * voxel at v is solid; otherwise moves v to the far corner (in ray direction)
 * of the largest empty aligned cell and returns 0. */
static inline int get_step_pos(const oracle_world* w, int32_t v[3], const float d[3], uint32_t* descents) {
    uint32_t si = sector_index(w, v[0], v[1], v[2]);
    uint64_t m = w->sector_masks[si];
    uint32_t bi = mask_index(v[0] >> 3, v[1] >> 3, v[2] >> 3);
    uint32_t half = bi < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
    uint32_t lod = 3;
    if ((half >> (bi & 31)) & 1) {
        (*descents)++;
        uint32_t cell = si * 512u + bi * 8u + cell_index(v[0] >> 2, v[1] >> 2, v[2] >> 2);
        m = w->occupancy[cell];
        bi = mask_index(v[0], v[1], v[2]);
        half = bi < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
        lod = 0;
        if ((half >> (bi & 31)) & 1) return 1;
    }
    /* level4: whole 4x4x4 mask empty; level2: the 2x2x2 group holding bi is empty */
    if (m == 0)
        lod += 2;
    else if (((half >> (bi & 0xA)) & 0x00330033u) == 0)
        lod += 1;
    int32_t cm = (int32_t)((1u << lod) - 1u);
    for (int a = 0; a < 3; a++) v[a] = d[a] < 0 ? (v[a] & ~cm) : (v[a] | cm);
    return 0;
}

static inline uint32_t normal_code(const float side[3], const float d[3], float t) {
    int X = side[0] == t, Y = side[1] == t;
    int Z = !X && !Y;
    int m[3] = {X, Y, Z};
    uint32_t code = 0;
    for (int a = 0; a < 3; a++) {
        /* (dir & -0.0) ^ -1.0: +1 for a negative-signed direction, -1 otherwise */
        uint32_t c = m[a] ? (signbit(d[a]) ? 1u : 3u) : 0u;
        code |= c << (2 * a);
    }
    return code;
}

/* VHitResult::UV (CpuRenderer.cpp:217-220): fract(X ? pos.y : pos.x), fract(Z ? pos.y : pos.z)
 * with fract(x) = x - floor(x) (SIMD_AVX2.h:121), X and Z the normal's tie masks. */
static void hit_uv(const float side[3], float t, const float cur[3], float uv[2]) {
    int X = side[0] == t, Y = side[1] == t;
    int Z = !X && !Y;
    float a = X ? cur[1] : cur[0], b = Z ? cur[1] : cur[2];
    uv[0] = a - floorf(a);
    uv[1] = b - floorf(b);
}

static void cast_one(const oracle_world* w, const float o[3], const float d[3], const int32_t W[3],
                     uint32_t max_iters, float* t_out, oracle_hit* rec, float pos[3], uint32_t* desc_out,
                     float uv[2]) {
    float inv[3], tstart[3], side[3] = {0.0f, 0.0f, 0.0f}, cur[3];
    for (int a = 0; a < 3; a++) {
        inv[a] = 1.0f / d[a];
        tstart[a] = ((d[a] < 0 ? 0.0f : 1.0f) - o[a]) * inv[a];
        cur[a] = o[a];
    }
    uint32_t descents = 0, iters = 0;
    int hit = 0;
    int32_t v[3];
    for (uint32_t i = 0; i < max_iters; i++) {
        for (int a = 0; a < 3; a++) v[a] = wrap_add(W[a], floor2i(cur[a]));
        if (!((uint32_t)(v[0] | v[2]) < w->size_xz && (uint32_t)v[1] < w->size_y)) break;
        iters = i + 1;
        if (get_step_pos(w, v, d, &descents)) {
            hit = 1;
            break;
        }
        for (int a = 0; a < 3; a++) side[a] = fmaf((float)wrap_sub(v[a], W[a]), inv[a], tstart[a]);
        float tmin = min_x86(min_x86(side[0], side[1]), side[2]) + 0.001f;
        for (int a = 0; a < 3; a++) cur[a] = fmaf(tmin, d[a], o[a]);
    }
    if (hit) {
        float t = min_x86(min_x86(side[0], side[1]), side[2]);
        uint32_t mat = oracle_world_get_voxel(w, v[0], v[1], v[2]);
        *t_out = t;
        rec->x = v[0];
        rec->y = v[1];
        rec->z = v[2];
        rec->info = mat | normal_code(side, d, t) << 8 | (iters & 0xFFFFu) << 16;
        if (pos) {
            pos[0] = cur[0];
            pos[1] = cur[1];
            pos[2] = cur[2];
        }
        if (uv) hit_uv(side, t, cur, uv);
    } else {
        /* out of bounds after `iters` lookups, or the cap was exhausted (iters = max_iters) */
        *t_out = -1.0f;
        rec->x = rec->y = rec->z = 0;
        rec->info = (iters & 0xFFFFu) << 16;
        if (pos) pos[0] = pos[1] = pos[2] = 0.0f;
        if (uv) uv[0] = uv[1] = 0.0f;
    }
    if (desc_out) *desc_out = descents;
}

void oracle_cast(const oracle_world* w, const float* const o[3], const float* const d[3],
                 const int32_t world_origin[3], uint32_t begin, uint32_t end, uint32_t max_iters,
                 float* t, oracle_hit* rec, float* const pos[3], uint16_t* descents, float* const uv[2]) {
    for (uint32_t i = begin; i < end; i++) {
        float oi[3] = {o[0][i], o[1][i], o[2][i]};
        float di[3] = {d[0][i], d[1][i], d[2][i]};
        float p[3], tc[2];
        uint32_t desc;
        cast_one(w, oi, di, world_origin, max_iters, &t[i], &rec[i], pos ? p : NULL, &desc, uv ? tc : NULL);
        if (pos) {
            pos[0][i] = p[0];
            pos[1][i] = p[1];
            pos[2][i] = p[2];
        }
        if (uv) {
            uv[0][i] = tc[0];
            uv[1][i] = tc[1];
        }
        if (descents) descents[i] = (uint16_t)(desc > 0xFFFF ? 0xFFFF : desc);
    }
}

void oracle_step_pos(const oracle_world* w, uint32_t n, const int32_t* v_in, const float* d, int32_t* v_out,
                     uint8_t* solid, uint8_t* descended) {
    for (uint32_t i = 0; i < n; i++) {
        int32_t v[3] = {v_in[3 * i], v_in[3 * i + 1], v_in[3 * i + 2]};
        uint32_t desc = 0;
        solid[i] = (uint8_t)get_step_pos(w, v, d + 3 * i, &desc);
        descended[i] = (uint8_t)desc;
        v_out[3 * i] = v[0];
        v_out[3 * i + 1] = v[1];
        v_out[3 * i + 2] = v[2];
    }
}

void oracle_count(const oracle_world* w, const float* const o[3], const float* const d[3],
                  const int32_t world_origin[3], uint32_t n, uint32_t max_iters, oracle_counters* out) {
    memset(out, 0, sizeof(*out));
    for (uint32_t i = 0; i < n; i++) {
        float oi[3] = {o[0][i], o[1][i], o[2][i]};
        float di[3] = {d[0][i], d[1][i], d[2][i]};
        float t;
        oracle_hit r;
        uint32_t desc;
        cast_one(w, oi, di, world_origin, max_iters, &t, &r, NULL, &desc, NULL);
        out->rays++;
        out->iters += r.info >> 16;
        out->descents += desc;
        out->hits += (r.info & 0xFF) != 0;
    }
}
