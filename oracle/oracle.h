/*
 * oracle.h - CPU restatement of VoxelRT's brickmap ray traversal.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; it is never linked into, loaded by or called from
 * the product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg use it.
 *
 * World layout: the reference's FlatVoxelStorage (src/VoxelRT/CpuRenderer.cpp:20-84)
 * with the view extents as parameters instead of the fixed 2048x512x2048:
 *   SectorMasks[nsec]                 u64, bit = brick allocated
 *   StorageBuffer[nsec * 32768]       u8, sector * 32 KiB + brick * 512 + voxel
 *   OccupancyStorage[nsec * 512]      u64, sector * 512 + brick * 8 + cell
 * i.e. dense per sector, unlike the product's popcount-compacted slots, so the
 * two layouts check each other.
 *
 * Traversal contract (both the scalar and the SIMD restatement):
 *   FP32, IEEE division for 1/d, explicit fmaf where the reference's
 *   -ffast-math build contracts, x86 min semantics (a < b ? a : b), x86
 *   cvtps2dq floor2i semantics (NaN / out of range -> INT_MIN), wrapping int
 *   arithmetic, un-biased hit distance, absolute hit voxel for the material.
 *   Build with -ffp-contract=off and without -ffast-math.
 */
#ifndef VRT_ORACLE_H
#define VRT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_world {
    uint32_t shift_xz, shift_y;   /* log2 of the sector grid extents */
    uint32_t size_xz, size_y;     /* voxel extents */
    uint32_t nsec;
    uint64_t* sector_masks;       /* [nsec] */
    uint8_t* storage;             /* [nsec * 32768] */
    uint64_t* occupancy;          /* [nsec * 512] */
    uint64_t palette[256];
    /* sector-of-sectors masks (GpuRenderer.cpp:133-149): one u64 per 4x4x4 group of
     * sectors, bit = that sector has a brick; group grid of max(1, n / 4) per axis */
    uint32_t gshift_xz, gshift_y;
    uint64_t* sector_groups;      /* [1 << (2 * gshift_xz + gshift_y)] */
} oracle_world;

/* Hit record in the same packing as the product's vrt_hit (include/vrt.h). */
typedef struct oracle_hit {
    int32_t x, y, z;
    uint32_t info;      /* material | normal code << 8 | iterations << 16 */
} oracle_hit;

typedef struct oracle_counters {
    uint64_t rays, iters, descents, hits;
} oracle_counters;

/* Build from a sector stream: for each sector, its allocation mask and
 * popcount(mask) bricks of 512 bytes in increasing brick index. */
int oracle_world_create(uint32_t log2_size_xz, uint32_t log2_size_y, uint32_t nsec_in,
                        const uint32_t* sector_idx, const uint64_t* alloc_mask,
                        const uint8_t* bricks, oracle_world** out);
void oracle_world_destroy(oracle_world* w);
void oracle_world_set_palette(oracle_world* w, const uint64_t* palette);
uint8_t oracle_world_get_voxel(const oracle_world* w, int32_t x, int32_t y, int32_t z);

/* Scalar cast of rays [begin, end). t[i] = -1 on a miss. pos, descents and uv (the hit's
 * texture coordinates, CpuRenderer.cpp:217-220) may be NULL. */
void oracle_cast(const oracle_world* w, const float* const o[3], const float* const d[3],
                 const int32_t world_origin[3], uint32_t begin, uint32_t end, uint32_t max_iters,
                 float* t, oracle_hit* rec, float* const pos[3], uint16_t* descents, float* const uv[2]);

/* GetStepPos (CpuRenderer.cpp:135-171) alone, for n in-bounds voxel positions v_in[3 * i]
 * and directions d[3 * i]: solid[i] = 1 when the voxel is solid (v_out = v_in), else v_out
 * is the far corner, in the ray direction, of the empty cell the step skips; descended[i] = 1
 * when the step read the brick's occupancy (the brick is allocated). */
void oracle_step_pos(const oracle_world* w, uint32_t n, const int32_t* v_in, const float* d, int32_t* v_out,
                     uint8_t* solid, uint8_t* descended);

/* Counter totals of a scalar cast (same definitions as the product's counters). */
void oracle_count(const oracle_world* w, const float* const o[3], const float* const d[3],
                  const int32_t world_origin[3], uint32_t n, uint32_t max_iters,
                  oracle_counters* out);

/* SIMD restatement of the reference's vector loop (CpuRenderer.cpp:172-224):
 * 16 lanes on AVX-512, else 8 lanes on AVX2, consecutive rays per vector,
 * `threads` worker threads (0 = all cores).  Results identical to oracle_cast
 * (pos, optional: the hit position relative to world_origin, 0 on a miss).
 * Returns the lane width used. */
int oracle_cast_simd(const oracle_world* w, const float* const o[3], const float* const d[3],
                     const int32_t world_origin[3], uint32_t n, uint32_t max_iters,
                     uint32_t threads, float* t, oracle_hit* rec, float* const pos[3]);
/* Lane width oracle_cast_simd would use on this host (16, 8, or 0 if unsupported). */
int oracle_simd_width(void);
int oracle_num_cores(void);

/* Scalar FP64 cast of VoxelMap::RayCast (VoxelMap.cpp:125-170). distance <= 0: miss. */
typedef struct oracle_hit_f64 {
    double distance;
    float normal[3];
    float uv[2];
    int32_t voxel[3];
} oracle_hit_f64;
void oracle_cast_f64(const oracle_world* w, const double origin[3], const double dir[3],
                     uint32_t max_iters, oracle_hit_f64* out);

/* Brute-force FP64 Amanatides-Woo DDA without LOD: first solid voxel along the
 * ray from (world_origin + o), up to max_steps voxel steps.  Returns 1 on a hit
 * and writes the voxel and the distance of the entry face; also returns the
 * smallest gap between the two nearest face crossings seen (ambiguity margin). */
int oracle_dda_f64(const oracle_world* w, const double o[3], const double d[3],
                   const int32_t world_origin[3], uint32_t max_steps,
                   int32_t voxel[3], double* t, double* min_gap);

/* ---- shader method: the GLSL traversal (src/VoxelRT/Shaders/VoxelTraversal.glsl) ----
 *
 * Contract (FP32, IEEE division; GLSL's a*b+c taken as contracted):
 *   inv = 1/d; tStart = (step(0,d) - o) * inv;
 *   clipRayToAABB (:133-145) against [1 - W, size - W - 1], per-axis min/max = fminf/fmaxf,
 *     start = fmaf(d, tmin, o) when 0 < tmin < tmax, else o;
 *   rayCast (:162-204): v = W + floor(start); each iteration side = fmaf(v - W, inv, tStart),
 *     tmin = bits(fminf(fminf(side))) + 5 ulp, cur = fmaf(tmin, d, o), v = W + floor(cur),
 *     bounds, getStepPos; cap 256; distance = biased tmin; iterations = loop index;
 *   rayCastCoarse (:206-243): o clipped first, tStart from the clipped origin, +0.001 bias,
 *     cap 96, coarse hits (findAnyOccupiedPos, :40-52) from loop index 31 on;
 *   getStepPos (:92-131): brick / voxel levels as the CPU path, plus the sector-of-sectors
 *     level (scale 32: LOD 32/64/128) when the sector mask is 0; with the ray-cell
 *     interaction LUT (:116-121) the mask is first ANDed with the octant's entry;
 *   normal = tmin >= side ? -sign(d) : 0 per axis (:194-196), in the code of vrt_hit;
 *   floor to int as floor2i.  A miss after the cap reports iterations = cap.
 * method: 0 = rayCast, 1 = rayCastCoarse.  max_iters 0 = the shader's cap. */
void oracle_ray_cell_lut(uint64_t table[512]);
void oracle_cast_shader(const oracle_world* w, const float* const o[3], const float* const d[3],
                        const int32_t world_origin[3], uint32_t begin, uint32_t end, uint32_t method,
                        uint32_t lut, uint32_t max_iters, float* t, oracle_hit* rec, float* const pos[3]);
/* One ray of the shader method; lut = the ray-cell table or NULL; cap > 0. */
void oracle_shader_ray(const oracle_world* w, const float o[3], const float d[3], const int32_t W[3],
                       uint32_t method, const uint64_t* lut, uint32_t cap, float* t, oracle_hit* rec, float pos[3]);

/* One frame of VoxelRender.comp:28-94 into G-buffer images, as vrt_render_frame
 * (include/vrt.h) states it: primary rays o, d given per pixel (row-major), casts
 * of the shader method, sun / bounce / sky shading, PCG bounce directions.
 * sun: the unit sun direction; pa, pb: the depth projection terms
 * -(f + n) / (f - n) and -(2 f n) / (f - n); fwd: the unit view axis. */
void oracle_render_frame(const oracle_world* w, uint32_t width, uint32_t height, const float* const o[3],
                         const float* const d[3], const int32_t W[3], const float fwd[3], const float sun[3],
                         float pa, float pb, uint32_t max_bounces, uint32_t frame, uint32_t lut,
                         const float sky[3], uint32_t* albedo_normal, float* irradiance, float* depth);

/* Internal: fills w->sector_groups from w->sector_masks. */
int oracle_world_build_groups(oracle_world* w);

/* PCG of RandomGen.glsl:6-20. */
uint32_t oracle_pcg_init(uint32_t px, uint32_t py, uint32_t seed);
uint32_t oracle_pcg_next(uint32_t* state);
/* sin and cos of 2*pi*u for u in [0, 1]. */
void oracle_sincos_turns(float u, float* s, float* c);

/* Bounce construction (CpuRenderer.cpp:389-392 with the PCG sample of
 * VoxelRender.comp:71): compacted bounce rays of the hits, in increasing i. */
uint32_t oracle_bounce(uint32_t n, const float* t, const oracle_hit* rec, const float* const pos[3],
                       const float* const d[3], uint32_t width, uint32_t seed,
                       float* const out_o[3], float* const out_d[3], uint32_t* src);

/* Backward of the Lambertian render in FP64: grad[v*3+c] += grad_image[p*3+c] * weight_p
 * over pixels p that hit voxel v; abssum accumulates |terms| (the error scale of an fp32 sum). */
void oracle_render_backward(uint32_t npix, const uint32_t* saved_voxel, const float* saved_weight,
                            const float* grad_image, double* grad, double* abssum);
/* L = sum_p sum_c (albedo[v_p][c] * weight_p - target_pc)^2 in FP64. */
double oracle_render_loss(uint32_t npix, const uint32_t* saved_voxel, const float* saved_weight,
                          const double* albedo, const float* target);

#ifdef __cplusplus
}
#endif
#endif
