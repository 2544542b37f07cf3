/*
 * vrt.h - C ABI of the MI355X brickmap ray traversal (libvoxelrt_hip.so).
 *
 * The world is a two-level brickmap: a sector is 4x4x4 bricks (32^3 voxels),
 * a brick is 8x8x8 voxels of one byte each (0 = empty, 1..255 = material id).
 * On the device it is held as structure-of-arrays buffers in HBM:
 *   sectors       16 B [nsec]      allocation mask (bit b = brick b allocated) and
 *                                  the slot of the sector's first allocated brick
 *   occupancy     u64 [slots * 8]  per brick, one mask per 4x4x4 voxel cell
 *   voxels        u8  [slots * 512]
 *   palette       u64 [256]
 * Brick b of a sector lives in slot base + popcount(mask & ((1 << b) - 1)).
 *
 * Every entry point returns a vrt_status; on failure vrt_last_error() gives a
 * thread-local message.  No C++ exception crosses this boundary.  Device work
 * is asynchronous on the given stream (a hipStream_t, NULL = the null stream);
 * buffers passed to casts are device pointers owned by the caller.  A call with
 * nothing to do (no rays, views or pixels) checks its options, launches nothing
 * and returns VRT_OK; its buffer pointers may then be NULL.  A world lives on the
 * device current at vrt_world_create; calls that take the world (or a batch plan
 * of it) need that device current, and return VRT_ERR_STATE otherwise.
 *
 * Reference interfaces replaced (paths relative to the VoxelRT source tree):
 *   vrt_world_*            GpuVoxelStorage / FlatVoxelStorage SyncBuffers
 *                          (src/VoxelRT/GpuRenderer.cpp:17-167,
 *                           src/VoxelRT/CpuRenderer.cpp:20-84)
 *   vrt_cast               RayCast / GetStepPos / GetVoxelMaterial
 *                          (src/VoxelRT/CpuRenderer.cpp:110-224)
 *   vrt_cast_one_f64       HitResult VoxelMap::RayCast(dvec3, dvec3, uint32_t)
 *                          (src/VoxelRT/VoxelMap.h:213-214)
 *   vrt_bounce_generate    diffuse bounce construction in RenderRow
 *                          (src/VoxelRT/CpuRenderer.cpp:342-392,
 *                           src/VoxelRT/Shaders/VoxelRender.comp:63-86);
 *                          vrt_bounce_generate_ex also groups the rays by direction
 *   vrt_cast_shader        rayCast / rayCastCoarse of the GPU renderer
 *                          (src/VoxelRT/Shaders/VoxelTraversal.glsl:92-243)
 *   vrt_render_frame       GpuRenderer::RenderFrame's trace and shading
 *                          (src/VoxelRT/GpuRenderer.cpp:239-273,
 *                           src/VoxelRT/Shaders/VoxelRender.comp:28-94)
 * vrt_stream_release frees the per-stream state of the persistent and compact cast
 * schedules; it has no counterpart in the reference.
 * The C++ interface of the reference (HitResult, VoxelMap::RayCast, a Renderer's
 * RenderFrame) sits over these calls in vrt.hpp; world files and seeded worlds
 * are in vrt_scene.h.
 */
#ifndef VRT_H
#define VRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VRT_ABI_VERSION 2

/* ---------------------------------------------------------------- errors */

typedef enum vrt_status {
    VRT_OK = 0,
    VRT_ERR_INVALID_ARGUMENT = 1, /* null or short buffer, bad extent, n = 0 where n > 0 is needed */
    VRT_ERR_CAPACITY = 2,         /* brick slots exhausted ("Could not allocate brick slots") */
    VRT_ERR_HIP = 3,              /* a HIP runtime call failed */
    VRT_ERR_NO_DEVICE = 4,        /* no usable gfx950 device */
    VRT_ERR_STATE = 5             /* call not valid in the object's current state */
} vrt_status;

/* Message of the last failed call on this thread ("" if none). */
const char* vrt_last_error(void);
int vrt_abi_version(void);

/* ---------------------------------------------------------------- world */

typedef struct vrt_world vrt_world;

typedef struct vrt_world_desc {
    uint32_t log2_size_xz;   /* voxel extent along x and z is 1 << log2_size_xz, 5..14 */
    uint32_t log2_size_y;    /* voxel extent along y is 1 << log2_size_y, 5..14 */
    uint32_t initial_slots;  /* device brick slots reserved up front (0 = grow on demand) */
    uint32_t max_slots;      /* slot arena capacity (0 = 128 per sector: twice the bricks the grid holds, so a
                                sector's run of up to 64 slots can move even in a nearly full world; device
                                memory follows the highest slot in use, not this bound); at most 2^28 */
} vrt_world_desc;

/* Shape of a created world (all derived from the descriptor). */
typedef struct vrt_world_info {
    uint32_t size_xz, size_y;      /* voxel extents */
    uint32_t sectors_xz, sectors_y;/* sector grid (32 voxels per sector edge) */
    uint32_t num_sectors;
    uint32_t num_bricks;           /* bricks currently allocated */
    uint32_t slot_capacity;        /* device slots currently backed by memory */
    uint64_t device_bytes;         /* HBM held by the world buffers */
    uint32_t slot_high_water;      /* one past the highest allocated slot: per-voxel arrays in
                                      slot order (the render's albedo) need this many * 512 entries */
    uint32_t generation;           /* changes whenever the occupancy and voxel buffers move (slot
                                      growth): device pointers taken before are then stale */
    uint64_t upload_bytes;         /* host-to-device bytes sent by syncs since creation */
    uint64_t upload_brick_bytes;   /* of those, brick bytes: 512 per uploaded brick */
    uint64_t moved_bricks;         /* bricks moved between slots on the device (contents kept) */
} vrt_world_info;

int vrt_world_create(const vrt_world_desc* desc, void* stream, vrt_world** out);
int vrt_world_destroy(vrt_world* world);
int vrt_world_get_info(const vrt_world* world, vrt_world_info* out);

/*
 * Update a set of sectors: the DirtyLocs protocol of GpuRenderer.cpp:45-167, where
 * the caller records DirtyLocs[sectorIdx] |= brickMask (VoxelMap.h:185) and the
 * renderer's SyncBuffers consumes it.  For each i < nsec:
 *   sector_idx[i]  linear sector index x | z << sx | y << 2sx (sx = log2 sectors_xz)
 *   alloc_mask[i]  bricks present in the sector after the update (0 = sector emptied)
 *   dirty_mask[i]  bricks whose bytes follow in `bricks` (only its bits inside
 *                  alloc_mask count, as dirtyMask &= allocMask, GpuRenderer.cpp:62);
 *                  a brick newly set in alloc_mask must be dirty
 * bricks holds, sector after sector in the order given, the dirty bricks of 512
 * bytes each in increasing brick index; brick voxel order is x | z << 3 | y << 6.
 * Slots are (re)allocated on the host; only dirty bricks are uploaded and have
 * their occupancy masks rebuilt on the device.  Bricks that stay allocated and
 * are not dirty keep their device contents: when the sector's slot run moves or
 * their slot shifts with the allocation, they are moved on the device (the
 * reference re-uploads the whole sector then, BrickSlotAllocator.cpp:7-23).
 * Uploads go through staging buffers the world owns (8 MiB pinned halves, kept
 * from the first sync), so the call allocates nothing after its first use except
 * when slots or moves outgrow their buffers, and does not wait for the device
 * unless a staging half is still in use.  Host pointers; the call returns once
 * the host buffers may be reused, the device work is ordered on `stream`, and
 * casts on other streams must not overlap it.  On VRT_ERR_CAPACITY or
 * VRT_ERR_INVALID_ARGUMENT the world is unchanged; a HIP failure once device
 * work has begun leaves the world failed (every later call on it returns
 * VRT_ERR_STATE).  Bricks that are entirely empty may be listed; they stay
 * allocated.
 */
int vrt_world_sync_bricks(vrt_world* world, uint32_t nsec, const uint32_t* sector_idx,
                          const uint64_t* alloc_mask, const uint64_t* dirty_mask,
                          const uint8_t* bricks, void* stream);

/* vrt_world_sync_bricks with every brick of each sector dirty: bricks holds
 * popcount(alloc_mask[i]) bricks per sector (MarkAllDirty, VoxelMap.h:207-211). */
int vrt_world_sync(vrt_world* world, uint32_t nsec, const uint32_t* sector_idx,
                   const uint64_t* alloc_mask, const uint8_t* bricks, void* stream);

/* Palette entries (Material::GetEncoded layout, VoxelMap.h:27-41). Host pointer, 256 entries. */
int vrt_world_set_palette(vrt_world* world, const uint64_t* palette, void* stream);

/* One sector (GpuRenderer.cpp:27-33 AllocMasks + BaseSlots, packed so a lookup is one 16-byte load). */
typedef struct vrt_sector_record {
    uint64_t mask;       /* bit b: brick b (x | z << 2 | y << 4 in the sector) is allocated */
    uint32_t base_slot;  /* slot of the first allocated brick */
    uint32_t empty_group_cell; /* 15, or 31 when mask is 0: the cell mask (edge - 1) of a step that
                                  stays at the sector level where the 2x2x2 brick group is empty,
                                  i.e. the LOD of GetStepPos (CpuRenderer.cpp:158-160) precomputed
                                  per sector; the step takes half when the group holds a brick */
} vrt_sector_record;

/* Device views of the world buffers (read-only use by callers). */
typedef struct vrt_world_buffers {
    const vrt_sector_record* sectors; /* [num_sectors] */
    const uint64_t* occupancy;    /* [slot_capacity * 8] */
    const uint8_t* voxels;        /* [slot_capacity * 512] */
    const uint64_t* palette;      /* [256] */
} vrt_world_buffers;
int vrt_world_get_buffers(const vrt_world* world, vrt_world_buffers* out);

/* ---------------------------------------------------------------- rays and hits */

/* n rays in structure-of-arrays form; origins are relative to world_origin. */
typedef struct vrt_rays {
    const float* ox; const float* oy; const float* oz;
    const float* dx; const float* dy; const float* dz;
    int32_t world_origin[3];
    uint32_t n;
} vrt_rays;

/*
 * Hit record, 16 bytes per ray.
 *   x, y, z  absolute coordinates of the hit voxel (0 on a miss)
 *   info     bits 0-7   material id of the hit voxel (0 = miss)
 *            bits 8-13  face normal, 2 bits per axis (x, y, z) in two's
 *                       complement: 0 -> 0, 1 -> +1, 3 -> -1
 *            bits 16-31 traversal iterations (mask lookups) spent on the ray
 */
typedef struct vrt_hit {
    int32_t x, y, z;
    uint32_t info;
} vrt_hit;

typedef struct vrt_hits {
    float* t;        /* [n] distance along d to the entry face; -1 on a miss */
    vrt_hit* rec;    /* [n] */
    float* pos;      /* optional [3 * n] SoA hit position relative to world_origin (NULL = skip) */
    /* VHitResult's other fields (CpuRenderer.cpp:88-108, 210-221), filled by vrt_cast and
     * vrt_cast_ex only (the other entry points reject them):
     *   uv             optional [2 * n] SoA: fract(X ? pos.y : pos.x), fract(Z ? pos.y : pos.z)
     *                  with X / Z the normal's tie masks and fract(x) = x - floor(x); 0 on a miss
     *   material_data  optional [n]: the low 32 bits of the hit material's palette entry
     *                  (vrt_world_set_palette), 0 on a miss */
    float* uv;
    uint32_t* material_data;
} vrt_hits;

/* Totals over a launch, accumulated into device memory when requested. */
typedef struct vrt_cast_counters {
    unsigned long long rays;      /* rays traced */
    unsigned long long iters;     /* sector-mask lookups (GetStepPos evaluations) */
    unsigned long long descents;  /* lookups that descended into an allocated brick */
    unsigned long long hits;      /* rays that hit a voxel */
} vrt_cast_counters;

/*
 * Plain per-ray cast: the drop-in for RayCast (CpuRenderer.cpp:172-224),
 * FP32, un-biased t, +0.001 step bias, 2-level mask LOD.  max_iters is the
 * iteration cap (the reference uses 128), 1..65535.
 */
int vrt_cast(const vrt_world* world, const vrt_rays* rays, uint32_t max_iters,
             vrt_hits* hits, void* stream);

/* Kernel variants of the same cast (results are identical for every variant). */
typedef enum vrt_cast_schedule {
    VRT_SCHED_AUTO = 0,        /* the default (GRID) */
    VRT_SCHED_GRID = 1,        /* one ray per lane, one lane per ray, grid covers n */
    VRT_SCHED_PERSISTENT = 2,  /* persistent waves pull ray chunks from a queue and refill idle lanes;
                                  the queue counters belong to the launch's (device, stream) pair, so
                                  persistent casts on different streams never share them */
    VRT_SCHED_COMPACT = 3      /* cap and compact: a grid pass stops every ray after pass_cap steps;
                                  the rays still running are packed 64 to a wave and resumed from
                                  their saved state by a second pass.  For incoherent ray sets, whose
                                  lanes finish at very different steps.  The survivor state (24 bytes
                                  per ray of capacity and a 64-bit ballot per 64 rays) belongs to the
                                  (device, stream) pair; the first call on a stream, and a call with
                                  more rays than any before on it, allocates it (with a stream
                                  synchronization), so make one such call before capturing a graph.
                                  A captured compact cast keeps using its capture stream's state:
                                  replays must not overlap compact casts on that stream. */
} vrt_cast_schedule;

/* Order in which the grid schedule dispatches its blocks of 64 rays (results are identical). */
typedef enum vrt_cast_order {
    VRT_ORDER_FORWARD = 0,     /* block 0 first */
    VRT_ORDER_REVERSE = 1      /* last block first: cheap rays at the start of the buffer (a camera's
                                  sky rows) then run in the launch's tail, where the chip drains */
} vrt_cast_order;

typedef struct vrt_cast_opts {
    uint32_t schedule;          /* vrt_cast_schedule */
    uint32_t refill;            /* persistent: free lanes that trigger a retire/refill, 1..64 (0 = 16) */
    uint32_t chunk;             /* persistent: rays per queue grab, multiple of 64 (0 = 64) */
    uint32_t waves_per_cu;      /* persistent occupancy target (0 = default) */
    vrt_cast_counters* counters;/* optional device pointer, accumulated (not cleared) */
    const uint32_t* count;      /* optional device pointer: rays to trace, read when the kernel
                                   runs (rays->n is then the capacity and the upper bound) */
    uint32_t order;             /* grid and compact (first pass): vrt_cast_order (0 = forward) */
    uint32_t pass_cap;          /* compact: steps of the first pass (0 = 32); a cap at or above
                                   max_iters makes the cast one grid pass */
} vrt_cast_opts;

int vrt_cast_ex(const vrt_world* world, const vrt_rays* rays, uint32_t max_iters,
                vrt_hits* hits, const vrt_cast_opts* opts, void* stream);

/*
 * Frees the state the library keeps per (device, stream) for `stream` on every device: the
 * persistent schedule's queue counters and the compact schedule's survivor state.  It
 * synchronizes the stream first.  Call it before destroying a stream that ran such casts;
 * a later cast on the stream allocates the state again.  A stream without state is a
 * no-op.  No counterpart in the reference (its casts keep no per-stream state).
 */
int vrt_stream_release(void* stream);

/*
 * Scalar FP64 cast of one ray on the device, the semantics of VoxelMap::RayCast
 * (VoxelMap.cpp:125-170): bias +0.0001, fixed 32/8/1 levels, normal from
 * tmin >= sideDist, distance = biased tmin.  Synchronous; host in/out.
 * Returns the hit in *out (distance <= 0 means a miss).
 */
typedef struct vrt_hit_f64 {
    double distance;           /* -1 on a miss */
    float normal[3];
    float uv[2];
    int32_t voxel[3];
} vrt_hit_f64;
int vrt_cast_one_f64(const vrt_world* world, const double origin[3], const double dir[3],
                     uint32_t max_iters, vrt_hit_f64* out);

/* ---------------------------------------------------------------- shader method */

/*
 * The GLSL traversal (src/VoxelRT/Shaders/VoxelTraversal.glsl) on the same world:
 * a different method from vrt_cast, not bit-compatible with it.  Replaces the
 * shader's rayCast (:162-204) and rayCastCoarse (:206-243):
 *   - the ray starts at its origin clipped to the world box inset by one voxel
 *     (clipRayToAABB, :133-145) and first steps out of that voxel;
 *   - getStepPos (:92-131) adds a sector-of-sectors level: a sector whose brick
 *     mask is 0 steps 32/64/128 voxels from the 4x4x4-sector group mask;
 *   - ray_cell_lut = 1 first ANDs each mask with the ray-cell interaction table
 *     of the ray's octant (u_UseAnisotropicLods, :116-121; GpuRenderer.cpp:192-210);
 *   - full: +5 ulp bias on tmin, cap 256; coarse: +0.001 bias, cap 96, and from
 *     the 31st iteration a non-empty 4x4x4 voxel cell counts as a hit at its first
 *     solid voxel (findAnyOccupiedPos, :40-52);
 *   - t is the biased tmin; normal = -sign(d) on every axis with tmin >= side;
 *     iterations = the loop index at the hit or miss (the cap when exhausted).
 * FP32 with IEEE division, a*b+c as fmaf, min/max as fminf/fmaxf.
 */
typedef enum vrt_shader_method {
    VRT_SHADER_FULL = 0,    /* rayCast */
    VRT_SHADER_COARSE = 1   /* rayCastCoarse (shadow and bounce rays) */
} vrt_shader_method;

typedef struct vrt_shader_opts {
    uint32_t method;        /* vrt_shader_method */
    uint32_t ray_cell_lut;  /* 1 = mask with the ray-cell interaction table */
    uint32_t max_iters;     /* 0 = the shader's cap (256 full, 96 coarse), else 1..65535 */
    uint32_t reserved;      /* 0 */
} vrt_shader_opts;

/* At most 2^32 - 64 rays per call. */
int vrt_cast_shader(const vrt_world* world, const vrt_rays* rays, const vrt_shader_opts* opts,
                    vrt_hits* hits, void* stream);

/* ---------------------------------------------------------------- bounce rays */

/*
 * Diffuse bounce rays for the hits of one primary set (CpuRenderer.cpp:389-392,
 * VoxelRender.comp:63-71): for every hit i, origin = pos_i + N_i * 0.01,
 * direction = normalize(N_i + s), s uniform on the unit sphere from the PCG
 * stream random_init(uvec3(i % width, i / width, seed)) (RandomGen.glsl:6-20);
 * a degenerate N + s = 0 falls back to N.  Bounce rays are written compacted
 * in increasing i; src[k] is the primary index of bounce k; *count (device)
 * receives the number written.  hits must have pos.  scratch: device buffer of
 * vrt_bounce_scratch_bytes(n) bytes.
 */
typedef struct vrt_ray_buffers {
    float* ox; float* oy; float* oz;
    float* dx; float* dy; float* dz;
} vrt_ray_buffers;

size_t vrt_bounce_scratch_bytes(uint32_t n);
int vrt_bounce_generate(const vrt_rays* primary, const vrt_hits* hits, uint32_t width,
                        uint32_t seed, vrt_ray_buffers* out, uint32_t* src, uint32_t* count,
                        void* scratch, void* stream);

/*
 * vrt_bounce_generate with a choice of output order (the rays, src entries and count are the
 * same; only their order differs):
 *   VRT_BOUNCE_PRIMARY_ORDER    increasing primary index (vrt_bounce_generate);
 *   VRT_BOUNCE_DIRECTION_ORDER  grouped by the direction's octant, key 0..7 (bit a set when
 *                               component a of N + s is >= 0, read off the signs of its terms:
 *                               the normal's on its axis, elsewhere s's from its two PCG
 *                               samples; the octant of d except at sign boundaries), in
 *                               increasing primary index within a key.  Neighbouring bounce
 *                               rays then share a direction octant as well as a region, which
 *                               makes a following cast of them faster.  At most 8388608
 *                               primaries.
 * No counterpart in the reference (its bounce rays are traced per pixel).
 */
typedef enum vrt_bounce_order {
    VRT_BOUNCE_PRIMARY_ORDER = 0,
    VRT_BOUNCE_DIRECTION_ORDER = 1
} vrt_bounce_order;

int vrt_bounce_generate_ex(const vrt_rays* primary, const vrt_hits* hits, uint32_t width,
                           uint32_t seed, uint32_t order, vrt_ray_buffers* out, uint32_t* src,
                           uint32_t* count, void* scratch, void* stream);

/*
 * vrt_bounce_generate without a workspace, and allocating nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The per-block hit counts live in a
 * static device array of 8 KiB (2048 counts, one per 4096 primaries) that every GPU the
 * library runs on holds until the process ends; its address is looked up once per device, on
 * the first call with that device current.  At most 8388608 primaries (a 3840x2160 frame
 * fits); more is VRT_ERR_INVALID_ARGUMENT.  Each call writes every count it reads, so nothing
 * needs clearing between calls.  All calls on one GPU share the array: calls that overlap in
 * time (on different streams) race.  Results are identical to vrt_bounce_generate's.
 */
int vrt_bounce_generate_noalloc(const vrt_rays* primary, const vrt_hits* hits, uint32_t width,
                                uint32_t seed, vrt_ray_buffers* out, uint32_t* src, uint32_t* count,
                                void* stream);

/* ---------------------------------------------------------------- cameras */

/*
 * Pinhole camera.  Ray for pixel (px, py), before tiling:
 *   u = (px + 0.5) * inv_w * 2 - 1,  v = 1 - (py + 0.5) * inv_h * 2
 *   d = normalize(forward + u * right + v * up)   (right/up pre-scaled by the FOV)
 * a zero component of d is replaced by 2^-24; origin = eye (relative to
 * world_origin).  Rays are numbered tile by tile: tiles of tile_w x tile_h
 * pixels in row-major tile order, pixels row-major inside a tile.
 */
typedef struct vrt_camera {
    int32_t world_origin[3];
    float eye[3];
    float forward[3];
    float right[3];
    float up[3];
    uint32_t width, height;
    uint32_t tile_w, tile_h;   /* 1x1 = plain row-major */
} vrt_camera;

/* Generate the primary rays of a camera into device buffers (n = width*height, at most
 * 2^32 - 256). */
int vrt_camera_rays(const vrt_camera* cam, vrt_ray_buffers* out, void* stream);

/* ---------------------------------------------------------------- fused frame */

/*
 * Primary ray plus one diffuse bounce per hit, traced per pixel in one kernel: the structure of
 * the reference GPU renderer (Shaders/VoxelRender.comp:28-94, one invocation per pixel traces
 * the primary ray, then its bounce), without shading.  `primary` receives exactly what vrt_cast
 * returns for the camera's rays (vrt_camera_rays); `bounce` is indexed by pixel: entry i holds
 * the hit of the bounce ray vrt_bounce_generate builds for primary hit i (width = camera width,
 * `seed`), or, when primary i missed, a record with info = VRT_HIT_NO_RAY and t = -1.
 * primary->pos is optional.  Schedules: VRT_SCHED_GRID (one lane per pixel) or
 * VRT_SCHED_PERSISTENT (waves refill lanes in batches of `refill`: primary rays are built from
 * the camera in registers, so a refill waits on no load).
 */
#define VRT_HIT_NO_RAY (1u << 14)

typedef struct vrt_frame_opts {
    uint32_t schedule;            /* VRT_SCHED_GRID (default) or VRT_SCHED_PERSISTENT */
    uint32_t refill;              /* persistent: parked lanes that trigger a retire/refill (0 = 16) */
    uint32_t chunk;               /* persistent: pixels per queue grab, multiple of 64 (0 = 256) */
    vrt_cast_counters* counters;  /* optional device totals over primary and bounce rays */
    uint32_t order;               /* grid: vrt_cast_order over blocks of 64 pixels (0 = forward) */
} vrt_frame_opts;

int vrt_cast_frame(const vrt_world* world, const vrt_camera* cam, uint32_t max_iters, uint32_t seed,
                   vrt_hits* primary, vrt_hits* bounce, const vrt_frame_opts* opts, void* stream);

/* ---------------------------------------------------------------- renderer */

/*
 * One frame of the reference GPU renderer into G-buffer images: the body of
 * VoxelRender.comp:28-94 per pixel, the work GpuRenderer::RenderFrame
 * (GpuRenderer.cpp:239-273) dispatches in 8x8 groups.  The primary ray is
 * traced with the shader method's rayCast (vrt_cast_shader); on a hit the
 * palette gives the albedo (RGB565 / (31, 63, 31), squared) and the emission
 * (the half in bits 16..31); with max_bounces > 0 a coarse shadow ray toward the
 * sun normalize(0.3, 0.9, -0.28) adds (1.2, 1.1, 1.0) * 5 or halves the
 * throughput, then up to max_bounces coarse diffuse bounces (the first two also
 * test the sun) accumulate emission; a miss adds the sky.  Deviations from the
 * shader, all stated: rays come from vrt_camera's pinhole model (pixel centres,
 * image row-major, tile fields ignored), the sky is a constant radiance instead
 * of a cubemap, bounce directions take PCG samples of random_init(uvec3(px, py,
 * frame)) instead of blue noise, bounce directions get the zero nudge of
 * vrt_camera rays, irradiance is fp32 instead of fp16, and a ray that exhausts
 * its cap reports the cap as its iterations.
 * Depth is (pa * -q + pb) / q with q = dot(forward, pos / 16) (pos relative to
 * world_origin), pa = -(f + n) / (f - n), pb = -(2 f n) / (f - n), the z row of
 * glm::perspective (Camera.h:34); forward must be the unit view axis.
 */
typedef struct vrt_render_opts {
    uint32_t max_bounces;   /* u_MaxBounces, 0..16 */
    uint32_t frame;         /* u_FrameNo: seeds the bounce directions */
    uint32_t ray_cell_lut;  /* u_UseAnisotropicLods: 0 or 1 */
    uint32_t order;         /* vrt_cast_order of the 8x8 pixel tiles (same images) */
    float sky[3];           /* sky radiance */
    float near_z, far_z;    /* depth projection; both 0 = glim::Camera's 0.01 / 1000 */
} vrt_render_opts;

typedef struct vrt_gbuffer {
    uint32_t* albedo_normal;  /* [height][width] RGBA8 (u_AlbedoNormalTex): albedo as unorm8 round(clamp(x) * 255),
                                 alpha = packGNormal: (nx + 1) | (ny + 1) << 2 | (nz + 1) << 4, 21 on a miss */
    float* irradiance;        /* [height][width][4] (u_IrradianceTex): rgb, iterations of the last traced ray */
    float* depth;             /* [height][width] (u_DepthTex), -1 on a miss */
} vrt_gbuffer;

int vrt_render_frame(const vrt_world* world, const vrt_camera* cam, const vrt_render_opts* opts,
                     vrt_gbuffer* out, void* stream);

/* ---------------------------------------------------------------- batched */

typedef enum vrt_segment_kind {
    VRT_SEG_CAMERA = 0,  /* rays generated in-kernel from a vrt_camera */
    VRT_SEG_RAYS = 1     /* rays read from buffers (e.g. bounce sets) */
} vrt_segment_kind;

/*
 * One segment of a batch.  Hits of segment s are written at hit_offset .. +count.
 * For VRT_SEG_RAYS, rays are read from the batch's ray buffers at ray_offset,
 * with the segment's own world origin.  count_ptr, when non-NULL, is a device
 * u32 that overrides count at run time (count is then the capacity).
 */
typedef struct vrt_segment {
    uint32_t kind;
    uint32_t camera;          /* index into the camera table (VRT_SEG_CAMERA) */
    uint32_t ray_offset;      /* first ray in the ray buffers (VRT_SEG_RAYS) */
    uint32_t hit_offset;      /* first hit in the output */
    uint32_t count;           /* rays in the segment (capacity when count_ptr is set) */
    int32_t world_origin[3];  /* VRT_SEG_RAYS */
    const uint32_t* count_ptr;
} vrt_segment;

typedef struct vrt_batch {
    const vrt_camera* cameras; uint32_t ncameras;  /* host; each camera's tile fixes its hit order */
    const vrt_segment* segments; uint32_t nsegments; /* host */
    vrt_rays rays;           /* buffers for VRT_SEG_RAYS segments (world_origin, n unused) */
    vrt_hits hits;           /* output, indexed by hit_offset; pos (optional) is [3][H] with
                                H = max over segments of hit_offset + count */
    uint32_t max_iters;
} vrt_batch;

typedef struct vrt_batch_tuning {
    uint32_t tile_w, tile_h;   /* lane -> pixel tile inside camera segments: 8x8, 16x4, 4x16 or 64x1
                                  (camera sizes must be multiples of it); the hit order does not change */
    uint32_t rays_per_lane;    /* 1 or 2 rays traced interleaved per lane */
    uint32_t chunk;            /* 0: one wave per 64*rays_per_lane rays, balanced by the hardware
                                  dispatcher; else persistent waves claim this many rays per grab */
} vrt_batch_tuning;

typedef struct vrt_batch_plan vrt_batch_plan;

/* Build a plan (device segment/camera tables, queue scratch) for a batch. */
int vrt_batch_plan_create(const vrt_world* world, const vrt_batch* batch,
                          const vrt_batch_tuning* tuning, void* stream, vrt_batch_plan** out);
/* Trace every segment of the plan in one persistent launch. */
int vrt_batch_launch(vrt_batch_plan* plan, vrt_cast_counters* counters, void* stream);
/* Capture vrt_batch_launch into a hipGraph (once) and replay it. */
int vrt_batch_graph_launch(vrt_batch_plan* plan, void* stream);
int vrt_batch_plan_destroy(vrt_batch_plan* plan);

/*
 * Time every tuning in the search space on this batch (tile 8x8/16x4, rays per
 * lane 1/2, chunk 0/64/256) and return the fastest.  Results are cached per
 * device name and batch signature in-process and, when cache_path is not NULL,
 * in that text file.
 */
int vrt_batch_autotune(const vrt_world* world, const vrt_batch* batch, const char* cache_path,
                       vrt_batch_tuning* best, float* best_ms, void* stream);

/* ---------------------------------------------------------------- differentiable */

/*
 * Lambertian render of camera views, one primary ray per pixel, no shadows:
 *   radiance_c = albedo[v][c] * max(0, dot(N, L)) * intensity   (0 on a miss)
 * albedo is fp32 RGB per voxel in slot order: index (slot * 512 + voxel) * 3 + c; it
 * holds at least vrt_world_info.slot_high_water * 512 voxels (the library cannot check
 * this: the array's length is not passed).
 * image: [sum of width*height over views] * 3, views concatenated, pixels in
 * each camera's ray order.  saved_voxel (u32, ~0 on a miss) and saved_weight
 * (f32) are per pixel and are consumed by the backward pass.  At most 2^32 - 4096
 * pixels over all views.
 */
int vrt_render_fwd(const vrt_world* world, const vrt_camera* cameras, uint32_t nviews,
                   uint32_t max_iters, const float light_dir[3], float intensity,
                   const float* albedo, float* image, uint32_t* saved_voxel,
                   float* saved_weight, void* stream);

/* grad_albedo[v][c] += sum over pixels hitting v of grad_image[p][c] * saved_weight[p].
 * grad_albedo has the albedo's length; the world is unchanged since the forward pass.
 * npixels is at most 2^32 - 256.  A scatter-add with wave-aggregated float atomics. */
int vrt_render_bwd(const vrt_world* world, uint32_t npixels, const float* grad_image,
                   const uint32_t* saved_voxel, const float* saved_weight,
                   float* grad_albedo, void* stream);

/*
 * The same gradient, every one of the nvox * 3 entries written (no zeroing beforehand):
 * grad_albedo[v][c] = sum over pixels hitting v of grad_image[p][c] * saved_weight[p], 0 for
 * voxels no pixel hit.  Hit pixels are bucketed by brick slot (a counting sort), then each
 * slot's contributions are summed in LDS and its 512 rows written once; no global float
 * atomics.  nvox >= vrt_world_info.slot_high_water * 512 (the albedo's length); grad_albedo
 * 16-byte aligned; workspace (256-byte aligned) of vrt_render_bwd_workspace_bytes(npixels,
 * nvox) bytes.  The summation order within a voxel is not fixed (LDS atomics).
 */
size_t vrt_render_bwd_workspace_bytes(uint32_t npixels, uint64_t nvox);
int vrt_render_bwd_bricks(const vrt_world* world, uint32_t npixels, const float* grad_image,
                          const uint32_t* saved_voxel, const float* saved_weight, uint64_t nvox,
                          float* grad_albedo, void* workspace, size_t workspace_bytes, void* stream);

/*
 * The gradient as npixels unsummed entries (a sparse COO gradient with repeated indices):
 * index[p] = the voxel pixel p hit and values[p][c] = grad_image[p][c] * saved_weight[p], or
 * index 0 and values 0 for a pixel that contributes nothing (a miss or zero weight).  Summing
 * the values by index gives vrt_render_bwd's gradient.
 */
int vrt_render_bwd_sparse(const vrt_world* world, uint32_t npixels, const float* grad_image,
                          const uint32_t* saved_voxel, const float* saved_weight, int64_t* index,
                          float* values, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VRT_H */
