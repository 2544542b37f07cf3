/*
 * vrt_scene.h - host-side synthetic worlds and rays for libvoxelrt_hip.so.
 *
 * These replace the reference's world producers (TerrainGenerator,
 * src/VoxelRT/TerrainGenerator.cpp:5-146, FastNoise2-based) with seeded,
 * fully specified generators, and its ray producers (GetPrimaryRay,
 * src/VoxelRT/CpuRenderer.cpp:226-233) with a pinhole camera built from
 * IEEE operations only, so host and device generate identical rays.
 * All pointers are host pointers; nothing here touches the GPU.
 */
#ifndef VRT_SCENE_H
#define VRT_SCENE_H

#include <stdint.h>

#include "vrt.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * A world as a sector stream (the input format of vrt_world_sync):
 * sectors in increasing index, each with its allocation mask and
 * popcount(mask) bricks of 512 bytes in increasing brick index.
 * Host memory owned by the vrt_scene.
 */
typedef struct vrt_sector_stream {
    uint32_t log2_size_xz, log2_size_y;
    uint32_t nsec;
    const uint32_t* sector_idx;
    const uint64_t* alloc_mask;
    const uint8_t* bricks;
    uint64_t nbricks;
    uint64_t solid_voxels;
} vrt_sector_stream;

typedef struct vrt_scene vrt_scene;

/*
 * Heightmap terrain.  Height H(x, z) = base + amp * fbm2(x, z), fbm2 = value
 * noise fBm (octaves, frequency, gain, lacunarity 2) normalised to [0, 1).
 * A voxel (x, y, z) is solid when y < H(x, z) and, if cave_threshold > 0, the
 * 3D value-noise fBm (cave_octaves, cave_frequency, gain 0.5) at the voxel
 * centre is above cave_threshold.  Material = 1 + (y / band) % 255.
 */
typedef struct vrt_terrain_params {
    uint32_t log2_size_xz, log2_size_y;
    uint64_t seed;
    uint32_t octaves;
    float frequency;
    float gain;
    float base, amp;          /* in voxels */
    float cave_threshold;     /* 0 = no caves */
    uint32_t cave_octaves;
    float cave_frequency;
    uint32_t band;            /* material band height in voxels */
} vrt_terrain_params;

int vrt_scene_terrain(const vrt_terrain_params* p, uint32_t threads, vrt_scene** out);

/*
 * Spheres: `count` centres uniform in [0, size), radius uniform in
 * [rmin, rmax), material 1 + (hash % 255); a voxel belongs to a sphere when
 * its centre is within the radius; later spheres overwrite earlier ones.
 */
int vrt_scene_spheres(uint32_t log2_size_xz, uint32_t log2_size_y, uint32_t count, float rmin, float rmax,
                      uint64_t seed, uint32_t threads, vrt_scene** out);

/* A scene from a dense voxel grid [y][z][x] of size_y * size_xz * size_xz bytes. */
int vrt_scene_from_dense(uint32_t log2_size_xz, uint32_t log2_size_y, const uint8_t* voxels, vrt_scene** out);

/* The scene's sector stream.  Edits made with vrt_scene_set_voxel or
 * vrt_scene_delete_empty_bricks are folded in first.  The pointers stay valid until the
 * scene is destroyed or the next call that folds new edits in (this one, vrt_scene_save_cvox,
 * vrt_world_from_scene after an edit): a caller that edits the scene takes the stream again.
 * A scene is not safe to use from several threads at once. */
int vrt_scene_get(const vrt_scene* scene, vrt_sector_stream* out);
/* Voxel at absolute coordinates (0 outside the world). */
uint8_t vrt_scene_voxel(const vrt_scene* scene, int32_t x, int32_t y, int32_t z);

/*
 * Editing the host world: the reference's VoxelMap (src/VoxelRT/VoxelMap.h:180-211).
 * The scene keeps a dirty set, VoxelMap::DirtyLocs (VoxelMap.h:185): sector index ->
 * mask of the bricks changed since the last vrt_world_sync_scene.
 *
 * vrt_scene_voxel is VoxelMap::Get (VoxelMap.h:191-194): 0 outside the scene or in an
 * absent brick.
 *
 * vrt_scene_set_voxel is VoxelMap::Set (VoxelMap.h:195-201): writes one voxel,
 * allocating its brick (zero-filled) when absent, also for value 0, and marks the
 * brick dirty (GetBrick(create, markAsDirty), VoxelMap.cpp:98-123).  A position
 * outside the scene's extents is ignored and returns VRT_OK, as Set ignores positions
 * outside the reference's sector grid.  O(1) per call: edited sectors are held aside
 * and folded into the sector stream by the next call that reads the stream
 * (vrt_scene_get, vrt_scene_save_cvox, vrt_world_from_scene).
 */
int vrt_scene_set_voxel(vrt_scene* scene, int32_t x, int32_t y, int32_t z, uint8_t value);

/* n writes in order, each exactly as vrt_scene_set_voxel: xyz holds x, y, z of write k at
 * 3k .. 3k + 2, values[k] its material (a region edit, as VoxelMap::RegionDispatchSIMD makes,
 * VoxelMap.h:223-263, without a call per voxel). */
int vrt_scene_set_voxels(vrt_scene* scene, uint64_t n, const int32_t* xyz, const uint8_t* values);

/* VoxelMap::MarkAllDirty (VoxelMap.h:207-211): every sector of the scene becomes dirty
 * with its whole allocation mask. */
int vrt_scene_mark_all_dirty(vrt_scene* scene);

/* Deletes the scene's bricks that hold no solid voxel and marks them dirty; a sector left
 * without bricks leaves the scene (the garbage collection of VoxelMap::RegionDispatchSIMD,
 * VoxelMap.h:236-263, over the whole scene; Sector::DeleteEmptyBricks, VoxelMap.cpp:67-79).
 * *deleted (may be NULL) receives the number of bricks deleted. */
int vrt_scene_delete_empty_bricks(vrt_scene* scene, uint64_t* deleted);

/* Number of sectors in the scene's dirty set. */
uint32_t vrt_scene_dirty_sectors(const vrt_scene* scene);

/*
 * The reference's world file, "cvox 0004" (VoxelMap::Deserialize / Serialize,
 * src/VoxelRT/VoxelMap.cpp:207-274; zstd blobs of BinaryIO.cpp:8-60): a palette
 * of 256 Material records and the sectors of a VoxelMap, indexed by
 * WorldSectorIndexer<12,8,signed> (VoxelMap.h:100, signed x and z in 12 bits,
 * y in 8).  Loading places world sector s at scene sector s - sector_offset
 * (NULL = 0) of a scene of the given extents and skips (and counts) sectors
 * outside it; the palette becomes the scene's, encoded as Material::GetEncoded
 * (VoxelMap.h:27-41: RGB565 | half(emission) << 16 | fuzziness << 32; the half
 * rounds to nearest even), and vrt_world_from_scene uploads it.  Saving writes a
 * scene in the same layout (scene sector s at world sector s + sector_offset,
 * increasing index, packs of at most 16 MiB plus a final one) with the given
 * 256 x 16-byte Material records (NULL = zeros).  zstd is libzstd.so.1, opened
 * at run time; without it both calls fail with VRT_ERR_STATE.
 */
typedef struct vrt_cvox_info {
    uint32_t sectors_in_file;
    uint32_t sectors_loaded;
    uint32_t sectors_outside;   /* non-empty sectors outside the scene's extents */
    uint64_t bricks_loaded;
} vrt_cvox_info;

int vrt_scene_load_cvox(const char* path, uint32_t log2_size_xz, uint32_t log2_size_y,
                        const int32_t sector_offset[3], vrt_scene** out, vrt_cvox_info* info);
int vrt_scene_save_cvox(const vrt_scene* scene, const int32_t sector_offset[3], const uint8_t* materials,
                        const char* path);
/* The scene's encoded palette (a loaded file's, or one set below); VRT_ERR_STATE for a
 * generated scene without one. */
int vrt_scene_get_palette(const vrt_scene* scene, uint64_t palette[256]);
/* Sets the scene's palette (VoxelMap::Palette, VoxelMap.h:187, encoded as
 * Material::GetEncoded): vrt_world_sync_scene uploads it with every sync. */
int vrt_scene_set_palette(vrt_scene* scene, const uint64_t palette[256]);
void vrt_scene_destroy(vrt_scene* scene);

/* Convenience: vrt_world_create + vrt_world_sync of the whole scene. */
int vrt_world_from_scene(const vrt_scene* scene, void* stream, vrt_world** out);

/*
 * GpuVoxelStorage::SyncBuffers(VoxelMap&) (src/VoxelRT/GpuRenderer.cpp:45-167): consumes
 * the scene's dirty set.  Every dirty sector is synced with the scene's current allocation
 * mask (0 for a sector the scene no longer holds) and its dirty bricks that are still
 * allocated (dirtyMask &= allocMask, GpuRenderer.cpp:62), through vrt_world_sync_bricks;
 * then the scene's palette, if it has one (a loaded file's or vrt_scene_set_palette's), is
 * uploaded, on every call as SyncBuffers does (GpuRenderer.cpp:100-102); without one the
 * world's palette is left as it is.  The world's extents must equal the scene's (the same sector
 * indices).  On success the dirty set is cleared; on failure it is kept.
 */
int vrt_world_sync_scene(vrt_world* world, vrt_scene* scene, void* stream);

/*
 * Camera set-up: eye at `pos` (absolute, double), yaw/pitch in degrees
 * (yaw 0 looks along +z, pitch > 0 looks up), vertical field of view fov_deg.
 * world_origin = floor(pos), eye = pos - world_origin.
 */
int vrt_camera_look(const double pos[3], double yaw_deg, double pitch_deg, double fov_deg, uint32_t width,
                    uint32_t height, uint32_t tile_w, uint32_t tile_h, vrt_camera* out);

/* Host generation of a camera's rays (bit-identical to vrt_camera_rays). out: 6 arrays of width*height floats. */
int vrt_camera_rays_host(const vrt_camera* cam, float* const out[6]);

/* n rays with origins uniform in [0, size) and directions uniform on the sphere; world_origin 0. */
int vrt_incoherent_rays_host(uint32_t n, uint32_t size_xz, uint32_t size_y, uint64_t seed, float* const out[6]);

/* Host twin of vrt_bounce_generate for checking it.  Returns the bounce count via *count. */
int vrt_bounce_host(uint32_t n, const vrt_hit* rec, const float* pos /* 3*n SoA */, uint32_t width,
                    uint32_t seed, float* const out[6], uint32_t* src, uint32_t* count);

#ifdef __cplusplus
}
#endif
#endif
