// vrt.hpp - C++ side of the drop-in: the reference's voxel-map / ray-cast
// interface over the C ABI of libvoxelrt_hip.so (vrt.h).  Header only, C++17.
//
//   vrt::HitResult      HitResult (src/VoxelRT/VoxelMap.h:171-178)
//   vrt::World          the device world that GpuVoxelStorage / FlatVoxelStorage
//                       hold (GpuRenderer.cpp:17-167), synced sector by sector as
//                       SyncBuffers consumes VoxelMap::DirtyLocs
//   vrt::VoxelMapCaster VoxelMap::RayCast(dvec3 origin, dvec3 dir, uint32_t maxIters = 1024)
//                       (VoxelMap.h:213-214), the picking cast of Main.cpp:248 and
//                       Brush.cpp:47
//   vrt::HipRenderer    GpuRenderer::RenderFrame (GpuRenderer.cpp:239-273) into G-buffer
//                       images, for the Renderer interface of Renderer.h:15-20
//
// Errors become vrt::Error exceptions (the reference throws std::runtime_error,
// e.g. BrickSlotAllocator.cpp:15); vrt::CapacityError for exhausted brick slots.
// Vector arguments are any type with x, y, z members (glm::dvec3 fits).
#pragma once

#include <stdexcept>
#include <string>
#include <utility>

#include "vrt.h"

namespace vrt {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};
struct CapacityError : Error {
    using Error::Error;
};

inline void check(int rc) {
    if (rc == VRT_OK) return;
    std::string msg = vrt_last_error();
    if (rc == VRT_ERR_CAPACITY) throw CapacityError(rc, msg);
    throw Error(rc, msg);
}

struct Vec2 {
    float x = 0.0f, y = 0.0f;
};
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct IVec3 {
    int32_t x = 0, y = 0, z = 0;
};

// HitResult (VoxelMap.h:171-178): Distance <= 0 is a miss.
struct HitResult {
    double Distance = -1.0;
    Vec3 Normal;
    Vec2 UV;
    IVec3 VoxelPos;

    bool IsMiss() const { return Distance <= 0.0; }
};

// Owner of a vrt_world (move-only).
class World {
public:
    World(uint32_t log2_size_xz, uint32_t log2_size_y, uint32_t max_slots = 0, void* stream = nullptr) {
        vrt_world_desc d{log2_size_xz, log2_size_y, 0, max_slots};
        check(vrt_world_create(&d, stream, &w_));
    }
    explicit World(vrt_world* adopt) : w_(adopt) {}
    World(const World&) = delete;
    World& operator=(const World&) = delete;
    World(World&& o) noexcept : w_(std::exchange(o.w_, nullptr)) {}
    World& operator=(World&& o) noexcept {
        if (this != &o) {
            reset();
            w_ = std::exchange(o.w_, nullptr);
        }
        return *this;
    }
    ~World() { reset(); }

    vrt_world* handle() const { return w_; }

    // The sectors of one DirtyLocs batch, each sent whole: allocation mask and
    // popcount(mask) bricks of 512 bytes in increasing brick index.
    void Sync(uint32_t nsec, const uint32_t* sector_idx, const uint64_t* alloc_mask, const uint8_t* bricks,
              void* stream = nullptr) {
        check(vrt_world_sync(w_, nsec, sector_idx, alloc_mask, bricks, stream));
    }
    // The sectors of one DirtyLocs batch with their dirty masks: only the dirty bricks
    // (dirty_mask & alloc_mask) follow in `bricks`; the rest keep their device contents.
    void SyncBricks(uint32_t nsec, const uint32_t* sector_idx, const uint64_t* alloc_mask, const uint64_t* dirty_mask,
                    const uint8_t* bricks, void* stream = nullptr) {
        check(vrt_world_sync_bricks(w_, nsec, sector_idx, alloc_mask, dirty_mask, bricks, stream));
    }
    // Material::GetEncoded of the 256 palette entries (VoxelMap.h:27-41).
    void SetPalette(const uint64_t palette[256], void* stream = nullptr) {
        check(vrt_world_set_palette(w_, palette, stream));
    }
    vrt_world_info Info() const {
        vrt_world_info info{};
        check(vrt_world_get_info(w_, &info));
        return info;
    }
    // Plain per-ray cast of device ray buffers (RayCast, CpuRenderer.cpp:172-224).
    void Cast(const vrt_rays& rays, uint32_t max_iters, vrt_hits& hits, void* stream = nullptr) const {
        check(vrt_cast(w_, &rays, max_iters, &hits, stream));
    }

private:
    void reset() {
        if (w_) vrt_world_destroy(w_);
        w_ = nullptr;
    }
    vrt_world* w_ = nullptr;
};

// VoxelMap::RayCast (VoxelMap.cpp:140-170) on the device world: FP64, bias 1e-4,
// fixed 32 / 8 / 1 voxel levels, normal from tmin >= sideDist, synchronous.
class VoxelMapCaster {
public:
    explicit VoxelMapCaster(const World& world) : world_(world.handle()) {}
    explicit VoxelMapCaster(const vrt_world* world) : world_(world) {}

    template <class V3>
    HitResult RayCast(const V3& origin, const V3& dir, uint32_t maxIters = 1024) const {
        const double o[3] = {(double)origin.x, (double)origin.y, (double)origin.z};
        const double d[3] = {(double)dir.x, (double)dir.y, (double)dir.z};
        vrt_hit_f64 h{};
        check(vrt_cast_one_f64(world_, o, d, maxIters, &h));
        HitResult r;
        r.Distance = h.distance;
        r.Normal = {h.normal[0], h.normal[1], h.normal[2]};
        r.UV = {h.uv[0], h.uv[1]};
        r.VoxelPos = {h.voxel[0], h.voxel[1], h.voxel[2]};
        return r;
    }

private:
    const vrt_world* world_;
};

// GpuRenderer::RenderFrame's trace (VoxelRender.comp:28-94) into caller-owned
// device G-buffer images; settings as GpuRenderer::DrawSettings exposes them.
class HipRenderer {
public:
    explicit HipRenderer(const World& world) : world_(world.handle()) {}

    uint32_t NumLightBounces = 1;      // "Light Bounces"
    bool UseAnisotropicLods = false;   // "Anisotropic LODs": the ray-cell interaction table
    float Sky[3] = {0.6f, 0.75f, 1.0f};
    float NearZ = 0.01f, FarZ = 1000.0f;
    uint32_t FrameNo = 0;
    bool LastTileFirst = false;        // tile dispatch order (same images; see vrt_cast_order)

    void RenderFrame(const vrt_camera& cam, vrt_gbuffer& out, void* stream = nullptr) {
        vrt_render_opts o{};
        o.max_bounces = NumLightBounces;
        o.frame = FrameNo++;
        o.ray_cell_lut = UseAnisotropicLods ? 1u : 0u;
        for (int c = 0; c < 3; c++) o.sky[c] = Sky[c];
        o.near_z = NearZ;
        o.far_z = FarZ;
        o.order = LastTileFirst ? VRT_ORDER_REVERSE : VRT_ORDER_FORWARD;
        check(vrt_render_frame(world_, &cam, &o, &out, stream));
    }

private:
    const vrt_world* world_;
};

}  // namespace vrt
