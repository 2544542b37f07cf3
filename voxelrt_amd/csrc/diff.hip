// diff.hip - differentiable Lambertian render: forward (cast + shade, saving
// the hit voxel and shading weight of every pixel) and backward (scatter-add
// of dL/dpixel x weight into per-voxel RGB albedo gradients).
//
// No reference counterpart (the reference renders, it does not differentiate);
// the shading term is the reference's sun light: direction
// normalize(0.3, 0.9, -0.28), intensity 5 (Shaders/VoxelRender.comp:133-134)
// with a Lambert N.L factor.  The traversal stays discrete and exact (the same
// cast as vrt_cast), so radiance is linear in the albedo and
//   d radiance_pc / d albedo[v][c] = weight_p  if pixel p hit voxel v.
#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>

#include "traverse.h"
#include "vrt_cast_impl.h"

namespace vrt {

// Up to kFwdViews views per launch, their cameras passed in the kernel arguments: one
// launch over many small views fills the chip where one launch per view (a 512x512
// view is half a round of resident waves) leaves it draining after every view.
constexpr uint32_t kFwdViews = 16;
struct FwdViews {
    vrt_camera cam[kFwdViews];
    uint32_t block_begin[kFwdViews + 1];  // first block of 64 pixels of each view; [nv] = all blocks
    uint32_t pixel_base[kFwdViews];       // first output pixel of each view
    uint32_t nv;
};

__global__ __launch_bounds__(64) void render_fwd_kernel(WorldView w, FwdViews views, uint32_t max_iters,
                                                        float lx, float ly, float lz, float intensity,
                                                        const float* __restrict__ albedo, float* __restrict__ image,
                                                        uint32_t* __restrict__ saved_voxel,
                                                        float* __restrict__ saved_weight) {
    // the wave's view (blocks never straddle views)
    const uint32_t b = blockIdx.x;
    uint32_t k = 0;
    while (k + 1 < views.nv && b >= views.block_begin[k + 1]) k++;
    const vrt_camera& cam = views.cam[k];
    const uint32_t n = cam.width * cam.height;
    const uint32_t li = (b - views.block_begin[k]) * 64u + threadIdx.x;  // pixel within the view
    const bool valid = li < n;
    const uint32_t i = views.pixel_base[k] + li;                          // output pixel
    RayState st;
    int32_t v[3] = {0, 0, 0};
    bool tame = true;
    if (valid) {
        uint32_t px, py;
        tile_pixel(li, cam.width, cam.tile_w, cam.tile_h, px, py);
        float d[3];
        camera_ray(cam.forward, cam.right, cam.up, 1.0f / (float)cam.width, 1.0f / (float)cam.height, px, py, d);
        ray_setup(st, cam.eye[0], cam.eye[1], cam.eye[2], d[0], d[1], d[2], cam.world_origin[0], cam.world_origin[1],
                  cam.world_origin[2]);
        tame = ray_is_tame(st);
    }
    bool wave_tame = __all(tame);
    if (valid) {
        uint32_t slot = 0;
        bool hit = ray_trace_wave(w, st, max_iters, v, slot, wave_tame,
                                  origin_kind(cam.world_origin[0], cam.world_origin[1], cam.world_origin[2]));
        float wgt = 0.0f;
        uint32_t gv = 0xFFFFFFFFu;
        if (hit) {
            float t;
            int4 rec;
            ray_hit_record(w, st, v, slot, t, rec);
            float nx = normal_component((uint32_t)rec.w, 0), ny = normal_component((uint32_t)rec.w, 1),
                  nz = normal_component((uint32_t)rec.w, 2);
            float ndl = nx * lx + ny * ly + nz * lz;
            wgt = fmaxf(ndl, 0.0f) * intensity;
            uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
            gv = slot * 512u + ((ux & 7u) | (uz & 7u) << 3 | (uy & 7u) << 6);
        }
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (gv != 0xFFFFFFFFu) {
            r = albedo[(size_t)gv * 3 + 0] * wgt;
            g = albedo[(size_t)gv * 3 + 1] * wgt;
            b = albedo[(size_t)gv * 3 + 2] * wgt;
        }
        image[(size_t)i * 3 + 0] = r;
        image[(size_t)i * 3 + 1] = g;
        image[(size_t)i * 3 + 2] = b;
        saved_voxel[i] = gv;
        saved_weight[i] = wgt;
    }
}

__device__ __forceinline__ float wave_masked_sum(float x, bool take) {
    float s = take ? x : 0.0f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Scatter-add with wave aggregation: lanes that hit the same voxel are summed
// in registers (ballot on the key of the first pending lane, shuffle
// reduction) and one lane issues the atomics; after kRounds unique keys the
// remaining lanes add directly.
__global__ __launch_bounds__(256) void render_bwd_kernel(uint32_t n, const float* __restrict__ grad_image,
                                                         const uint32_t* __restrict__ saved_voxel,
                                                         const float* __restrict__ saved_weight,
                                                         float* __restrict__ grad_albedo) {
    const int kRounds = 4;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gv = 0xFFFFFFFFu;
    float gr = 0.0f, gg = 0.0f, gb = 0.0f;
    if (i < n) {
        gv = saved_voxel[i];
        float wgt = saved_weight[i];
        if (gv != 0xFFFFFFFFu && wgt != 0.0f) {
            gr = grad_image[(size_t)i * 3 + 0] * wgt;
            gg = grad_image[(size_t)i * 3 + 1] * wgt;
            gb = grad_image[(size_t)i * 3 + 2] * wgt;
        } else {
            gv = 0xFFFFFFFFu;
        }
    }
    bool pending = gv != 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    for (int round = 0; round < kRounds; round++) {
        uint64_t todo = __ballot(pending);
        if (todo == 0) break;
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
        uint32_t key = __shfl(gv, (int)leader, 64);
        bool mine = pending && gv == key;
        float sr = wave_masked_sum(gr, mine), sg = wave_masked_sum(gg, mine), sb = wave_masked_sum(gb, mine);
        if (lane == leader) {
            atomicAdd(&grad_albedo[(size_t)key * 3 + 0], sr);
            atomicAdd(&grad_albedo[(size_t)key * 3 + 1], sg);
            atomicAdd(&grad_albedo[(size_t)key * 3 + 2], sb);
        }
        if (mine) pending = false;
    }
    if (pending) {
        atomicAdd(&grad_albedo[(size_t)gv * 3 + 0], gr);
        atomicAdd(&grad_albedo[(size_t)gv * 3 + 1], gg);
        atomicAdd(&grad_albedo[(size_t)gv * 3 + 2], gb);
    }
}

}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_render_fwd(const vrt_world* world, const vrt_camera* cameras, uint32_t nviews, uint32_t max_iters,
                   const float light_dir[3], float intensity, const float* albedo, float* image,
                   uint32_t* saved_voxel, float* saved_weight, void* stream) {
    return guarded([&]() -> int {
        if (!world || (nviews && !cameras) || !light_dir)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: null argument");
        if (int rcw = check_world(world, "vrt_render_fwd"); rcw != VRT_OK) return rcw;
        if (max_iters == 0 || max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: max_iters must be 1..65535");
        if ((uint64_t)world->slot_cap * 512 >= 0xFFFFFFFFull)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: world too large for 32-bit voxel ids");
        hipStream_t s = (hipStream_t)stream;
        WorldView w = world->view();
        uint64_t total = 0;
        for (uint32_t k = 0; k < nviews; k++) {
            const vrt_camera& c = cameras[k];
            if (!c.width || !c.height || !c.tile_w || !c.tile_h || c.width % c.tile_w || c.height % c.tile_h)
                return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: camera size must be a multiple of its tile");
            total += (uint64_t)c.width * c.height;
        }
        // each launch's 64-lane blocks (at most one partial block per view) and every
        // output pixel index must fit in 32 bits
        if (total > 0xFFFFF000ull) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: more than 2^32 - 4096 pixels");
        if (total && (!albedo || !image || !saved_voxel || !saved_weight))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: null argument");
        uint32_t base = 0;
        for (uint32_t k0 = 0; k0 < nviews; k0 += kFwdViews) {
            FwdViews v{};
            v.nv = std::min(kFwdViews, nviews - k0);
            uint64_t blocks = 0;
            for (uint32_t k = 0; k < v.nv; k++) {
                const vrt_camera& c = cameras[k0 + k];
                v.cam[k] = c;
                v.block_begin[k] = (uint32_t)blocks;
                v.pixel_base[k] = base;
                uint32_t n = c.width * c.height;
                blocks += ((uint64_t)n + 63) / 64;
                base += n;
            }
            v.block_begin[v.nv] = (uint32_t)blocks;
            render_fwd_kernel<<<(uint32_t)blocks, 64, 0, s>>>(w, v, max_iters, light_dir[0], light_dir[1], light_dir[2],
                                                               intensity, albedo, image, saved_voxel, saved_weight);
            VRT_HIP(hipGetLastError());
        }
        return VRT_OK;
    });
}

int vrt_render_bwd(const vrt_world* world, uint32_t npixels, const float* grad_image, const uint32_t* saved_voxel,
                   const float* saved_weight, float* grad_albedo, void* stream) {
    return guarded([&]() -> int {
        if (!world || (npixels && (!grad_image || !saved_voxel || !saved_weight || !grad_albedo)))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd: null argument");
        if (int rcw = check_world(world, "vrt_render_bwd"); rcw != VRT_OK) return rcw;
        if (npixels == 0) return VRT_OK;
        if (npixels > 0xFFFFFF00u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd: more than 2^32 - 256 pixels");
        render_bwd_kernel<<<(npixels + 255) / 256, 256, 0, (hipStream_t)stream>>>(npixels, grad_image, saved_voxel,
                                                                                  saved_weight, grad_albedo);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}

}  // extern "C"
