// render.hip - vrt_render_frame: the reference GPU renderer's per-pixel shading
// (src/VoxelRT/Shaders/VoxelRender.comp:28-94) over the shader method, into
// G-buffer images (GBuffer.h:31-56 formats: RGBA8 albedo + packed normal,
// RGBA irradiance, R32F depth).
//
// One 64-lane group per 8x8 pixel tile, as GpuRenderer::RenderFrame dispatches
// (GpuRenderer.cpp:267-268).  Per lane: the primary rayCast, then the sun
// shadow ray and the bounce chain with rayCastCoarse (shader_trace.h).  The
// arithmetic follows the order stated for vrt_render_frame in vrt.h, so the
// images are reproducible bit for bit on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "shader_trace.h"
#include "vrt_cast_impl.h"

namespace vrt {

namespace {

struct RenderArgs {
    vrt_camera cam;
    float inv_w, inv_h;
    uint32_t max_bounces, frame;
    float sky[3];
    float sun[3];
    float pa, pb;
    const uint64_t* palette;
    uint32_t* albedo_normal;
    float* irradiance;
    float* depth;
};

__device__ __forceinline__ void normal_of(uint32_t code, float N[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int32_t c = (int32_t)((code >> (2 * a)) & 3u);
        N[a] = (float)((c ^ 2) - 2);
    }
}

__device__ __forceinline__ uint32_t unorm8(float x) { return (uint32_t)(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f + 0.5f); }

template <bool LUT>
__global__ __launch_bounds__(64) void render_kernel(ShaderView sv, RenderArgs a, uint32_t reverse) {
    // 8x8 tiles in row order, or last tile first (the image's cheap top rows then run in
    // the launch's tail; the pixels are the same)
    uint32_t tx = reverse ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
    uint32_t ty = reverse ? gridDim.y - 1u - blockIdx.y : blockIdx.y;
    uint32_t px = tx * 8 + (threadIdx.x & 7u), py = ty * 8 + (threadIdx.x >> 3);
    if (px >= a.cam.width || py >= a.cam.height) return;
    const uint32_t i = py * a.cam.width + px;
    const int32_t* W = a.cam.world_origin;
    const float sun_rgb[3] = {1.2f * 5.0f, 1.1f * 5.0f, 1.0f * 5.0f};
    uint32_t st = pcg_init(px, py, a.frame);
    float rd[3];
    camera_ray(a.cam.forward, a.cam.right, a.cam.up, a.inv_w, a.inv_h, px, py, rd);
    float albedo[3] = {1.0f, 1.0f, 1.0f}, irr[3], N[3] = {0.0f, 0.0f, 0.0f}, dep = -1.0f;
    ShaderHit hit;
    uint32_t iters;
    if (shader_trace<false, LUT>(sv, a.cam.eye, rd, W, 256u, hit)) {
        iters = hit.iters;
        uint64_t mat = a.palette[hit.mat];
        material_color(mat, albedo);
        normal_of(hit.code, N);
        float q = a.cam.forward[0] * (hit.cur[0] / 16.0f) + a.cam.forward[1] * (hit.cur[1] / 16.0f) +
                  a.cam.forward[2] * (hit.cur[2] / 16.0f);
        dep = (a.pa * -q + a.pb) / q;
        float em = material_emission(mat);
        float thr[3] = {1.0f, 1.0f, 1.0f};
#pragma unroll
        for (int c = 0; c < 3; c++) irr[c] = albedo[c] * em;
        if (a.max_bounces == 0) {
#pragma unroll
            for (int c = 0; c < 3; c++) irr[c] = albedo[c];
        } else {
            float so[3];
#pragma unroll
            for (int k = 0; k < 3; k++) so[k] = hit.cur[k] + N[k] * 0.01f;
            ShaderHit sh;
            if (!shader_trace<true, LUT>(sv, so, a.sun, W, 96u, sh)) {
#pragma unroll
                for (int c = 0; c < 3; c++) irr[c] = irr[c] + sun_rgb[c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) thr[c] = thr[c] * 0.5f;
            }
        }
        for (uint32_t b = 0; b < a.max_bounces; b++) {
            float bo[3], dd[3], bd[3];
#pragma unroll
            for (int k = 0; k < 3; k++) bo[k] = hit.cur[k] + N[k] * 0.01f;
            cosine_dir(N, st, dd);
#pragma unroll
            for (int k = 0; k < 3; k++) bd[k] = nudge_zero(dd[k]);
            bool bh = shader_trace<true, LUT>(sv, bo, bd, W, 96u, hit);
            iters = hit.iters;
            if (!bh) {
#pragma unroll
                for (int c = 0; c < 3; c++) irr[c] = irr[c] + thr[c] * a.sky[c];
                break;
            }
            normal_of(hit.code, N);
            uint64_t bm = a.palette[hit.mat];
            float col[3];
            material_color(bm, col);
#pragma unroll
            for (int c = 0; c < 3; c++) thr[c] = thr[c] * col[c];
            float bem = material_emission(bm);
            if (b < 2) {
                float so[3];
#pragma unroll
                for (int k = 0; k < 3; k++) so[k] = hit.cur[k] + N[k] * 0.01f;
                ShaderHit sh;
                if (!shader_trace<true, LUT>(sv, so, a.sun, W, 96u, sh)) {
#pragma unroll
                    for (int c = 0; c < 3; c++) thr[c] = thr[c] * sun_rgb[c];
                    bem = bem + 5.0f;
                }
            }
#pragma unroll
            for (int c = 0; c < 3; c++) irr[c] = irr[c] + thr[c] * bem;
        }
    } else {
        iters = hit.iters;
#pragma unroll
        for (int c = 0; c < 3; c++) irr[c] = a.sky[c];
    }
    uint32_t ncode = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) ncode |= (uint32_t)(int32_t)(N[k] + 1.0f) << (2 * k);
    a.albedo_normal[i] = unorm8(albedo[0]) | unorm8(albedo[1]) << 8 | unorm8(albedo[2]) << 16 | ncode << 24;
    *(float4*)(a.irradiance + (size_t)i * 4) = make_float4(irr[0], irr[1], irr[2], (float)iters);
    a.depth[i] = dep;
}

}  // namespace

// The renderer's fixed terms, computed on the host in FP32 (VoxelRender.comp:51,
// Camera.h:34): the unit sun direction and the depth projection pair.
void render_constants(float near_z, float far_z, float sun[3], float* pa, float* pb) {
    float x = 0.3f, y = 0.9f, z = -0.28f;
    float il = 1.0f / std::sqrt(x * x + y * y + z * z);
    sun[0] = x * il;
    sun[1] = y * il;
    sun[2] = z * il;
    *pa = -(far_z + near_z) / (far_z - near_z);
    *pb = -(2.0f * far_z * near_z) / (far_z - near_z);
}

}  // namespace vrt

using namespace vrt;

extern "C" int vrt_render_frame(const vrt_world* world, const vrt_camera* cam, const vrt_render_opts* opts,
                                vrt_gbuffer* out, void* stream) {
    return guarded([&]() -> int {
        if (!world || !cam || !opts || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: null argument");
        if (int rcw = check_world(world, "vrt_render_frame"); rcw != VRT_OK) return rcw;
        if (!cam->width || !cam->height) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: empty image");
        if ((uint64_t)cam->width * cam->height > 0x7FFFFFFFull)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: too many pixels");
        if (opts->max_bounces > 16) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: max_bounces must be 0..16");
        if (opts->ray_cell_lut > 1 || opts->order > VRT_ORDER_REVERSE)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: ray_cell_lut and order must be 0 or 1");
        float near_z = opts->near_z, far_z = opts->far_z;
        if (near_z == 0.0f && far_z == 0.0f) {
            near_z = 0.01f;
            far_z = 1000.0f;
        }
        if (!(near_z > 0.0f && far_z > near_z))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: need 0 < near_z < far_z");
        if (!out->albedo_normal || !out->irradiance || !out->depth)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: null image buffer");
        if (((uintptr_t)out->irradiance & 15) != 0)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_frame: irradiance must be 16-byte aligned");
        if (fault_injected("vrt_render_frame")) return fail(VRT_ERR_HIP, "vrt_render_frame: injected HIP fault");
        RenderArgs a;
        a.cam = *cam;
        a.inv_w = 1.0f / (float)cam->width;
        a.inv_h = 1.0f / (float)cam->height;
        a.max_bounces = opts->max_bounces;
        a.frame = opts->frame;
        for (int c = 0; c < 3; c++) a.sky[c] = opts->sky[c];
        render_constants(near_z, far_z, a.sun, &a.pa, &a.pb);
        a.palette = world->d_palette;
        a.albedo_normal = out->albedo_normal;
        a.irradiance = out->irradiance;
        a.depth = out->depth;
        ShaderView sv = world->shader_view();
        dim3 grid((cam->width + 7) / 8, (cam->height + 7) / 8);
        hipStream_t s = (hipStream_t)stream;
        uint32_t rev = opts->order == VRT_ORDER_REVERSE ? 1u : 0u;
        if (opts->ray_cell_lut)
            render_kernel<true><<<grid, 64, 0, s>>>(sv, a, rev);
        else
            render_kernel<false><<<grid, 64, 0, s>>>(sv, a, rev);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}
