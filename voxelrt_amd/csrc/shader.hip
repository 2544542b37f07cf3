// shader.hip - vrt_cast_shader: the GLSL traversal over a ray batch, one ray per lane.
#include <hip/hip_runtime.h>

#include <cstring>

#include "shader_trace.h"
#include "vrt_cast_impl.h"

namespace vrt {

namespace {

template <bool COARSE, bool LUT>
__global__ __launch_bounds__(256) void shader_cast_kernel(ShaderView sv, RaysIn r, HitsOut h, uint32_t cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.n) return;
    const float o[3] = {r.ox[i], r.oy[i], r.oz[i]};
    const float d[3] = {r.dx[i], r.dy[i], r.dz[i]};
    const int32_t W[3] = {r.W[0], r.W[1], r.W[2]};
    ShaderHit sh;
    // one trace per kernel: the tame loop pays for its second loop copy in both methods
    bool hit = shader_trace<COARSE, LUT, true>(sv, o, d, W, cap, sh);
    h.t[i] = sh.t;
    h.rec[i] = make_int4(sh.v[0], sh.v[1], sh.v[2], (int)(sh.mat | sh.code << 8 | (sh.iters & 0xFFFFu) << 16));
    if (h.pos_x) {
        h.pos_x[i] = hit ? sh.cur[0] : 0.0f;
        h.pos_y[i] = hit ? sh.cur[1] : 0.0f;
        h.pos_z[i] = hit ? sh.cur[2] : 0.0f;
    }
}

}  // namespace

}  // namespace vrt

using namespace vrt;

extern "C" int vrt_cast_shader(const vrt_world* world, const vrt_rays* rays, const vrt_shader_opts* opts,
                               vrt_hits* hits, void* stream) {
    return guarded([&]() -> int {
        if (!world || !opts) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_shader: null world or options");
        if (int rcw = check_world(world, "vrt_cast_shader"); rcw != VRT_OK) return rcw;
        if (opts->method > VRT_SHADER_COARSE) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_shader: unknown method");
        if (opts->ray_cell_lut > 1 || opts->reserved != 0)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_shader: ray_cell_lut must be 0 or 1, reserved 0");
        if (opts->max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_shader: max_iters must be 0..65535");
        if (fault_injected("vrt_cast_shader")) return fail(VRT_ERR_HIP, "vrt_cast_shader: injected HIP fault");
        RaysIn r;
        HitsOut h;
        int rc = check_rays_hits(rays, hits, r, h);
        if (rc != VRT_OK) return rc;
        if (r.n == 0) return VRT_OK;
        // 64-lane blocks: the grid's lanes (blocks * 64) must fit in 32 bits
        if (r.n > 0xFFFFFFC0u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_shader: too many rays");
        bool coarse = opts->method == VRT_SHADER_COARSE;
        uint32_t cap = opts->max_iters ? opts->max_iters : (coarse ? 96u : 256u);
        ShaderView sv = world->shader_view();
        hipStream_t s = (hipStream_t)stream;
        uint32_t blocks = (r.n + 63) / 64;
        if (coarse) {
            if (opts->ray_cell_lut)
                shader_cast_kernel<true, true><<<blocks, 64, 0, s>>>(sv, r, h, cap);
            else
                shader_cast_kernel<true, false><<<blocks, 64, 0, s>>>(sv, r, h, cap);
        } else {
            if (opts->ray_cell_lut)
                shader_cast_kernel<false, true><<<blocks, 64, 0, s>>>(sv, r, h, cap);
            else
                shader_cast_kernel<false, false><<<blocks, 64, 0, s>>>(sv, r, h, cap);
        }
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}
