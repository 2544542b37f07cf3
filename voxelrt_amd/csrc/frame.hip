// frame.hip - primary ray plus one diffuse bounce per pixel in one kernel (vrt_cast_frame).
//
// The reference GPU renderer traces a pixel's primary ray and then its bounces
// in one shader invocation (src/VoxelRT/Shaders/VoxelRender.comp:28-94,
// dispatched as 8x8 groups by GpuRenderer::RenderFrame, GpuRenderer.cpp:267-268).
// Here the per-ray arithmetic is the plain cast (traverse.h) and the bounce ray
// is the one vrt_bounce_generate builds (vrt_math.h bounce_ray), so the primary
// hits equal vrt_cast on the camera's rays and the bounce hits equal vrt_cast on
// the compacted bounce rays; only the scheduling differs:
//   GRID        one lane per pixel: primary, then (on a hit) the bounce ray;
//   PERSISTENT  waves keep 64 lanes busy: a lane whose ray finished parks; once
//               `refill` lanes are parked or idle, the wave writes the parked
//               results, turns parked primary hits into bounce rays and gives
//               idle lanes new pixels.  Primary and bounce rays are built in
//               registers (camera_ray, bounce_ray), so a refill waits on no load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "traverse.h"
#include "vrt_cast_impl.h"

namespace vrt {

__device__ __forceinline__ void store_hit(const HitsOut& h, uint32_t i, const float* cur, float t, int4 rec) {
    h.t[i] = t;
    h.rec[i] = rec;
    if (h.pos_x) {
        bool hit = (rec.w & 0xFF) != 0;
        h.pos_x[i] = hit ? cur[0] : 0.0f;
        h.pos_y[i] = hit ? cur[1] : 0.0f;
        h.pos_z[i] = hit ? cur[2] : 0.0f;
    }
}

__device__ __forceinline__ void finish_record(const WorldView& w, const RayState& st, int res, const int32_t v[3],
                                              uint32_t slot, float& t, int4& rec) {
    if (res == STEP_HIT)
        ray_hit_record(w, st, v, slot, t, rec);
    else
        ray_miss_record(st, t, rec);
}

struct FrameArgs {
    vrt_camera cam;
    float inv_w, inv_h;
    uint32_t n;
    uint32_t max_iters;
    uint32_t seed;
    HitsOut hp, hb;
};

__device__ __forceinline__ void setup_primary(const FrameArgs& a, uint32_t i, RayState& st) {
    uint32_t px, py;
    tile_pixel(i, a.cam.width, a.cam.tile_w, a.cam.tile_h, px, py);
    float d[3];
    camera_ray(a.cam.forward, a.cam.right, a.cam.up, a.inv_w, a.inv_h, px, py, d);
    ray_setup(st, a.cam.eye[0], a.cam.eye[1], a.cam.eye[2], d[0], d[1], d[2], a.cam.world_origin[0],
              a.cam.world_origin[1], a.cam.world_origin[2]);
}

__device__ __forceinline__ void setup_bounce(const FrameArgs& a, uint32_t i, uint32_t info, RayState& st) {
    float p[3] = {st.cur[0], st.cur[1], st.cur[2]};
    float o[3], d[3];
    bounce_ray(info, p, i, a.cam.width, a.seed, o, d);
    ray_setup(st, o[0], o[1], o[2], d[0], d[1], d[2], a.cam.world_origin[0], a.cam.world_origin[1],
              a.cam.world_origin[2]);
}

template <bool COUNT>
__device__ __forceinline__ void count_ray(unsigned long long* c, const RayState& st, int4 rec) {
    if constexpr (COUNT) {
        c[0]++;
        c[1] += st.iters;
        c[2] += st.descents;
        c[3] += (rec.w & 0xFF) != 0;
    }
}

__device__ __forceinline__ void flush4(vrt_cast_counters* ctr, unsigned long long* c) {
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) c[k] += __shfl_xor(c[k], off, 64);
    if (__lane_id() == 0) {
        atomicAdd(&ctr->rays, c[0]);
        atomicAdd(&ctr->iters, c[1]);
        atomicAdd(&ctr->descents, c[2]);
        atomicAdd(&ctr->hits, c[3]);
    }
}

// ---------------------------------------------------------------- GRID

// Traces the wave's active rays, with the fast step when all of them are tame;
// true on a hit, with v the hit voxel.
__device__ __forceinline__ bool trace_wave(const WorldView& w, RayState& st, bool active, uint32_t max_iters,
                                           int origin, int32_t v[3], uint32_t& slot) {
    bool tame = __all(!active || ray_is_tame(st));
    if (!active) return false;
    return ray_trace_wave(w, st, max_iters, v, slot, tame, origin);
}

__device__ __forceinline__ void finish_trace(const WorldView& w, const RayState& st, bool hit, const int32_t v[3],
                                             uint32_t slot, float& t, int4& rec) {
    if (hit)
        ray_hit_record(w, st, v, slot, t, rec);
    else
        ray_miss_record(st, t, rec);
}

template <bool COUNT>
__global__ __launch_bounds__(256) void frame_grid_kernel(WorldView w, FrameArgs a, vrt_cast_counters* ctr,
                                                         uint32_t reverse) {
    // blocks of pixels in order or last block first, as the plain cast's grid
    uint32_t blk = reverse ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
    uint32_t i = blk * blockDim.x + threadIdx.x;
    bool valid = i < a.n;
    unsigned long long c[4] = {0, 0, 0, 0};
    RayState st;
    int32_t v[3] = {0, 0, 0};
    uint32_t slot = 0;
    if (valid) setup_primary(a, i, st);
    const int origin = origin_kind(a.cam.world_origin[0], a.cam.world_origin[1], a.cam.world_origin[2]);
    bool hit = trace_wave(w, st, valid, a.max_iters, origin, v, slot);
    bool bounce = false;
    uint32_t info = 0;
    if (valid) {
        float t;
        int4 rec;
        finish_trace(w, st, hit, v, slot, t, rec);
        store_hit(a.hp, i, st.cur, t, rec);
        count_ray<COUNT>(c, st, rec);
        info = (uint32_t)rec.w;
        bounce = (info & 0xFF) != 0;
        if (bounce) setup_bounce(a, i, info, st);
    }
    hit = trace_wave(w, st, bounce, a.max_iters, origin, v, slot);
    if (valid) {
        if (bounce) {
            float t;
            int4 rec;
            finish_trace(w, st, hit, v, slot, t, rec);
            store_hit(a.hb, i, st.cur, t, rec);
            count_ray<COUNT>(c, st, rec);
        } else {
            float z[3] = {0.0f, 0.0f, 0.0f};
            store_hit(a.hb, i, z, -1.0f, make_int4(0, 0, 0, (int)VRT_HIT_NO_RAY));
        }
    }
    if constexpr (COUNT) flush4(ctr, c);
}

// ---------------------------------------------------------------- PERSISTENT

enum LanePhase : int { LANE_IDLE = 0, LANE_PRIMARY = 1, LANE_BOUNCE = 2, LANE_PARKED_PRIMARY = 3, LANE_PARKED_BOUNCE = 4 };

template <bool COUNT>
__global__ __launch_bounds__(256) void frame_persistent_kernel(WorldView w, FrameArgs a, uint32_t chunk,
                                                               uint32_t refill_min, ChunkQueue cq,
                                                               vrt_cast_counters* ctr) {
    const uint32_t n = a.n;
    const uint32_t home = blockIdx.x % kQueues;
    const uint32_t wave_in_part = (blockIdx.x / kQueues) * (blockDim.x / 64) + threadIdx.x / 64;
    uint32_t probe = 0;
    uint32_t first = part_begin(home, cq.nchunks) + wave_in_part;
    uint32_t pool_next = 0, pool_end = 0;  // wave-uniform
    if (wave_in_part < cq.waves_per_q && first < part_begin(home + 1, cq.nchunks)) {
        pool_next = min(first * chunk, n);
        pool_end = min(pool_next + chunk, n);
    }
    bool exhausted = false;
    bool tame = true;
    int phase = LANE_IDLE;
    int res = STEP_CONTINUE;
    uint32_t idx = 0, slot = 0;
    int32_t v[3] = {0, 0, 0};
    RayState st;
    unsigned long long c[4] = {0, 0, 0, 0};
    for (;;) {
        uint64_t freem = __ballot(phase == LANE_IDLE || phase >= LANE_PARKED_PRIMARY);
        if ((uint32_t)__popcll(freem) >= refill_min) {
            if (phase == LANE_PARKED_PRIMARY) {
                float t;
                int4 rec;
                finish_record(w, st, res, v, slot, t, rec);
                store_hit(a.hp, idx, st.cur, t, rec);
                count_ray<COUNT>(c, st, rec);
                if ((rec.w & 0xFF) != 0) {
                    setup_bounce(a, idx, (uint32_t)rec.w, st);
                    phase = LANE_BOUNCE;
                } else {
                    float z[3] = {0.0f, 0.0f, 0.0f};
                    store_hit(a.hb, idx, z, -1.0f, make_int4(0, 0, 0, (int)VRT_HIT_NO_RAY));
                    phase = LANE_IDLE;
                }
            } else if (phase == LANE_PARKED_BOUNCE) {
                float t;
                int4 rec;
                finish_record(w, st, res, v, slot, t, rec);
                store_hit(a.hb, idx, st.cur, t, rec);
                count_ray<COUNT>(c, st, rec);
                phase = LANE_IDLE;
            }
            uint64_t idle = __ballot(phase == LANE_IDLE);
            uint32_t nidle = (uint32_t)__popcll(idle);
            if (pool_next == pool_end && !exhausted && nidle) {
                uint32_t ch = claim_chunk(cq, home, probe);
                if (ch == ~0u) {
                    exhausted = true;
                } else {
                    pool_next = min(ch * chunk, n);
                    pool_end = min(pool_next + chunk, n);
                }
            }
            uint32_t avail = pool_end - pool_next;
            if (phase == LANE_IDLE) {
                uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                if (rank < avail) {
                    idx = pool_next + rank;
                    setup_primary(a, idx, st);
                    phase = LANE_PRIMARY;
                }
            }
            pool_next += min(avail, nidle);
            bool running = phase == LANE_PRIMARY || phase == LANE_BOUNCE;
            tame = __ballot(running && !ray_is_tame(st)) == 0;
            if (!__any(running)) {
                if (exhausted) break;
                continue;
            }
        }
        if (phase == LANE_PRIMARY || phase == LANE_BOUNCE) {
#pragma unroll
            for (int k = 0; k < 2; k++) {
                if (phase == LANE_PRIMARY || phase == LANE_BOUNCE) {
                    if (st.iters >= a.max_iters)
                        res = STEP_MISS;
                    else if (tame)
                        res = ray_step<false>(w, st, v, &slot);
                    else
                        res = ray_step<true>(w, st, v, &slot);
                    if (res != STEP_CONTINUE) phase += 2;  // primary -> parked primary, bounce -> parked bounce
                }
            }
        }
    }
    if constexpr (COUNT) flush4(ctr, c);
}

__global__ void frame_fill_kernel(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace vrt

using namespace vrt;

extern "C" int vrt_cast_frame(const vrt_world* world, const vrt_camera* cam, uint32_t max_iters, uint32_t seed,
                              vrt_hits* primary, vrt_hits* bounce, const vrt_frame_opts* opts, void* stream) {
    return guarded([&]() -> int {
        if (!world || !cam || !primary || !bounce) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: null argument");
        if (int rcw = check_world(world, "vrt_cast_frame"); rcw != VRT_OK) return rcw;
        if (max_iters == 0 || max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: max_iters must be 1..65535");
        if (!cam->width || !cam->height || !cam->tile_w || !cam->tile_h || cam->width % cam->tile_w ||
            cam->height % cam->tile_h)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: camera size must be a multiple of its tile");
        if (!primary->t || !primary->rec || !bounce->t || !bounce->rec)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: null hit buffer");
        if (((uintptr_t)primary->rec & 15) || ((uintptr_t)bounce->rec & 15))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: hit records must be 16-byte aligned");
        uint64_t n64 = (uint64_t)cam->width * cam->height;
        if (n64 > 0x7FFFFFFFull) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: too many pixels");
        vrt_frame_opts o;
        std::memset(&o, 0, sizeof(o));
        if (opts) o = *opts;
        hipStream_t s = (hipStream_t)stream;
        FrameArgs a;
        a.cam = *cam;
        a.inv_w = 1.0f / (float)cam->width;
        a.inv_h = 1.0f / (float)cam->height;
        a.n = (uint32_t)n64;
        a.max_iters = max_iters;
        a.seed = seed;
        if (primary->uv || primary->material_data || bounce->uv || bounce->material_data)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: uv and material_data are filled by vrt_cast only");
        auto hits_view = [&](const vrt_hits* hh, HitsOut& out) {
            out = HitsOut{};
            out.t = hh->t;
            out.rec = (int4*)hh->rec;
            out.pos_x = hh->pos;
            out.pos_y = hh->pos ? hh->pos + a.n : nullptr;
            out.pos_z = hh->pos ? hh->pos + 2 * (size_t)a.n : nullptr;
        };
        hits_view(primary, a.hp);
        hits_view(bounce, a.hb);
        WorldView w = world->view();
        uint32_t sched = o.schedule == VRT_SCHED_AUTO ? VRT_SCHED_GRID : o.schedule;
        if (sched == VRT_SCHED_GRID) {
            // one wave per workgroup, as the plain cast's grid
            if (o.order > VRT_ORDER_REVERSE) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: unknown order");
            uint32_t blocks = (a.n + 63) / 64;
            uint32_t rev = o.order == VRT_ORDER_REVERSE ? 1u : 0u;
            if (o.counters)
                frame_grid_kernel<true><<<blocks, 64, 0, s>>>(w, a, o.counters, rev);
            else
                frame_grid_kernel<false><<<blocks, 64, 0, s>>>(w, a, nullptr, rev);
            VRT_HIP(hipGetLastError());
            return VRT_OK;
        }
        if (sched != VRT_SCHED_PERSISTENT) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: schedule must be grid or persistent");
        uint32_t chunk = o.chunk ? o.chunk : 256;
        if (chunk % 64 != 0 || chunk > (1u << 20)) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_frame: chunk must be a multiple of 64");
        int cus = 0;
        uint32_t* queue = nullptr;
        int rc = device_info(world->device, &cus, &queue, s);
        if (rc != VRT_OK) return rc;
        uint32_t blocks = std::max<uint32_t>(kQueues, ((uint32_t)cus * 32 / 4) / kQueues * kQueues);
        ChunkQueue cq;
        cq.ctr = queue;
        cq.nchunks = (a.n + chunk - 1) / chunk;
        cq.waves_per_q = blocks / kQueues * 4;
        uint32_t refill_min = o.refill ? std::min<uint32_t>(std::max<uint32_t>(o.refill, 1), 64) : 16;
        frame_fill_kernel<<<(kQueues * kQueueStride + 255) / 256, 256, 0, s>>>(queue, kQueues * kQueueStride, 0u);
        if (o.counters)
            frame_persistent_kernel<true><<<blocks, 256, 0, s>>>(w, a, chunk, refill_min, cq, o.counters);
        else
            frame_persistent_kernel<false><<<blocks, 256, 0, s>>>(w, a, chunk, refill_min, cq, nullptr);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}
