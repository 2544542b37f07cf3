// batch.hip - many cameras and ray sets of uneven sizes in one launch, its
// hipGraph capture and the autotuner (vrt_batch_* in vrt.h).
//
// No reference counterpart: the reference traces one camera per frame
// (GpuRenderer::RenderFrame, src/VoxelRT/GpuRenderer.cpp:239-273, one
// dispatch of 8x8 groups; CpuRenderer::RenderFrame, CpuRenderer.cpp:415-474).
// Per ray the arithmetic is the same cast (traverse.h), so every segment's
// hits equal a plain vrt_cast of the same rays.
//
// Layout of the launch: the batch is a virtual index space of wave blocks of
// 64 * rays_per_lane rays; every segment is padded to whole blocks so a wave
// never straddles two segments and its segment lookup is wave-uniform (a
// binary search over the block prefix, scalar loads).  Camera segments build
// their rays in the kernel (vrt_math.h camera_ray): lane -> pixel follows the
// tuned tile (8x8 or 16x4), and the hit goes to the pixel's index in the
// camera's own ray order, so the output does not depend on the tuning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "traverse.h"
#include "vrt_cast_impl.h"

namespace vrt {

struct BatchSegDev {
    uint32_t kind, cam, ray_offset, hit_offset, count, block_begin;
    int32_t W[3];
    uint32_t pad;
    const uint32_t* count_ptr;
};

struct BatchArgs {
    const BatchSegDev* segs;
    uint32_t nseg;
    uint32_t nblocks;
    const vrt_camera* cams;
    const float* ox; const float* oy; const float* oz;
    const float* dx; const float* dy; const float* dz;
    HitsOut h;
    uint32_t max_iters;
    uint32_t tile_w, tile_h;
};

__device__ __forceinline__ uint32_t find_segment(const BatchArgs& a, uint32_t block) {
    uint32_t lo = 0, hi = a.nseg;  // largest s with segs[s].block_begin <= block
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (a.segs[mid].block_begin <= block) lo = mid; else hi = mid;
    }
    return lo;
}

// Index of pixel (px, py) in the camera's own ray order (inverse of tile_pixel).
__device__ __forceinline__ uint32_t camera_order_index(const vrt_camera& c, uint32_t px, uint32_t py) {
    uint32_t tx = px / c.tile_w, ty = py / c.tile_h;
    return (ty * (c.width / c.tile_w) + tx) * (c.tile_w * c.tile_h) + (py - ty * c.tile_h) * c.tile_w +
           (px - tx * c.tile_w);
}

struct LaneRay {
    bool valid;
    uint32_t out;  // hit index
};

// Loads (or builds) ray k of the lane in block `local_block` of segment `sg`.
__device__ __forceinline__ LaneRay lane_ray(const BatchArgs& a, const BatchSegDev& sg, uint32_t count, uint32_t local,
                                            RayState& st) {
    LaneRay lr;
    lr.valid = local < count;
    lr.out = 0;
    if (!lr.valid) return lr;
    if (sg.kind == VRT_SEG_CAMERA) {
        const vrt_camera& c = a.cams[sg.cam];
        uint32_t px, py;
        tile_pixel(local, c.width, a.tile_w, a.tile_h, px, py);
        float d[3];
        camera_ray(c.forward, c.right, c.up, 1.0f / (float)c.width, 1.0f / (float)c.height, px, py, d);
        ray_setup(st, c.eye[0], c.eye[1], c.eye[2], d[0], d[1], d[2], c.world_origin[0], c.world_origin[1],
                  c.world_origin[2]);
        lr.out = sg.hit_offset + camera_order_index(c, px, py);
    } else {
        uint32_t i = sg.ray_offset + local;
        ray_setup(st, a.ox[i], a.oy[i], a.oz[i], a.dx[i], a.dy[i], a.dz[i], sg.W[0], sg.W[1], sg.W[2]);
        lr.out = sg.hit_offset + local;
    }
    return lr;
}

// The origin kind of the segment's rays (the camera's world origin for camera segments).
__device__ __forceinline__ int segment_origin(const BatchArgs& a, const BatchSegDev& sg) {
    if (sg.kind == VRT_SEG_CAMERA) {
        const vrt_camera& c = a.cams[sg.cam];
        return origin_kind(c.world_origin[0], c.world_origin[1], c.world_origin[2]);
    }
    return origin_kind(sg.W[0], sg.W[1], sg.W[2]);
}

template <int RPL, bool COUNT>
__device__ __forceinline__ void trace_block(const WorldView& w, const BatchArgs& a, uint32_t block,
                                            vrt_cast_counters* ctr) {
    uint32_t s = find_segment(a, block);
    BatchSegDev sg = a.segs[s];
    uint32_t count = sg.count_ptr ? min(*sg.count_ptr, sg.count) : sg.count;
    uint32_t lane = __lane_id();
    RayState st[RPL];
    LaneRay lr[RPL];
    bool done[RPL];
    bool tame = true;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        uint32_t local = (block - sg.block_begin) * (64 * RPL) + (uint32_t)k * 64 + lane;
        lr[k] = lane_ray(a, sg, count, local, st[k]);
        done[k] = !lr[k].valid;
        tame = tame && (done[k] || ray_is_tame(st[k]));
    }
    bool wave_tame = __all(tame);
    bool hit[RPL];
    int32_t v[RPL][3];
    uint32_t slot[RPL];
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        hit[k] = false;
        slot[k] = 0;
    }
    if constexpr (RPL == 1) {
        if (!done[0])
            hit[0] = ray_trace_wave(w, st[0], a.max_iters, v[0], slot[0], wave_tame, segment_origin(a, sg));
    } else {
        // the lane's rays step in lockstep, so one counter caps them all
        for (uint32_t it = 0; it < a.max_iters; it++) {
            bool any_running = false;
#pragma unroll
            for (int k = 0; k < RPL; k++) {
                if (!done[k]) {
                    bool cont = wave_tame ? ray_advance<false>(w, st[k], v[k], hit[k], &slot[k])
                                          : ray_advance<true>(w, st[k], v[k], hit[k], &slot[k]);
                    done[k] = !cont;
                }
                any_running = any_running || !done[k];
            }
            if (!any_running) break;
        }
    }
    unsigned long long c_rays = 0, c_iters = 0, c_desc = 0, c_hits = 0;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        if (lr[k].valid) {
            float t;
            int4 rec;
            if (hit[k])
                ray_hit_record(w, st[k], v[k], slot[k], t, rec);
            else
                ray_miss_record(st[k], t, rec);
            a.h.t[lr[k].out] = t;
            a.h.rec[lr[k].out] = rec;
            if (a.h.pos_x) {
                bool hit = (rec.w & 0xFF) != 0;
                a.h.pos_x[lr[k].out] = hit ? st[k].cur[0] : 0.0f;
                a.h.pos_y[lr[k].out] = hit ? st[k].cur[1] : 0.0f;
                a.h.pos_z[lr[k].out] = hit ? st[k].cur[2] : 0.0f;
            }
            if constexpr (COUNT) {
                c_rays++;
                c_iters += st[k].iters;
                c_desc += st[k].descents;
                c_hits += (rec.w & 0xFF) != 0;
            }
        }
    }
    if constexpr (COUNT) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            c_rays += __shfl_xor(c_rays, off, 64);
            c_iters += __shfl_xor(c_iters, off, 64);
            c_desc += __shfl_xor(c_desc, off, 64);
            c_hits += __shfl_xor(c_hits, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&ctr->rays, c_rays);
            atomicAdd(&ctr->iters, c_iters);
            atomicAdd(&ctr->descents, c_desc);
            atomicAdd(&ctr->hits, c_hits);
        }
    }
}

// One wave per block of the virtual index space; the hardware dispatcher balances waves.
template <int RPL, bool COUNT>
__global__ __launch_bounds__(256) void batch_grid_kernel(WorldView w, BatchArgs a, vrt_cast_counters* ctr) {
    uint32_t block = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (block >= a.nblocks) return;  // wave-uniform
    trace_block<RPL, COUNT>(w, a, block, ctr);
}

// Persistent: a device-sized grid whose waves claim blocks from a counter.
template <int RPL, bool COUNT>
__global__ __launch_bounds__(256) void batch_persistent_kernel(WorldView w, BatchArgs a, uint32_t* queue,
                                                               uint32_t claim, vrt_cast_counters* ctr) {
    for (;;) {
        uint32_t b = 0;
        if (__lane_id() == 0) b = atomicAdd(queue, claim);
        b = __shfl(b, 0, 64);
        if (b >= a.nblocks) break;
        uint32_t e = min(b + claim, a.nblocks);
        for (uint32_t k = b; k < e; k++) trace_block<RPL, COUNT>(w, a, k, ctr);
    }
}

__global__ void batch_fill_kernel(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

}  // namespace vrt

using namespace vrt;

struct vrt_batch_plan {
    const vrt_world* world = nullptr;
    int device = 0;
    vrt_batch_tuning tuning{};
    uint32_t schedule = VRT_SCHED_GRID;
    BatchArgs args{};
    BatchSegDev* d_segs = nullptr;
    vrt_camera* d_cams = nullptr;
    uint32_t* d_queue = nullptr;
    uint64_t total_rays = 0;
    int cus = 256;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_generation = 0;  // the world's generation when the graph was captured
};

static void destroy_plan(vrt_batch_plan* p) {
    if (!p) return;
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    (void)hipFree(p->d_segs);
    (void)hipFree(p->d_cams);
    (void)hipFree(p->d_queue);
    delete p;
}

static int validate_tuning(const vrt_batch_tuning& t) {
    if (!((t.tile_w == 8 && t.tile_h == 8) || (t.tile_w == 16 && t.tile_h == 4) || (t.tile_w == 4 && t.tile_h == 16) ||
          (t.tile_w == 64 && t.tile_h == 1)))
        return fail(VRT_ERR_INVALID_ARGUMENT, "batch tuning: tile must be 8x8, 16x4, 4x16 or 64x1");
    if (t.rays_per_lane != 1 && t.rays_per_lane != 2)
        return fail(VRT_ERR_INVALID_ARGUMENT, "batch tuning: rays_per_lane must be 1 or 2");
    if (t.chunk != 0 && (t.chunk % 64 != 0 || t.chunk > 65536))
        return fail(VRT_ERR_INVALID_ARGUMENT, "batch tuning: chunk must be 0 (grid) or a multiple of 64");
    return VRT_OK;
}

static int launch_plan(vrt_batch_plan* p, vrt_cast_counters* ctr, hipStream_t s) {
    if (p->args.nblocks == 0) return VRT_OK;
    WorldView w = p->world->view();
    const uint32_t rpl = p->tuning.rays_per_lane;
    if (p->tuning.chunk == 0) {
        // one wave (one ray block) per workgroup, as the plain cast's grid
        uint32_t grid = p->args.nblocks;
        if (rpl == 1) {
            if (ctr) batch_grid_kernel<1, true><<<grid, 64, 0, s>>>(w, p->args, ctr);
            else batch_grid_kernel<1, false><<<grid, 64, 0, s>>>(w, p->args, nullptr);
        } else {
            if (ctr) batch_grid_kernel<2, true><<<grid, 64, 0, s>>>(w, p->args, ctr);
            else batch_grid_kernel<2, false><<<grid, 64, 0, s>>>(w, p->args, nullptr);
        }
    } else {
        uint32_t claim = std::max<uint32_t>(1, p->tuning.chunk / (64 * rpl));
        uint32_t grid = std::min<uint32_t>((uint32_t)p->cus * 8, (p->args.nblocks + 3) / 4);
        batch_fill_kernel<<<1, 64, 0, s>>>(p->d_queue, 0u);
        if (rpl == 1) {
            if (ctr) batch_persistent_kernel<1, true><<<grid, 256, 0, s>>>(w, p->args, p->d_queue, claim, ctr);
            else batch_persistent_kernel<1, false><<<grid, 256, 0, s>>>(w, p->args, p->d_queue, claim, nullptr);
        } else {
            if (ctr) batch_persistent_kernel<2, true><<<grid, 256, 0, s>>>(w, p->args, p->d_queue, claim, ctr);
            else batch_persistent_kernel<2, false><<<grid, 256, 0, s>>>(w, p->args, p->d_queue, claim, nullptr);
        }
    }
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

static int create_plan(const vrt_world* world, const vrt_batch* b, const vrt_batch_tuning* tun, hipStream_t s,
                       vrt_batch_plan** out) {
    if (!world || !b || !tun || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_batch_plan_create: null argument");
    if (int rcw = check_world(world, "vrt_batch_plan_create"); rcw != VRT_OK) return rcw;
    *out = nullptr;
    int rc = validate_tuning(*tun);
    if (rc != VRT_OK) return rc;
    if (b->max_iters == 0 || b->max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: max_iters must be 1..65535");
    if (b->nsegments && !b->segments) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: null segment table");
    if (b->ncameras && !b->cameras) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: null camera table");
    if (((uintptr_t)b->hits.rec & 15) != 0) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: hit records must be 16-byte aligned");
    if (b->hits.uv || b->hits.material_data)
        return fail(VRT_ERR_INVALID_ARGUMENT, "batch: uv and material_data are filled by vrt_cast only");
    const uint32_t per_block = 64 * tun->rays_per_lane;
    std::vector<BatchSegDev> segs(b->nsegments);
    uint64_t blocks = 0, rays = 0;
    for (uint32_t i = 0; i < b->nsegments; i++) {
        const vrt_segment& g = b->segments[i];
        BatchSegDev& d = segs[i];
        std::memset(&d, 0, sizeof(d));
        d.kind = g.kind;
        d.cam = g.camera;
        d.ray_offset = g.ray_offset;
        d.hit_offset = g.hit_offset;
        d.count = g.count;
        d.W[0] = g.world_origin[0];
        d.W[1] = g.world_origin[1];
        d.W[2] = g.world_origin[2];
        d.count_ptr = g.count_ptr;
        if (g.kind == VRT_SEG_CAMERA) {
            if (g.camera >= b->ncameras) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: segment camera out of range");
            const vrt_camera& c = b->cameras[g.camera];
            if (!c.width || !c.height || !c.tile_w || !c.tile_h || c.width % c.tile_w || c.height % c.tile_h)
                return fail(VRT_ERR_INVALID_ARGUMENT, "batch: camera size must be a multiple of its tile");
            if (c.width % tun->tile_w || c.height % tun->tile_h)
                return fail(VRT_ERR_INVALID_ARGUMENT, "batch: camera size must be a multiple of the tuned tile");
            // pixel indices are 32-bit in the kernel
            if ((uint64_t)c.width * c.height > 0xFFFFFF00ull)
                return fail(VRT_ERR_INVALID_ARGUMENT, "batch: camera with more than 2^32 - 256 pixels");
            d.count = c.width * c.height;
            if (g.count_ptr) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: camera segments have a fixed count");
        } else if (g.kind == VRT_SEG_RAYS) {
            if (d.count && (!b->rays.ox || !b->rays.oy || !b->rays.oz || !b->rays.dx || !b->rays.dy || !b->rays.dz))
                return fail(VRT_ERR_INVALID_ARGUMENT, "batch: ray segments need ray buffers");
        } else {
            return fail(VRT_ERR_INVALID_ARGUMENT, "batch: unknown segment kind");
        }
        // hit_offset + index and ray_offset + index are 32-bit in the kernel
        if ((uint64_t)d.hit_offset + d.count > 0x100000000ull)
            return fail(VRT_ERR_INVALID_ARGUMENT, "batch: segment " + std::to_string(i) +
                                                      " reaches past hit index 2^32 - 1");
        if (g.kind == VRT_SEG_RAYS && (uint64_t)d.ray_offset + d.count > 0x100000000ull)
            return fail(VRT_ERR_INVALID_ARGUMENT, "batch: segment " + std::to_string(i) +
                                                      " reaches past ray index 2^32 - 1");
        if (d.count > 0xFFFFFF00u)
            return fail(VRT_ERR_INVALID_ARGUMENT, "batch: segment " + std::to_string(i) + " has more than 2^32 - 256 rays");
        d.block_begin = (uint32_t)blocks;
        blocks += ((uint64_t)d.count + per_block - 1) / per_block;
        rays += d.count;
        if (blocks > 0x7FFFFFFFull) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: too many rays");
    }
    if (rays && (!b->hits.t || !b->hits.rec)) return fail(VRT_ERR_INVALID_ARGUMENT, "batch: null hit buffers");
    auto* p = new vrt_batch_plan();
    p->world = world;
    p->device = world->device;
    p->tuning = *tun;
    p->total_rays = rays;
    int cus = 256;
    rc = device_info(world->device, &cus, nullptr, s);
    if (rc != VRT_OK) {
        destroy_plan(p);
        return rc;
    }
    p->cus = cus;
    auto bad = [&](hipError_t e, const char* what) {
        int r = hip_fail(e, what);
        destroy_plan(p);
        return r;
    };
    hipError_t e;
    if (!segs.empty()) {
        if ((e = hipMalloc(&p->d_segs, segs.size() * sizeof(BatchSegDev))) != hipSuccess) return bad(e, "hipMalloc(segments)");
        if ((e = hipMemcpyAsync(p->d_segs, segs.data(), segs.size() * sizeof(BatchSegDev), hipMemcpyHostToDevice, s)) != hipSuccess)
            return bad(e, "hipMemcpyAsync(segments)");
    }
    if (b->ncameras) {
        if ((e = hipMalloc(&p->d_cams, b->ncameras * sizeof(vrt_camera))) != hipSuccess) return bad(e, "hipMalloc(cameras)");
        if ((e = hipMemcpyAsync(p->d_cams, b->cameras, b->ncameras * sizeof(vrt_camera), hipMemcpyHostToDevice, s)) != hipSuccess)
            return bad(e, "hipMemcpyAsync(cameras)");
    }
    if ((e = hipMalloc(&p->d_queue, 256)) != hipSuccess) return bad(e, "hipMalloc(queue)");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return bad(e, "hipStreamSynchronize");
    BatchArgs& a = p->args;
    a.segs = p->d_segs;
    a.nseg = b->nsegments;
    a.nblocks = (uint32_t)blocks;
    a.cams = p->d_cams;
    a.ox = b->rays.ox; a.oy = b->rays.oy; a.oz = b->rays.oz;
    a.dx = b->rays.dx; a.dy = b->rays.dy; a.dz = b->rays.dz;
    a.h.t = b->hits.t;
    a.h.rec = (int4*)b->hits.rec;
    a.h.pos_x = b->hits.pos;
    a.h.pos_y = a.h.pos_z = nullptr;
    if (b->hits.pos) {
        // positions are [3][total] SoA over the batch's hit index space
        uint64_t hit_space = 0;
        for (uint32_t i = 0; i < b->nsegments; i++)
            hit_space = std::max<uint64_t>(hit_space, (uint64_t)segs[i].hit_offset + segs[i].count);
        a.h.pos_y = b->hits.pos + hit_space;
        a.h.pos_z = b->hits.pos + 2 * hit_space;
    }
    a.max_iters = b->max_iters;
    a.tile_w = tun->tile_w;
    a.tile_h = tun->tile_h;
    *out = p;
    return VRT_OK;
}

// ---------------------------------------------------------------- autotune cache

static std::mutex g_tune_mu;
static std::map<std::string, std::pair<vrt_batch_tuning, float>> g_tune_cache;

static std::string batch_signature(const vrt_world* w, const vrt_batch* b) {
    std::ostringstream os;
    os << "w" << w->log2_xz << "x" << w->log2_y << "b" << w->num_bricks << "i" << b->max_iters << "n" << b->nsegments;
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the segment shapes
    auto mix = [&](uint64_t x) {
        for (int i = 0; i < 8; i++) {
            h ^= (x >> (8 * i)) & 0xFF;
            h *= 1099511628211ull;
        }
    };
    for (uint32_t i = 0; i < b->nsegments; i++) {
        const vrt_segment& g = b->segments[i];
        mix(g.kind);
        if (g.kind == VRT_SEG_CAMERA && g.camera < b->ncameras) {
            mix(b->cameras[g.camera].width);
            mix(b->cameras[g.camera].height);
        } else {
            mix(g.count);
        }
    }
    os << "h" << std::hex << h;
    return os.str();
}

static std::string device_name(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string n = prop.gcnArchName;
    for (char& c : n)
        if (c == ' ' || c == ':') c = '_';
    return n + "_cu" + std::to_string(prop.multiProcessorCount);
}

extern "C" {

int vrt_batch_plan_create(const vrt_world* world, const vrt_batch* batch, const vrt_batch_tuning* tuning,
                          void* stream, vrt_batch_plan** out) {
    return guarded([&]() -> int { return create_plan(world, batch, tuning, (hipStream_t)stream, out); });
}

int vrt_batch_launch(vrt_batch_plan* plan, vrt_cast_counters* counters, void* stream) {
    return guarded([&]() -> int {
        if (!plan) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_batch_launch: null plan");
        if (int rcw = check_world(plan->world, "vrt_batch_launch"); rcw != VRT_OK) return rcw;
        if (fault_injected("vrt_batch_launch")) return fail(VRT_ERR_HIP, "vrt_batch_launch: injected HIP fault");
        return launch_plan(plan, counters, (hipStream_t)stream);
    });
}

int vrt_batch_graph_launch(vrt_batch_plan* plan, void* stream) {
    return guarded([&]() -> int {
        if (!plan) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_batch_graph_launch: null plan");
        if (int rcw = check_world(plan->world, "vrt_batch_graph_launch"); rcw != VRT_OK) return rcw;
        if (plan->args.nblocks == 0) return VRT_OK;  // nothing to cast, nothing to capture
        if (plan->graph_exec && plan->graph_generation != plan->world->generation) {
            // the world's occupancy and voxel buffers moved (slot growth) since the capture: the
            // graph's kernel arguments point at freed memory, so capture again
            VRT_HIP(hipDeviceSynchronize());  // earlier replays may be in flight on any stream
            (void)hipGraphExecDestroy(plan->graph_exec);
            (void)hipGraphDestroy(plan->graph);
            plan->graph_exec = nullptr;
            plan->graph = nullptr;
        }
        if (!plan->graph_exec) {
            // capture on a private stream (the null stream cannot be captured)
            if (!plan->capture_stream) VRT_HIP(hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking));
            plan->graph_generation = plan->world->generation;
            VRT_HIP(hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_plan(plan, nullptr, plan->capture_stream);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(plan->capture_stream, &g);
            if (rc != VRT_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            plan->graph = g;
            VRT_HIP(hipGraphInstantiate(&plan->graph_exec, plan->graph, nullptr, nullptr, 0));
        }
        VRT_HIP(hipGraphLaunch(plan->graph_exec, (hipStream_t)stream));
        return VRT_OK;
    });
}

int vrt_batch_plan_destroy(vrt_batch_plan* plan) {
    return guarded([&]() -> int {
        if (!plan) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_batch_plan_destroy: null plan");
        (void)hipDeviceSynchronize();
        destroy_plan(plan);
        return VRT_OK;
    });
}

int vrt_batch_autotune(const vrt_world* world, const vrt_batch* batch, const char* cache_path,
                       vrt_batch_tuning* best, float* best_ms, void* stream) {
    return guarded([&]() -> int {
        if (!world || !batch || !best) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_batch_autotune: null argument");
        if (int rcw = check_world(world, "vrt_batch_autotune"); rcw != VRT_OK) return rcw;
        hipStream_t s = (hipStream_t)stream;
        std::string key = device_name(world->device) + ":" + batch_signature(world, batch);
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune_cache.find(key);
            if (it == g_tune_cache.end() && cache_path) {
                std::ifstream in(cache_path);
                std::string k;
                vrt_batch_tuning t;
                float ms;
                while (in >> k >> t.tile_w >> t.tile_h >> t.rays_per_lane >> t.chunk >> ms)
                    if (k == key) g_tune_cache[key] = {t, ms};
                it = g_tune_cache.find(key);
            }
            if (it != g_tune_cache.end()) {
                *best = it->second.first;
                if (best_ms) *best_ms = it->second.second;
                return VRT_OK;
            }
        }
        const uint32_t tiles[2][2] = {{8, 8}, {16, 4}};
        const uint32_t rpls[2] = {1, 2};
        const uint32_t chunks[3] = {0, 64, 256};
        float best_t = 1e30f;
        vrt_batch_tuning bt{8, 8, 1, 0};
        hipEvent_t e0, e1;
        VRT_HIP(hipEventCreate(&e0));
        VRT_HIP(hipEventCreate(&e1));
        int rc = VRT_OK;
        for (auto& tl : tiles)
            for (uint32_t rpl : rpls)
                for (uint32_t ch : chunks) {
                    if (rc != VRT_OK) break;
                    vrt_batch_tuning t{tl[0], tl[1], rpl, ch};
                    vrt_batch_plan* p = nullptr;
                    int r = create_plan(world, batch, &t, s, &p);
                    if (r == VRT_ERR_INVALID_ARGUMENT) continue;  // e.g. a camera not divisible by this tile
                    if (r != VRT_OK) {
                        rc = r;
                        break;
                    }
                    float ms = 0.0f;
                    const int reps = 5;
                    r = launch_plan(p, nullptr, s);  // warm-up
                    if (r == VRT_OK) {
                        (void)hipEventRecord(e0, s);
                        for (int k = 0; k < reps && r == VRT_OK; k++) r = launch_plan(p, nullptr, s);
                        (void)hipEventRecord(e1, s);
                        (void)hipEventSynchronize(e1);
                        (void)hipEventElapsedTime(&ms, e0, e1);
                        ms /= reps;
                    }
                    destroy_plan(p);
                    if (r != VRT_OK) {
                        rc = r;
                        break;
                    }
                    if (ms < best_t) {
                        best_t = ms;
                        bt = t;
                    }
                }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != VRT_OK) return rc;
        if (best_t >= 1e30f) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_batch_autotune: no tuning fits this batch");
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[key] = {bt, best_t};
        }
        if (cache_path) {
            std::ofstream outf(cache_path, std::ios::app);
            outf << key << " " << bt.tile_w << " " << bt.tile_h << " " << bt.rays_per_lane << " " << bt.chunk << " "
                 << best_t << "\n";
        }
        *best = bt;
        if (best_ms) *best_ms = best_t;
        return VRT_OK;
    });
}

}  // extern "C"
