// cast.hip - plain per-ray cast kernels and their C ABI (vrt_cast, vrt_cast_ex,
// vrt_cast_one_f64).
//
// vrt_cast replaces the reference's RayCast (src/VoxelRT/CpuRenderer.cpp:172-224)
// for a buffer of rays; vrt_cast_one_f64 replaces VoxelMap::RayCast
// (src/VoxelRT/VoxelMap.cpp:140-170).  The schedules compute identical results:
//   GRID        one lane per ray, one wave per workgroup, the grid covers n rays
//               (blocks of 64 rays in order or last block first); the default;
//   PERSISTENT  a grid sized to the device; each wave takes chunks of rays from
//               a global counter into a wave-local pool and refills lanes whose
//               ray finished (ballot + mbcnt rank), so long rays do not keep
//               the other 63 lanes of their wave idle;
//   COMPACT     a GRID pass capped at K steps, then the rays still running packed
//               64 to a wave and resumed from their saved state by a second pass.
// On the measured workloads the hardware dispatcher balances GRID's plain waves
// better than the persistent queue does; PERSISTENT is kept as an option.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>

#include "traverse.h"
#include "vrt_cast_impl.h"

namespace vrt {

// ---------------------------------------------------------------- counters

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ __forceinline__ void flush_counters(vrt_cast_counters* ctr, unsigned long long rays, unsigned long long iters,
                                               unsigned long long desc, unsigned long long hits) {
    rays = wave_sum(rays);
    iters = wave_sum(iters);
    desc = wave_sum(desc);
    hits = wave_sum(hits);
    if (__lane_id() == 0) {
        atomicAdd(&ctr->rays, rays);
        atomicAdd(&ctr->iters, iters);
        atomicAdd(&ctr->descents, desc);
        atomicAdd(&ctr->hits, hits);
    }
}

__device__ __forceinline__ void store_result(const HitsOut& h, uint32_t i, const RayState& r, float t, int4 rec) {
    h.t[i] = t;
    h.rec[i] = rec;
    const bool hit = (rec.w & 0xFF) != 0;
    if (h.pos_x) {
        h.pos_x[i] = hit ? r.cur[0] : 0.0f;
        h.pos_y[i] = hit ? r.cur[1] : 0.0f;
        h.pos_z[i] = hit ? r.cur[2] : 0.0f;
    }
    if (h.uv_u) {
        // UV of CpuRenderer.cpp:217-220: fract(X ? pos.y : pos.x), fract(Z ? pos.y : pos.z), the
        // tie masks X and Z being the nonzero normal fields of the record
        bool X = ((uint32_t)rec.w >> 8) & 3u, Z = ((uint32_t)rec.w >> 12) & 3u;
        float a = X ? r.cur[1] : r.cur[0], b = Z ? r.cur[1] : r.cur[2];
        h.uv_u[i] = hit ? a - floorf(a) : 0.0f;
        h.uv_v[i] = hit ? b - floorf(b) : 0.0f;
    }
    // the palette entry's low 32 bits (GetVoxelMaterial, CpuRenderer.cpp:120-132)
    if (h.material_data) h.material_data[i] = hit ? (uint32_t)h.palette[rec.w & 0xFF] : 0u;
}

// Traces one ray to completion.
__device__ __forceinline__ void trace_full(const WorldView& w, RayState& st, uint32_t max_iters, bool wave_tame,
                                           int origin, float& t, int4& rec) {
    int32_t v[3];
    uint32_t slot;
    if (ray_trace_wave(w, st, max_iters, v, slot, wave_tame, origin))
        ray_hit_record(w, st, v, slot, t, rec);
    else
        ray_miss_record(st, t, rec);
}

// ---------------------------------------------------------------- GRID schedule

// One wave per workgroup: a finished wave frees its slot for the next one at once
// (with four-wave workgroups the dispatcher waits for room for all four), which
// measures 2-3% faster on the 1024^3 terrain's primary and bounce rays.  Workgroup b
// runs on XCD b % 8, so consecutive tiles spread over the XCDs; mapping each XCD to
// a contiguous eighth of the rays instead measured 1.2x (primary) and 1.9x (bounce)
// slower: the eighths differ in work and the XCDs finish unevenly.
constexpr uint32_t kGridThreads = 64;

// Ray blocks are dispatched in order, or last block first (reverse): the results
// are the same, but the waves dispatched last make up the launch's tail, where the
// chip drains.  A camera's last rows (the ground, with descents and hits) cost the
// most per wave and its first rows (sky) the least, so the reverse order leaves cheap
// waves for the tail: 0.118 against 0.139 ms for the 1920x1080 primary rays.  A random
// block order (0.145 ms) loses the L2 reuse between neighbouring tiles.
template <bool COUNT>
__global__ __launch_bounds__(kGridThreads) void cast_grid_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t max_iters,
                                                                 vrt_cast_counters* ctr, uint32_t reverse) {
    uint32_t blk = reverse ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
    uint32_t i = blk * kGridThreads + threadIdx.x;
    bool valid = i < ray_count(r);
    RayState st;
    float t = 0.0f;
    int4 rec = make_int4(0, 0, 0, 0);
    if (valid) {
        ray_setup(st, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.W[0], r.W[1], r.W[2]);
        // the fast loop when every ray of the wave is tame, else the exact one
        trace_full(w, st, max_iters, __all(ray_is_tame(st)), origin_kind(r.W[0], r.W[1], r.W[2]), t, rec);
        store_result(h, i, st, t, rec);
    }
    if constexpr (COUNT) {
        bool hit = valid && (rec.w & 0xFF) != 0;
        flush_counters(ctr, valid ? 1ull : 0ull, valid ? st.iters : 0ull, valid ? st.descents : 0ull, hit ? 1ull : 0ull);
    }
}

// ---------------------------------------------------------------- COMPACT schedule

// Cap and compact.  Pass 1 is the grid kernel with the iteration cap lowered to K: a
// lane still running after K steps saves its traversal state (position and side
// distances) at its ray index instead of a result, and each wave writes the ballot of
// its survivors.  Pass 2 repacks them: the survivors of a group of 64 pass-1 waves, in
// order, are resumed 64 to a wave with the step count at K.  A step reads the ray, its
// position and the iteration count, and a hit record reads the side distances of the
// step before the hit, so a resumed ray makes the single pass's steps and writes its
// record.  Incoherent rays finish at very different steps (a wave's lanes do 11-32% of
// its steps on the 2048^3 sphere worlds' incoherent sets and the 1024^3 terrain's bounce
// set); the long rays then run in full waves.  Every pass-1 wave writes its ballot and
// the packing is recomputed from the ballots, so there are no atomics (one shared
// counter serializes past L2 on this part) and nothing to clear between launches.
//
// A survivor is a ray that neither hit nor left the world in its K steps: in every loop
// form, a ray that stops earlier has made fewer than K steps that went on, or has hit.
constexpr uint32_t kCompactGroup = 64;  // pass-1 waves whose survivors one pass-2 group packs

struct SurvivorState {
    uint64_t* __restrict__ ballot;  // per ray block (pass-1 wave): the lanes that survived
    float4* __restrict__ a;         // per ray: position, side distance x
    float2* __restrict__ b;         // per ray: side distances y, z
};

__device__ __forceinline__ void finish_record(const WorldView& w, const RayState& st, bool hit, const int32_t v[3],
                                              uint32_t slot, float& t, int4& rec) {
    if (hit)
        ray_hit_record(w, st, v, slot, t, rec);
    else
        ray_miss_record(st, t, rec);
}

template <bool COUNT>
__global__ __launch_bounds__(kGridThreads) void cast_pass1_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t cap,
                                                                  vrt_cast_counters* ctr, uint32_t reverse,
                                                                  SurvivorState q) {
    uint32_t blk = reverse ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
    uint32_t i = blk * kGridThreads + threadIdx.x;
    bool valid = i < ray_count(r);
    RayState st;
    bool hit = false, survivor = false;
    if (valid) {
        ray_setup(st, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.W[0], r.W[1], r.W[2]);
        int32_t v[3];
        uint32_t slot = 0;
        hit = ray_trace_wave(w, st, cap, v, slot, __all(ray_is_tame(st)), origin_kind(r.W[0], r.W[1], r.W[2]));
        survivor = !hit && st.iters == cap;
        if (survivor) {
            q.a[i] = make_float4(st.cur[0], st.cur[1], st.cur[2], st.side[0]);
            q.b[i] = make_float2(st.side[1], st.side[2]);
        } else {
            float t;
            int4 rec;
            finish_record(w, st, hit, v, slot, t, rec);
            store_result(h, i, st, t, rec);
        }
    }
    const uint64_t b = __ballot(survivor);
    if (threadIdx.x == 0) q.ballot[blk] = b;
    if constexpr (COUNT) {
        // a survivor's steps and descents so far; pass 2 adds the rest and its hit
        flush_counters(ctr, valid ? 1ull : 0ull, valid ? st.iters : 0ull, valid ? st.descents : 0ull,
                       hit ? 1ull : 0ull);
    }
}

// Position of the k-th set bit of m (k < popcount(m)), by halving.
__device__ __forceinline__ uint32_t select_bit(uint64_t m, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t width = 32; width > 0; width >>= 1) {
        const uint64_t low = m & ((1ull << width) - 1ull);
        const uint32_t c = (uint32_t)__popcll(low);
        if (k >= c) {
            k -= c;
            m >>= width;
            pos += width;
        } else {
            m = low;
        }
    }
    return pos;
}

// Pass 2: wave (round, group) resumes survivors round*64 .. round*64+63 of its group of 64
// pass-1 waves, counted in ray order.  Rounds are ngroups waves apart, so every group's
// first round is dispatched before any second round; a round past the group's survivors
// exits after the 64 ballot loads.
template <bool COUNT>
__global__ __launch_bounds__(kGridThreads) void cast_pass2_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t cap,
                                                                  uint32_t max_iters, vrt_cast_counters* ctr,
                                                                  SurvivorState q, uint32_t nblocks,
                                                                  uint32_t ngroups) {
    const uint32_t g = blockIdx.x % ngroups, round = blockIdx.x / ngroups;
    const uint32_t lane = threadIdx.x;
    const uint32_t wv = g * kCompactGroup + lane;  // the pass-1 wave this lane sums
    const uint64_t m = wv < nblocks ? q.ballot[wv] : 0ull;
    const uint32_t c = (uint32_t)__popcll(m);
    uint32_t incl = c;  // survivors of pass-1 waves 0..lane of the group
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    if (round * 64u >= total) return;  // wave-uniform
    const uint32_t k = round * 64u + lane;  // this lane's survivor within the group
    // the pass-1 wave holding survivor k: the first one whose inclusive count exceeds k
    uint32_t src = 0;
#pragma unroll
    for (uint32_t step = 32; step > 0; step >>= 1) {
        if (__shfl(incl, (int)(src + step - 1), 64) <= k) src += step;
    }
    const uint32_t before = __shfl(incl, (int)src, 64) - __shfl(c, (int)src, 64);
    const uint64_t msrc = __shfl(m, (int)src, 64);
    unsigned long long c_iters = 0, c_desc = 0, c_hits = 0;
    if (k < total) {
        const uint32_t i = (g * kCompactGroup + src) * kGridThreads + select_bit(msrc, k - before);
        // every surviving lane of pass 1 had i below the ray count; the test keeps a ballot
        // that pass 1 did not rewrite (a failed launch) from reading past the rays
        if (i < r.n) {
            const float4 a = q.a[i];
            const float2 b = q.b[i];
            RayState st;
            ray_setup(st, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.W[0], r.W[1], r.W[2]);
            st.cur[0] = a.x;
            st.cur[1] = a.y;
            st.cur[2] = a.z;
            st.side[0] = a.w;
            st.side[1] = b.x;
            st.side[2] = b.y;
            st.iters = cap;
            int32_t v[3];
            uint32_t slot = 0;
            const bool hit = ray_trace_wave(w, st, max_iters, v, slot, __all(ray_is_tame(st)),
                                            origin_kind(r.W[0], r.W[1], r.W[2]), cap);
            float t;
            int4 rec;
            finish_record(w, st, hit, v, slot, t, rec);
            store_result(h, i, st, t, rec);
            if constexpr (COUNT) {
                c_iters = st.iters - cap;
                c_desc = st.descents;
                c_hits = hit ? 1ull : 0ull;
            }
        }
    }
    if constexpr (COUNT) flush_counters(ctr, 0ull, c_iters, c_desc, c_hits);
}

// ---------------------------------------------------------------- PERSISTENT schedule

// Work distribution: rays are cut into chunks; the chunk range is split into
// kQueues contiguous parts, one per XCD (workgroup b runs on XCD b % 8), so an
// XCD's L2 sees a compact part of the image.  Each wave owns its first chunk
// statically; further chunks come from its XCD's queue counter and, once that
// part is drained, from the other parts in turn.  The counters live on separate
// 256-byte lines, so claims from different XCDs do not serialize.
// Each wave keeps 64 ray slots.  A finished ray parks its result in registers;
// once at least refill_min lanes are free (or all of them), the wave writes the
// parked results and starts new rays in the free lanes.  Every lane also holds
// the origin and direction of its next ray, loaded at the previous refill, so a
// refill starts rays without waiting on memory; the lanes that consumed their
// prefetched ray issue the next loads and go on stepping while they arrive.
template <bool COUNT>
__global__ __launch_bounds__(256) void cast_persistent_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t max_iters,
                                                              uint32_t chunk, uint32_t refill_min, ChunkQueue cq,
                                                              vrt_cast_counters* ctr) {
    const uint32_t n = ray_count(r);
    const uint32_t home = blockIdx.x % kQueues;
    const uint32_t wave_in_part = (blockIdx.x / kQueues) * (blockDim.x / 64) + threadIdx.x / 64;
    uint32_t probe = 0;
    // static first chunk
    uint32_t first = part_begin(home, cq.nchunks) + wave_in_part;
    uint32_t pool_next = 0, pool_end = 0;  // wave-uniform
    if (wave_in_part < cq.waves_per_q && first < part_begin(home + 1, cq.nchunks)) {
        pool_next = min(first * chunk, n);  // chunks past a device-side count are empty
        pool_end = min(pool_next + chunk, n);
    }
    bool exhausted = false;  // wave-uniform: every part is drained
    // Gives the lanes in `need` consecutive ray indices from the wave's pool (one chunk
    // claim at most); a lane left over gets ~0u and asks again at the next refill.
    auto take = [&](bool need) -> uint32_t {
        uint64_t m = __ballot(need);
        if (m == 0) return ~0u;
        if (pool_next == pool_end && !exhausted) {
            uint32_t c = claim_chunk(cq, home, probe);
            if (c == ~0u) {
                exhausted = true;
            } else {
                pool_next = min(c * chunk, n);
                pool_end = min(pool_next + chunk, n);
            }
        }
        uint32_t avail = pool_end - pool_next;
        uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        uint32_t got = need && rank < avail ? pool_next + rank : ~0u;
        pool_next += min(avail, (uint32_t)__popcll(m));
        return got;
    };
    // the lane's next ray
    uint32_t pf = take(true);
    float po[3] = {0.0f, 0.0f, 0.0f}, pd[3] = {1.0f, 1.0f, 1.0f};
    if (pf != ~0u) {
        po[0] = r.ox[pf]; po[1] = r.oy[pf]; po[2] = r.oz[pf];
        pd[0] = r.dx[pf]; pd[1] = r.dy[pf]; pd[2] = r.dz[pf];
    }
    bool more = true;       // wave-uniform: some lane holds a next ray or the pool may refill
    bool tame = true;       // wave-uniform: every running ray is tame
    bool running = false;   // the lane is tracing a ray
    bool parked = false;    // the lane's ray finished; its result is not written yet
    bool hit = false;
    uint32_t idx = 0;
    int32_t v[3] = {0, 0, 0};
    RayState st;
    unsigned long long c_rays = 0, c_iters = 0, c_desc = 0, c_hits = 0;
    for (;;) {
        uint32_t nfree = (uint32_t)__popcll(__ballot(!running));
        if (nfree == 64 || (nfree >= refill_min && more)) {
            if (parked) {
                float t;
                int4 rec;
                if (hit)
                    ray_hit_record(w, st, v, brick_slot_at(w, v), t, rec);
                else
                    ray_miss_record(st, t, rec);
                store_result(h, idx, st, t, rec);
                if constexpr (COUNT) {
                    c_rays++;
                    c_iters += st.iters;
                    c_desc += st.descents;
                    c_hits += hit ? 1ull : 0ull;
                }
                parked = false;
            }
            if (!running && pf != ~0u) {
                idx = pf;
                ray_setup(st, po[0], po[1], po[2], pd[0], pd[1], pd[2], r.W[0], r.W[1], r.W[2]);
                running = true;
                hit = false;
                pf = ~0u;
            }
            uint32_t next = take(pf == ~0u);
            if (pf == ~0u && next != ~0u) {
                pf = next;
                po[0] = r.ox[pf]; po[1] = r.oy[pf]; po[2] = r.oz[pf];
                pd[0] = r.dx[pf]; pd[1] = r.dy[pf]; pd[2] = r.dz[pf];
            }
            more = __any(pf != ~0u) || !exhausted;
            tame = !__any(running && !ray_is_tame(st));
            if (!__any(running)) {
                if (!more) break;
                continue;
            }
        }
        if (running) {
            uint32_t unused;
            bool cont = tame ? ray_advance<false>(w, st, v, hit, &unused) : ray_advance<true>(w, st, v, hit, &unused);
            if (!cont || st.iters >= max_iters) {
                running = false;
                parked = true;
            }
        }
    }
    if constexpr (COUNT) flush_counters(ctr, c_rays, c_iters, c_desc, c_hits);
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- device properties and queues

static std::mutex g_dev_mu;
static int g_cus[64];
// Queue counters of persistent launches, one block per (device, stream): launches on one
// stream run one after another, so they can share a block; launches on different streams
// never do.  Blocks are kept until the process ends (2 KiB each).
static std::map<std::pair<int, hipStream_t>, uint32_t*> g_queues;

int device_info(int dev, int* cus, uint32_t** queue, hipStream_t s) {
    if (dev < 0 || dev >= 64) return fail(VRT_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (!g_cus[dev]) {
        hipDeviceProp_t prop;
        VRT_HIP(hipGetDeviceProperties(&prop, dev));
        g_cus[dev] = prop.multiProcessorCount;
    }
    *cus = g_cus[dev];
    if (queue) {
        uint32_t*& q = g_queues[{dev, s}];
        if (!q) VRT_HIP(hipMalloc(&q, (size_t)kQueues * kQueueStride * sizeof(uint32_t)));
        *queue = q;
    }
    return VRT_OK;
}

// Survivor state of compact casts, one buffer per (device, stream): a ballot per ray block,
// then 16 + 8 bytes per ray of capacity.  A cast with more rays than the buffer holds
// replaces it, after synchronizing the stream (whose queued launches may still use the old
// one).  Kept until the process ends.
struct SurvivorBuffer {
    void* base = nullptr;
    uint64_t cap = 0;  // rays, a multiple of the block size
};
static std::map<std::pair<int, hipStream_t>, SurvivorBuffer> g_survivors;

static int survivor_state(int dev, hipStream_t s, uint32_t n, SurvivorState* q) {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    SurvivorBuffer& b = g_survivors[{dev, s}];
    const uint64_t need = ((uint64_t)n + kGridThreads - 1) / kGridThreads * kGridThreads;
    if (b.cap < need) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        VRT_HIP(hipStreamIsCapturing(s, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(VRT_ERR_STATE, "vrt_cast_ex: the compact schedule's survivor buffer for this stream is "
                                       "smaller than the ray count; make one such cast before capturing");
        if (b.base) {
            VRT_HIP(hipStreamSynchronize(s));
            VRT_HIP(hipFree(b.base));
            b.base = nullptr;
            b.cap = 0;
        }
        void* p = nullptr;
        VRT_HIP(hipMalloc(&p, (size_t)need / kGridThreads * 8 + (size_t)need * 24));
        b.base = p;
        b.cap = need;
    }
    const size_t nblk = (size_t)b.cap / kGridThreads;
    q->ballot = (uint64_t*)b.base;
    q->a = (float4*)((char*)b.base + nblk * 8);
    q->b = (float2*)((char*)b.base + nblk * 8 + (size_t)b.cap * 16);
    return VRT_OK;
}

// ---------------------------------------------------------------- launch

// First-pass steps of the compact schedule when the caller gives 0: within 1% of the best
// of 24-64 on 2 M incoherent rays in the 2048^3 world of 20 000 spheres (0.317 ms, 0.314 at
// 24, against 0.341 ms for the grid); 64 is best in the 587-sphere world (0.432 against
// 0.448 ms).
constexpr uint32_t kDefaultPassCap = 32;

// Option checks that do not depend on the ray count, so a call is rejected or
// accepted the same way whatever n is (n = 0 included).
int check_cast_opts(const vrt_cast_opts& o) {
    uint32_t sched = o.schedule == VRT_SCHED_AUTO ? VRT_SCHED_GRID : o.schedule;
    if (sched != VRT_SCHED_GRID && sched != VRT_SCHED_PERSISTENT && sched != VRT_SCHED_COMPACT)
        return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: unknown schedule");
    if (o.order > VRT_ORDER_REVERSE) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: unknown order");
    if (sched == VRT_SCHED_PERSISTENT && (o.chunk % 64 != 0 || o.chunk > (1u << 20)))
        return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: chunk must be a multiple of 64");
    return VRT_OK;
}

int launch_cast(const vrt_world* world, const RaysIn& r, const HitsOut& h, uint32_t max_iters,
                const vrt_cast_opts& o, hipStream_t s) {
    WorldView w = world->view();
    int rc0 = check_cast_opts(o);
    if (rc0 != VRT_OK) return rc0;
    if (r.n == 0) return VRT_OK;
    uint32_t sched = o.schedule == VRT_SCHED_AUTO ? VRT_SCHED_GRID : o.schedule;
    const uint32_t cap = o.pass_cap ? o.pass_cap : kDefaultPassCap;
    // a first pass as long as the cast is the whole cast
    if (sched == VRT_SCHED_COMPACT && cap >= max_iters) sched = VRT_SCHED_GRID;
    if (sched == VRT_SCHED_COMPACT) {
        SurvivorState q;
        int rc;
        rc = survivor_state(world->device, s, r.n, &q);
        if (rc != VRT_OK) return rc;
        const uint32_t blocks = (uint32_t)(((uint64_t)r.n + kGridThreads - 1) / kGridThreads);
        const uint32_t rev = o.order == VRT_ORDER_REVERSE ? 1u : 0u;
        // pass 2: 64 rounds per group of 64 ray blocks (a round past a group's survivors exits)
        const uint32_t ngroups = (blocks + kCompactGroup - 1) / kCompactGroup;
        const uint32_t waves2 = ngroups * 64u;
        if (o.counters)
            cast_pass1_kernel<true><<<blocks, kGridThreads, 0, s>>>(w, r, h, cap, o.counters, rev, q);
        else
            cast_pass1_kernel<false><<<blocks, kGridThreads, 0, s>>>(w, r, h, cap, nullptr, rev, q);
        // pass 2 reads the ballots pass 1 writes: no pass 2 after a failed pass 1
        VRT_HIP(hipGetLastError());
        if (o.counters)
            cast_pass2_kernel<true><<<waves2, kGridThreads, 0, s>>>(w, r, h, cap, max_iters, o.counters, q, blocks,
                                                                   ngroups);
        else
            cast_pass2_kernel<false><<<waves2, kGridThreads, 0, s>>>(w, r, h, cap, max_iters, nullptr, q, blocks,
                                                                    ngroups);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    if (sched == VRT_SCHED_GRID) {
        uint32_t blocks = (uint32_t)(((uint64_t)r.n + kGridThreads - 1) / kGridThreads);
        uint32_t rev = o.order == VRT_ORDER_REVERSE ? 1u : 0u;
        if (o.counters)
            cast_grid_kernel<true><<<blocks, kGridThreads, 0, s>>>(w, r, h, max_iters, o.counters, rev);
        else
            cast_grid_kernel<false><<<blocks, kGridThreads, 0, s>>>(w, r, h, max_iters, nullptr, rev);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    uint32_t chunk = o.chunk ? o.chunk : 64;
    if (r.n > 0x7FFFFFFFu) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: persistent schedule takes at most 2^31 - 1 rays");
    int cus = 0;
    uint32_t* queue = nullptr;
    int rc = device_info(world->device, &cus, &queue, s);
    if (rc != VRT_OK) return rc;
    uint32_t wpc = o.waves_per_cu ? std::min<uint32_t>(o.waves_per_cu, 32) : 32;
    // whole rounds of kQueues workgroups so every XCD part gets the same number of waves
    uint32_t blocks = std::max<uint32_t>(kQueues, ((uint32_t)cus * wpc / 4) / kQueues * kQueues);
    ChunkQueue cq;
    cq.ctr = queue;
    cq.nchunks = (r.n + chunk - 1) / chunk;
    cq.waves_per_q = blocks / kQueues * 4;
    uint32_t refill_min = o.refill ? std::min<uint32_t>(std::max<uint32_t>(o.refill, 1), 64) : 16;
    // a fill kernel, not a memset: the launch may be captured into a graph
    fill_u32_kernel<<<(kQueues * kQueueStride + 255) / 256, 256, 0, s>>>(queue, kQueues * kQueueStride, 0u);
    if (o.counters)
        cast_persistent_kernel<true><<<blocks, 256, 0, s>>>(w, r, h, max_iters, chunk, refill_min, cq, o.counters);
    else
        cast_persistent_kernel<false><<<blocks, 256, 0, s>>>(w, r, h, max_iters, chunk, refill_min, cq, nullptr);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int check_rays_hits(const vrt_rays* rays, const vrt_hits* hits, RaysIn& r, HitsOut& h, bool extras) {
    if (!rays || !hits) return fail(VRT_ERR_INVALID_ARGUMENT, "null rays or hits");
    if (!extras && (hits->uv || hits->material_data))
        return fail(VRT_ERR_INVALID_ARGUMENT, "uv and material_data are filled by vrt_cast / vrt_cast_ex only");
    if (rays->n && (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz))
        return fail(VRT_ERR_INVALID_ARGUMENT, "null ray buffer");
    if (rays->n && (!hits->t || !hits->rec)) return fail(VRT_ERR_INVALID_ARGUMENT, "null hit buffer");
    if (((uintptr_t)hits->rec & 15) != 0) return fail(VRT_ERR_INVALID_ARGUMENT, "hit records must be 16-byte aligned");
    r.ox = rays->ox; r.oy = rays->oy; r.oz = rays->oz;
    r.dx = rays->dx; r.dy = rays->dy; r.dz = rays->dz;
    r.W[0] = rays->world_origin[0]; r.W[1] = rays->world_origin[1]; r.W[2] = rays->world_origin[2];
    r.n = rays->n;
    r.count = nullptr;
    h.t = hits->t;
    h.rec = (int4*)hits->rec;
    if (hits->pos) {
        h.pos_x = hits->pos;
        h.pos_y = hits->pos + rays->n;
        h.pos_z = hits->pos + 2 * (size_t)rays->n;
    } else {
        h.pos_x = h.pos_y = h.pos_z = nullptr;
    }
    h.uv_u = hits->uv;
    h.uv_v = hits->uv ? hits->uv + rays->n : nullptr;
    h.material_data = hits->material_data;
    h.palette = nullptr;
    return VRT_OK;
}

// ---------------------------------------------------------------- FP64 picking cast

// VoxelMap::RayCast (VoxelMap.cpp:140-170) with GetStepLevel (:125-139): a
// single ray in FP64, bias 1e-4, fixed step levels 32 / 8 / 1 voxels.  Sector
// coordinates wrap like WorldSectorIndexer<12,8,signed> (VoxelMap.h:100).
__device__ __forceinline__ int32_t d2i_floor(double x) {
    double f = floor(x);
    return fabs(f) < 2147483648.0 ? (int32_t)f : INT32_MIN;
}

__global__ void cast_one_f64_kernel(WorldView w, double ox, double oy, double oz, double dx, double dy, double dz,
                                    uint32_t max_iters, vrt_hit_f64* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double o[3] = {ox, oy, oz}, d[3] = {dx, dy, dz}, inv[3], ts[3];
    int32_t p[3];
    for (int a = 0; a < 3; a++) {
        inv[a] = 1.0 / d[a];
        ts[a] = ((d[a] < 0.0 ? 0.0 : 1.0) - o[a]) * inv[a];
        p[a] = d2i_floor(o[a]);
    }
    vrt_hit_f64 res;
    res.distance = -1.0;
    for (int a = 0; a < 3; a++) { res.normal[a] = 0.0f; res.voxel[a] = 0; }
    res.uv[0] = res.uv[1] = 0.0f;
    uint32_t nxz = 1u << w.shift_xz, ny = 1u << w.shift_y;
    for (uint32_t i = 0; i < max_iters; i++) {
        double side[3], hp[3];
        for (int a = 0; a < 3; a++) side[a] = fma((double)p[a], inv[a], ts[a]);
        double tmin = side[0];
        tmin = side[1] < tmin ? side[1] : tmin;  // glm::min(a, b) = b < a ? b : a
        tmin = side[2] < tmin ? side[2] : tmin;
        tmin += 0.0001;
        for (int a = 0; a < 3; a++) {
            hp[a] = fma(tmin, d[a], o[a]);
            p[a] = d2i_floor(hp[a]);
        }
        // GetStepLevel
        int k;
        uint32_t sx = (uint32_t)(p[0] >> 5) & 4095u, sy = (uint32_t)(p[1] >> 5) & 255u, sz = (uint32_t)(p[2] >> 5) & 4095u;
        uint64_t m = 0;
        uint32_t si = 0;
        if (sx < nxz && sz < nxz && sy < ny) {
            si = sx | sz << w.shift_xz | sy << (2 * w.shift_xz);
            m = (uint64_t)w.sec[si].x | (uint64_t)w.sec[si].y << 32;
        }
        if (m == 0) {
            k = 5;
        } else {
            uint32_t bi = ((uint32_t)(p[0] >> 3) & 3u) | ((uint32_t)(p[2] >> 3) & 3u) << 2 | ((uint32_t)(p[1] >> 3) & 3u) << 4;
            if (!((m >> bi) & 1ull)) {
                k = 3;
            } else {
                uint32_t slot = w.sec[si].z + (uint32_t)__popcll(m & ((1ull << bi) - 1ull));
                uint32_t vi = ((uint32_t)p[0] & 7u) | ((uint32_t)p[2] & 7u) << 3 | ((uint32_t)p[1] & 7u) << 6;
                k = w.vox[(size_t)slot * 512 + vi] == 0 ? 0 : -1;
            }
        }
        if (k < 0) {
            bool sm[3];
            for (int a = 0; a < 3; a++) {
                sm[a] = tmin >= side[a];
                double sg = d[a] > 0 ? 1.0 : (d[a] < 0 ? -1.0 : 0.0);
                res.normal[a] = sm[a] ? (float)(-sg) : 0.0f;
                res.voxel[a] = p[a];
            }
            float u = sm[0] ? (float)hp[1] : (float)hp[0];
            float v = sm[2] ? (float)hp[1] : (float)hp[2];
            res.uv[0] = u - floorf(u);
            res.uv[1] = v - floorf(v);
            res.distance = tmin;
            break;
        }
        int32_t mk = (int32_t)((1u << k) - 1u);
        for (int a = 0; a < 3; a++) p[a] = d[a] < 0 ? (p[a] & ~mk) : (p[a] | mk);
    }
    *out = res;
}

}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_cast(const vrt_world* world, const vrt_rays* rays, uint32_t max_iters, vrt_hits* hits, void* stream) {
    vrt_cast_opts o;
    std::memset(&o, 0, sizeof(o));
    return vrt_cast_ex(world, rays, max_iters, hits, &o, stream);
}

int vrt_stream_release(void* stream) {
    return guarded([&]() -> int {
        const hipStream_t s = (hipStream_t)stream;
        std::lock_guard<std::mutex> lk(g_dev_mu);
        bool synced = false;
        auto sync_once = [&]() -> int {
            if (!synced) VRT_HIP(hipStreamSynchronize(s));
            synced = true;
            return VRT_OK;
        };
        for (auto it = g_queues.begin(); it != g_queues.end();) {
            if (it->first.second != s) {
                ++it;
                continue;
            }
            if (int rc = sync_once(); rc != VRT_OK) return rc;
            VRT_HIP(hipFree(it->second));
            it = g_queues.erase(it);
        }
        for (auto it = g_survivors.begin(); it != g_survivors.end();) {
            if (it->first.second != s) {
                ++it;
                continue;
            }
            if (int rc = sync_once(); rc != VRT_OK) return rc;
            VRT_HIP(hipFree(it->second.base));
            it = g_survivors.erase(it);
        }
        return VRT_OK;
    });
}

int vrt_cast_ex(const vrt_world* world, const vrt_rays* rays, uint32_t max_iters, vrt_hits* hits,
                const vrt_cast_opts* opts, void* stream) {
    return guarded([&]() -> int {
        if (!world) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast: null world");
        if (int rcw = check_world(world, "vrt_cast"); rcw != VRT_OK) return rcw;
        if (max_iters == 0 || max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast: max_iters must be 1..65535");
        if (fault_injected("vrt_cast")) return fail(VRT_ERR_HIP, "vrt_cast: injected HIP fault");
        RaysIn r;
        HitsOut h;
        int rc = check_rays_hits(rays, hits, r, h, true);
        if (rc != VRT_OK) return rc;
        h.palette = world->d_palette;
        vrt_cast_opts o;
        std::memset(&o, 0, sizeof(o));
        if (opts) o = *opts;
        r.count = o.count;
        return launch_cast(world, r, h, max_iters, o, (hipStream_t)stream);
    });
}

int vrt_cast_one_f64(const vrt_world* world, const double origin[3], const double dir[3], uint32_t max_iters,
                     vrt_hit_f64* out) {
    return guarded([&]() -> int {
        if (!world || !origin || !dir || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_one_f64: null argument");
        if (int rcw = check_world(world, "vrt_cast_one_f64"); rcw != VRT_OK) return rcw;
        if (world->log2_y > 13) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_one_f64: world taller than 8192 voxels");
        // the loop has no exit for leaving the world (VoxelMap.cpp:140-170): every step up to the
        // cap runs on one lane, so the cap is bounded like every other cast's
        if (max_iters == 0 || max_iters > 65535)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_one_f64: max_iters must be 1..65535");
        vrt_hit_f64* d_out = nullptr;
        VRT_HIP(hipMalloc(&d_out, sizeof(vrt_hit_f64)));
        cast_one_f64_kernel<<<1, 64, 0, 0>>>(world->view(), origin[0], origin[1], origin[2], dir[0], dir[1], dir[2],
                                             max_iters, d_out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(out, d_out, sizeof(vrt_hit_f64), hipMemcpyDeviceToHost);
        (void)hipFree(d_out);
        if (e != hipSuccess) return hip_fail(e, "vrt_cast_one_f64");
        return VRT_OK;
    });
}

}  // extern "C"
