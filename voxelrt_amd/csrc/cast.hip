// cast.hip - plain per-ray cast kernels and their C ABI (vrt_cast, vrt_cast_ex,
// vrt_cast_one_f64).
//
// vrt_cast replaces the reference's RayCast (src/VoxelRT/CpuRenderer.cpp:172-224)
// for a buffer of rays; vrt_cast_one_f64 replaces VoxelMap::RayCast
// (src/VoxelRT/VoxelMap.cpp:140-170).  Every schedule computes identical results:
//   GRID        one lane per ray, one wave per workgroup, the grid covers n rays
//               (blocks of 64 rays in order or last block first); the default;
//   PERSISTENT  a grid sized to the device; each wave takes chunks of rays from
//               a global counter into a wave-local pool and refills lanes whose
//               ray finished (ballot + mbcnt rank), so long rays do not keep
//               the other 63 lanes of their wave idle;
//   LANES       each lane traces K rays back to back, the waiting ones in LDS;
//   REPACK      workgroups repack their running rays through LDS every few lookups;
//   MULTIPASS   passes of a few lookups, running rays compacted between passes.
// The alternatives to GRID are kept as options: on the measured workloads the
// hardware dispatcher balances plain waves better than they do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
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
    if (h.pos_x) {
        h.pos_x[i] = rec.w & 0xFF ? r.cur[0] : 0.0f;
        h.pos_y[i] = rec.w & 0xFF ? r.cur[1] : 0.0f;
        h.pos_z[i] = rec.w & 0xFF ? r.cur[2] : 0.0f;
    }
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

// ---------------------------------------------------------------- LANES schedule

// Each lane traces K rays one after another: a wave owns 64 K consecutive rays and
// lane l takes rays l, l + 64, ..., l + 64 (K - 1) (one 8x8 tile after another), so
// a lane whose ray ends starts its next ray at once, and the wave runs for the
// longest per-lane sum of K ray lengths instead of the longest ray in each of K
// waves.  The waiting rays are set up before the loop and wait in LDS, and a
// finished ray parks its raw result there: LDS traffic counts on lgkmcnt, so the
// loop's vmcnt waits stay those of its own mask loads (global loads and stores in
// the loop would join that in-order queue).  Hit records are made after the loop.
constexpr uint32_t kLanesMaxRays = 8;
// LDS words per lane and ray (rows of 64 words per wave): waiting, the ray's o, d, 1/d,
// tStart and corner selectors; parked, its hit voxel, sideDist, position and lookups
// over o, 1/d and tStart (d stays for the normal's signs).
constexpr uint32_t kLaneWords = 15;
__host__ __device__ constexpr uint32_t lanes_lds_bytes(uint32_t k) { return k * kLaneWords * 64 * 4; }

// Row offsets (in words of the lane's column) of a ray's LDS slot.
enum : uint32_t {
    kLaneO = 0, kLaneD = 3 * 64, kLaneInv = 6 * 64, kLaneTs = 9 * 64, kLaneDm = 12 * 64,  // waiting
    kLaneV = 0, kLaneSide = 6 * 64, kLanePos = 9 * 64, kLaneIt = 12 * 64                  // parked
};

template <bool EXACT, bool CUBIC, bool COUNT>
__device__ __forceinline__ void lanes_trace(const WorldView& w, RayState& st, uint32_t* slot, const uint32_t* end,
                                            uint32_t max_iters, unsigned long long (&c)[4]) {
    // slot: the running ray's LDS slot; end: one past the lane's last slot
    while (slot != end) {
        int32_t v[3];
        bool hit = false;
        uint32_t unused;
        if (ray_advance<EXACT, CUBIC>(w, st, v, hit, &unused) && st.iters < max_iters) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            slot[kLaneV + a * 64] = (uint32_t)v[a];
            slot[kLaneSide + a * 64] = __float_as_uint(st.side[a]);
            slot[kLanePos + a * 64] = __float_as_uint(st.cur[a]);
        }
        slot[kLaneIt] = st.iters;
        if constexpr (COUNT) {
            c[0]++;
            c[1] += st.iters;
            c[2] += st.descents;
            c[3] += hit ? 1ull : 0ull;
        }
        slot += kLaneWords * 64;
        if (slot != end) {
            // sideDist keeps the last ray's values: it is only read after a step, and
            // a hit on the first lookup is told by its lookup count
#pragma unroll
            for (int a = 0; a < 3; a++) {
                st.o[a] = __uint_as_float(slot[kLaneO + a * 64]);
                st.cur[a] = st.o[a];
                st.d[a] = __uint_as_float(slot[kLaneD + a * 64]);
                st.inv[a] = __uint_as_float(slot[kLaneInv + a * 64]);
                st.ts[a] = __uint_as_float(slot[kLaneTs + a * 64]);
                st.dm[a] = (int32_t)slot[kLaneDm + a * 64];
            }
            st.iters = 0;
            st.descents = 0;
        }
    }
}

template <bool COUNT>
__global__ __launch_bounds__(64) void cast_lanes_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t max_iters,
                                                        uint32_t K, vrt_cast_counters* ctr) {
    extern __shared__ uint32_t lanes_lds[];
    const uint32_t lane = threadIdx.x;
    uint32_t* lds = lanes_lds + lane;
    const uint32_t n = ray_count(r);
    const uint32_t base = blockIdx.x * 64u * K + lane;
    // rays of this lane: base + 64 j for j < nrays
    const uint32_t nrays = base >= n ? 0u : min(K, (n - base + 63u) / 64u);
    RayState st;
    bool tame = true;
    for (uint32_t j = 0; j < K; j++) {
        if (j >= nrays) break;
        uint32_t i = base + 64u * j;
        RayState q;
        ray_setup(q, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.W[0], r.W[1], r.W[2]);
        tame = tame && ray_is_tame(q);
        uint32_t* sl = lds + j * kLaneWords * 64;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            sl[kLaneO + a * 64] = __float_as_uint(q.o[a]);
            sl[kLaneD + a * 64] = __float_as_uint(q.d[a]);
            sl[kLaneInv + a * 64] = __float_as_uint(q.inv[a]);
            sl[kLaneTs + a * 64] = __float_as_uint(q.ts[a]);
            sl[kLaneDm + a * 64] = (uint32_t)q.dm[a];
        }
        if (j == 0) st = q;
    }
    unsigned long long c[4] = {0, 0, 0, 0};
    uint32_t* end = lds + nrays * kLaneWords * 64;
    if (__all(tame)) {
        if (w.size_y == w.size_xz)
            lanes_trace<false, true, COUNT>(w, st, lds, end, max_iters, c);
        else
            lanes_trace<false, false, COUNT>(w, st, lds, end, max_iters, c);
    } else {
        lanes_trace<true, false, COUNT>(w, st, lds, end, max_iters, c);
    }
    // hit records from the parked results, in ray order per lane
    for (uint32_t j = 0; j < nrays; j++) {
        const uint32_t* sl = lds + j * kLaneWords * 64;
        RayState q;
        int32_t v[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = (int32_t)sl[kLaneV + a * 64];
            q.d[a] = __uint_as_float(sl[kLaneD + a * 64]);
            q.side[a] = __uint_as_float(sl[kLaneSide + a * 64]);
            q.cur[a] = __uint_as_float(sl[kLanePos + a * 64]);
        }
        q.iters = sl[kLaneIt];
        uint32_t bslot = 0;
        bool hit = solid_voxel_at(w, v, bslot);
        if (hit && q.iters == 1) {
            // a hit on the first lookup: sideDist was never set (RayCast starts it at 0)
#pragma unroll
            for (int a = 0; a < 3; a++) q.side[a] = 0.0f;
        }
        float t;
        int4 rec;
        if (hit)
            ray_hit_record(w, q, v, bslot, t, rec);
        else
            ray_miss_record(q, t, rec);
        store_result(h, base + 64u * j, q, t, rec);
    }
    if constexpr (COUNT) flush_counters(ctr, c[0], c[1], c[2], c[3]);
}

// ---------------------------------------------------------------- REPACK schedule

// A wave runs until its slowest ray.  Here a workgroup owns 256 consecutive rays
// and traces them in segments of `seg` lookups; after each segment the rays still
// running are written to LDS and read back densely into the workgroup's first
// waves, so the other waves idle at the barrier and their issue slots go to other
// workgroups.  A ray's state survives the move bit for bit: origin, direction,
// 1/d and sideDist are stored; tStart, the corner selector and the position are
// recomputed with the operations that produced them (the position from the exact
// minimum, which equals the fast one here).  Finished rays leave their results in
// LDS; the workgroup writes all 256 results, coalesced, at the end.
struct RepackLds {
    float o[3][256], d[3][256], inv[3][256], side[3][256];
    uint32_t it[256];    // lookups | descents << 16
    uint32_t slot[256];  // the ray's position in the workgroup
    int4 rec[256];
    float t[256];
    float pos[3][256];
    uint32_t wave_cnt[4];
};

// Up to seg lookups of one ray; true when it finished (hit, left the world or reached the cap).
template <bool EXACT>
__device__ __forceinline__ bool trace_segment(const WorldView& w, RayState& st, uint32_t seg, uint32_t max_iters,
                                              int32_t v[3], bool& hit) {
    uint32_t unused;
    for (uint32_t k = 0; k < seg; k++)
        if (!ray_advance<EXACT>(w, st, v, hit, &unused) || st.iters >= max_iters) return true;
    return false;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void cast_repack_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t max_iters,
                                                          uint32_t seg, vrt_cast_counters* ctr) {
    __shared__ RepackLds L;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t n = ray_count(r);
    const uint32_t base = blockIdx.x * 256u;
    const bool with_pos = h.pos_x != nullptr;
    uint32_t my = tid;
    bool active = base + tid < n;
    RayState st;
    if (active) {
        uint32_t i = base + tid;
        ray_setup(st, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.W[0], r.W[1], r.W[2]);
    }
    unsigned long long c_rays = 0, c_iters = 0, c_desc = 0, c_hits = 0;
    int32_t v[3] = {0, 0, 0};
    for (;;) {
        bool tame = __all(!active || ray_is_tame(st));
        bool hit = false, fin = false;
        if (active) {
            if (tame)
                fin = trace_segment<false>(w, st, seg, max_iters, v, hit);
            else
                fin = trace_segment<true>(w, st, seg, max_iters, v, hit);
        }
        if (fin) {
            float t;
            int4 rec;
            if (hit)
                ray_hit_record(w, st, v, brick_slot_at(w, v), t, rec);
            else
                ray_miss_record(st, t, rec);
            L.t[my] = t;
            L.rec[my] = rec;
            if (with_pos) {
#pragma unroll
                for (int a = 0; a < 3; a++) L.pos[a][my] = hit ? st.cur[a] : 0.0f;
            }
            if constexpr (COUNT) {
                c_rays++;
                c_iters += st.iters;
                c_desc += st.descents;
                c_hits += hit ? 1ull : 0ull;
            }
            active = false;
        }
        // pack the running rays into slots 0 .. total-1
        uint64_t am = __ballot(active);
        if (lane == 0) L.wave_cnt[wave] = (uint32_t)__popcll(am);
        __syncthreads();
        uint32_t total = 0, before = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            uint32_t c = L.wave_cnt[k];
            before += k < wave ? c : 0u;
            total += c;
        }
        if (total == 0) break;
        if (active) {
            uint32_t rank = before + __builtin_amdgcn_mbcnt_hi((uint32_t)(am >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)am, 0u));
#pragma unroll
            for (int a = 0; a < 3; a++) {
                L.o[a][rank] = st.o[a];
                L.d[a][rank] = st.d[a];
                L.inv[a][rank] = st.inv[a];
                L.side[a][rank] = st.side[a];
            }
            L.it[rank] = st.iters | st.descents << 16;
            L.slot[rank] = my;
        }
        __syncthreads();
        active = tid < total;
        if (active) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                st.o[a] = L.o[a][tid];
                st.d[a] = L.d[a][tid];
                st.inv[a] = L.inv[a][tid];
                st.side[a] = L.side[a][tid];
                st.ts[a] = ((st.d[a] < 0.0f ? 0.0f : 1.0f) - st.o[a]) * st.inv[a];
                st.dm[a] = st.d[a] < 0.0f ? 0 : -1;
                asm volatile("" : "+v"(st.dm[a]));
                st.W[a] = r.W[a];
            }
            float tmin = min_x86(min_x86(st.side[0], st.side[1]), st.side[2]) + 0.001f;
#pragma unroll
            for (int a = 0; a < 3; a++) st.cur[a] = fmaf(tmin, st.d[a], st.o[a]);
            uint32_t itd = L.it[tid];
            st.iters = itd & 0xFFFFu;
            st.descents = itd >> 16;
            my = L.slot[tid];
        }
    }
    // every ray of the workgroup has finished: write the results in ray order
    __syncthreads();
    if (base + tid < n) {
        uint32_t i = base + tid;
        h.t[i] = L.t[tid];
        h.rec[i] = L.rec[tid];
        if (with_pos) {
            h.pos_x[i] = L.pos[0][tid];
            h.pos_y[i] = L.pos[1][tid];
            h.pos_z[i] = L.pos[2][tid];
        }
    }
    if constexpr (COUNT) flush_counters(ctr, c_rays, c_iters, c_desc, c_hits);
}

// ---------------------------------------------------------------- MULTIPASS schedule

// A wave runs until its slowest ray, so rays with long-tailed iteration counts
// (diffuse bounces: median 21, p99 77 lookups on the 1024^3 terrain) keep most
// lanes idle.  The multi-pass schedule traces every ray for at most `pass_iters`
// further lookups per pass; rays still running are parked as 20-byte
// continuation records (ray index, lookups/descents, sideDist) and the next pass
// packs them densely into waves again.  The last pass runs to the cap.
// Continuations go to 8 regions (one per XCD: workgroup b % 8) through one
// atomic per workgroup, so claims never serialize on a single counter.
static const uint32_t kRegions = 8;

struct ContSet {
    uint32_t* idx;        // ray index
    uint32_t* it;         // lookups | descents << 16
    float* sx; float* sy; float* sz;  // sideDist
    uint32_t* cnt;        // kRegions counters, kQueueStride apart
    uint32_t region_cap;  // entries per region
};

__device__ __forceinline__ uint32_t cont_total(const ContSet& c) {
    uint32_t t = 0;
#pragma unroll
    for (uint32_t r = 0; r < kRegions; r++) t += min(c.cnt[r * kQueueStride], c.region_cap);
    return t;
}

template <bool FIRST, bool COUNT>
__global__ __launch_bounds__(256) void cast_pass_kernel(WorldView w, RaysIn r, HitsOut h, uint32_t cap,
                                                        uint32_t max_iters, ContSet in, ContSet out,
                                                        vrt_cast_counters* ctr) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    uint32_t counts[kRegions];
    uint32_t total;
    if (FIRST) {
        total = ray_count(r);
    } else {
        total = 0;
#pragma unroll
        for (uint32_t k = 0; k < kRegions; k++) {
            counts[k] = min(in.cnt[k * kQueueStride], in.region_cap);
            total += counts[k];
        }
    }
    const uint32_t nchunks = (total + 255) / 256;
    const uint32_t region = blockIdx.x % kRegions;
    const uint32_t wave = threadIdx.x / 64;
    unsigned long long c_rays = 0, c_iters = 0, c_desc = 0, c_hits = 0;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {  // uniform per workgroup
        uint32_t i = c * 256 + threadIdx.x;
        bool valid = i < total;
        RayState st;
        uint32_t idx = 0;
        if (valid) {
            if (FIRST) {
                idx = i;
                ray_setup(st, r.ox[idx], r.oy[idx], r.oz[idx], r.dx[idx], r.dy[idx], r.dz[idx], r.W[0], r.W[1], r.W[2]);
            } else {
                uint32_t e = i, k = 0;
                while (k + 1 < kRegions && e >= counts[k]) {
                    e -= counts[k];
                    k++;
                }
                e += k * in.region_cap;
                idx = in.idx[e];
                ray_setup(st, r.ox[idx], r.oy[idx], r.oz[idx], r.dx[idx], r.dy[idx], r.dz[idx], r.W[0], r.W[1], r.W[2]);
                uint32_t packed = in.it[e];
                st.iters = packed & 0xFFFFu;
                st.descents = packed >> 16;
                st.side[0] = in.sx[e];
                st.side[1] = in.sy[e];
                st.side[2] = in.sz[e];
                // the position after the parked step, recomputed exactly as the step computed it
                float tmin = min_x86(min_x86(st.side[0], st.side[1]), st.side[2]) + 0.001f;
#pragma unroll
                for (int a = 0; a < 3; a++) st.cur[a] = fmaf(tmin, st.d[a], st.o[a]);
            }
        }
        bool wave_tame = __all(!valid || ray_is_tame(st));
        int res = STEP_CONTINUE;
        int32_t v[3] = {0, 0, 0};
        uint32_t slot = 0;
        if (valid) {
            while (st.iters < cap) {
                res = wave_tame ? ray_step<false>(w, st, v, &slot) : ray_step<true>(w, st, v, &slot);
                if (res != STEP_CONTINUE) break;
            }
        }
        bool survivor = valid && res == STEP_CONTINUE && st.iters < max_iters;
        if (valid && !survivor) {
            float t;
            int4 rec;
            if (res == STEP_HIT)
                ray_hit_record(w, st, v, slot, t, rec);
            else
                ray_miss_record(st, t, rec);
            store_result(h, idx, st, t, rec);
            if constexpr (COUNT) {
                c_rays++;
                c_iters += st.iters;
                c_desc += st.descents;
                c_hits += (rec.w & 0xFF) != 0;
            }
        }
        // append survivors: one atomic per workgroup into this workgroup's region
        uint64_t sb = __ballot(survivor);
        if (__lane_id() == 0) s_wave[wave] = (uint32_t)__popcll(sb);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            s_base = tot ? atomicAdd(out.cnt + region * kQueueStride, tot) : 0u;
        }
        __syncthreads();
        if (survivor) {
            uint32_t off = s_base;
            for (uint32_t k = 0; k < wave; k++) off += s_wave[k];
            off += __builtin_amdgcn_mbcnt_hi((uint32_t)(sb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)sb, 0u));
            if (off < out.region_cap) {  // always true: a region holds every ray its workgroups can see
                uint32_t e = region * out.region_cap + off;
                out.idx[e] = idx;
                out.it[e] = (st.iters & 0xFFFFu) | (st.descents << 16);
                out.sx[e] = st.side[0];
                out.sy[e] = st.side[1];
                out.sz[e] = st.side[2];
            }
        }
        __syncthreads();  // s_wave / s_base are reused by the next chunk
    }
    if constexpr (COUNT) flush_counters(ctr, c_rays, c_iters, c_desc, c_hits);
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- device properties and queues

struct DeviceInfo {
    int cus = 256;
    uint32_t* queues = nullptr;  // ring of queue counters for persistent launches
    std::atomic<uint32_t> next{0};
};
static DeviceInfo g_dev[64];
static std::mutex g_dev_mu;

static const uint32_t kQueueRing = 256;  // concurrent persistent launches per device

int device_info(int dev, int* cus, uint32_t** queue) {
    if (dev < 0 || dev >= 64) return fail(VRT_ERR_NO_DEVICE, "device index out of range");
    DeviceInfo& di = g_dev[dev];
    {
        std::lock_guard<std::mutex> lk(g_dev_mu);
        if (!di.queues) {
            hipDeviceProp_t prop;
            VRT_HIP(hipGetDeviceProperties(&prop, dev));
            di.cus = prop.multiProcessorCount;
            VRT_HIP(hipMalloc(&di.queues, (size_t)kQueueRing * kQueues * kQueueStride * sizeof(uint32_t)));
        }
    }
    *cus = di.cus;
    if (queue) *queue = di.queues + (size_t)(di.next.fetch_add(1) % kQueueRing) * kQueues * kQueueStride;
    return VRT_OK;
}

// Library-owned continuation scratch, one per device, grown on demand.
struct PassScratch {
    void* buf = nullptr;
    size_t bytes = 0;
};
static PassScratch g_pass[64];
static std::mutex g_pass_mu;

// Continuation entries a region must hold for a cast of n rays over G workgroups.
static uint32_t region_capacity(uint32_t n, uint32_t grid) {
    uint32_t nchunks = (n + 255) / 256;
    (void)grid;
    return ((nchunks + kRegions - 1) / kRegions) * 256;
}

size_t pass_scratch_bytes(uint32_t n) {
    size_t per_set = (size_t)region_capacity(n, 0) * kRegions * 20 + kRegions * kQueueStride * 4;
    return 2 * per_set + 256;
}

static void carve(void* base, uint32_t region_cap, ContSet& a, ContSet& b) {
    size_t ents = (size_t)region_cap * kRegions;
    char* p = (char*)(((uintptr_t)base + 255) & ~(uintptr_t)255);
    ContSet* sets[2] = {&a, &b};
    for (ContSet* c : sets) {
        c->cnt = (uint32_t*)p; p += kRegions * kQueueStride * 4;
        c->idx = (uint32_t*)p; p += ents * 4;
        c->it = (uint32_t*)p; p += ents * 4;
        c->sx = (float*)p; p += ents * 4;
        c->sy = (float*)p; p += ents * 4;
        c->sz = (float*)p; p += ents * 4;
        c->region_cap = region_cap;
    }
}

// ---------------------------------------------------------------- launch

// Option checks that do not depend on the ray count, so a call is rejected or
// accepted the same way whatever n is (n = 0 included).
int check_cast_opts(const vrt_cast_opts& o) {
    uint32_t sched = o.schedule == VRT_SCHED_AUTO ? VRT_SCHED_GRID : o.schedule;
    if (sched > VRT_SCHED_LANES) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: unknown schedule");
    if (o.order > VRT_ORDER_REVERSE) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: unknown order");
    if (sched == VRT_SCHED_LANES && o.chunk > kLanesMaxRays)
        return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: lanes schedule takes 1..8 rays per lane");
    if (sched == VRT_SCHED_MULTIPASS && o.passes > 64)
        return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: at most 64 passes");
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
    if (sched == VRT_SCHED_LANES) {
        uint32_t K = o.chunk ? o.chunk : 4;
        if (r.n > 0xFFFFF000u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: too many rays");
        uint32_t blocks = (uint32_t)(((uint64_t)r.n + 64u * K - 1) / (64u * K));
        size_t lds = lanes_lds_bytes(K);
        if (o.counters)
            cast_lanes_kernel<true><<<blocks, 64, lds, s>>>(w, r, h, max_iters, K, o.counters);
        else
            cast_lanes_kernel<false><<<blocks, 64, lds, s>>>(w, r, h, max_iters, K, nullptr);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    if (sched == VRT_SCHED_REPACK) {
        uint32_t seg = o.pass_iters ? o.pass_iters : 16;
        if (r.n > 0xFFFFFF00u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: too many rays");
        uint32_t blocks = (r.n + 255) / 256;
        if (o.counters)
            cast_repack_kernel<true><<<blocks, 256, 0, s>>>(w, r, h, max_iters, seg, o.counters);
        else
            cast_repack_kernel<false><<<blocks, 256, 0, s>>>(w, r, h, max_iters, seg, nullptr);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    if (sched == VRT_SCHED_MULTIPASS) {
        uint32_t K = o.pass_iters ? o.pass_iters : 16;
        uint32_t P = o.passes ? o.passes : 5;
        int cus = 0;
        int rc = device_info(world->device, &cus, nullptr);
        if (rc != VRT_OK) return rc;
        size_t need = pass_scratch_bytes(r.n);
        void* base = o.scratch;
        if (base) {
            if (o.scratch_bytes < need) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: scratch smaller than vrt_cast_scratch_bytes(n)");
        } else {
            std::lock_guard<std::mutex> lk(g_pass_mu);
            PassScratch& ps = g_pass[world->device];
            if (ps.bytes < need) {
                VRT_HIP(hipStreamSynchronize(s));
                (void)hipFree(ps.buf);
                ps.buf = nullptr;
                ps.bytes = 0;
                VRT_HIP(hipMalloc(&ps.buf, need));
                ps.bytes = need;
            }
            base = ps.buf;
        }
        ContSet a, b;
        carve(base, region_capacity(r.n, 0), a, b);
        uint32_t grid = std::max<uint32_t>(kRegions, std::min<uint32_t>((uint32_t)cus * 8, (r.n + 255) / 256));
        grid = (grid + kRegions - 1) / kRegions * kRegions;
        ContSet* cur_in = &a;
        ContSet* cur_out = &b;
        for (uint32_t p = 0; p < P; p++) {
            uint32_t cap = p + 1 == P ? max_iters : std::min<uint64_t>((uint64_t)K * (p + 1), max_iters);
            fill_u32_kernel<<<(kRegions * kQueueStride + 255) / 256, 256, 0, s>>>(cur_out->cnt, kRegions * kQueueStride, 0u);
            if (p == 0) {
                if (o.counters)
                    cast_pass_kernel<true, true><<<grid, 256, 0, s>>>(w, r, h, cap, max_iters, *cur_in, *cur_out, o.counters);
                else
                    cast_pass_kernel<true, false><<<grid, 256, 0, s>>>(w, r, h, cap, max_iters, *cur_in, *cur_out, nullptr);
            } else {
                if (o.counters)
                    cast_pass_kernel<false, true><<<grid, 256, 0, s>>>(w, r, h, cap, max_iters, *cur_in, *cur_out, o.counters);
                else
                    cast_pass_kernel<false, false><<<grid, 256, 0, s>>>(w, r, h, cap, max_iters, *cur_in, *cur_out, nullptr);
            }
            VRT_HIP(hipGetLastError());
            std::swap(cur_in, cur_out);
            if (cap >= max_iters) break;
        }
        return VRT_OK;
    }
    uint32_t chunk = o.chunk ? o.chunk : 64;
    if (r.n > 0x7FFFFFFFu) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_ex: persistent schedule takes at most 2^31 - 1 rays");
    int cus = 0;
    uint32_t* queue = nullptr;
    int rc = device_info(world->device, &cus, &queue);
    if (rc != VRT_OK) return rc;
    uint32_t wpc = o.waves_per_cu ? std::min<uint32_t>(o.waves_per_cu, 32) : 32;
    // whole rounds of kQueues workgroups so every XCD part gets the same number of waves
    uint32_t blocks = std::max<uint32_t>(kQueues, ((uint32_t)cus * wpc / 4) / kQueues * kQueues);
    ChunkQueue cq;
    cq.ctr = queue;
    cq.nchunks = (r.n + chunk - 1) / chunk;
    cq.waves_per_q = blocks / kQueues * 4;
    uint32_t refill_min = o.refill ? std::min<uint32_t>(std::max<uint32_t>(o.refill, 1), 64) : 16;
    fill_u32_kernel<<<(kQueues * kQueueStride + 255) / 256, 256, 0, s>>>(queue, kQueues * kQueueStride, 0u);
    if (o.counters)
        cast_persistent_kernel<true><<<blocks, 256, 0, s>>>(w, r, h, max_iters, chunk, refill_min, cq, o.counters);
    else
        cast_persistent_kernel<false><<<blocks, 256, 0, s>>>(w, r, h, max_iters, chunk, refill_min, cq, nullptr);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int check_rays_hits(const vrt_rays* rays, const vrt_hits* hits, RaysIn& r, HitsOut& h) {
    if (!rays || !hits) return fail(VRT_ERR_INVALID_ARGUMENT, "null rays or hits");
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

int vrt_cast_ex(const vrt_world* world, const vrt_rays* rays, uint32_t max_iters, vrt_hits* hits,
                const vrt_cast_opts* opts, void* stream) {
    return guarded([&]() -> int {
        if (!world) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast: null world");
        if (max_iters == 0 || max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast: max_iters must be 1..65535");
        if (fault_injected("vrt_cast")) return fail(VRT_ERR_HIP, "vrt_cast: injected HIP fault");
        RaysIn r;
        HitsOut h;
        int rc = check_rays_hits(rays, hits, r, h);
        if (rc != VRT_OK) return rc;
        vrt_cast_opts o;
        std::memset(&o, 0, sizeof(o));
        if (opts) o = *opts;
        r.count = o.count;
        return launch_cast(world, r, h, max_iters, o, (hipStream_t)stream);
    });
}

size_t vrt_cast_scratch_bytes(uint32_t n) { return pass_scratch_bytes(n); }

int vrt_cast_one_f64(const vrt_world* world, const double origin[3], const double dir[3], uint32_t max_iters,
                     vrt_hit_f64* out) {
    return guarded([&]() -> int {
        if (!world || !origin || !dir || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_one_f64: null argument");
        if (world->log2_y > 13) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_cast_one_f64: world taller than 8192 voxels");
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
