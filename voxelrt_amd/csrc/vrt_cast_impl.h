// vrt_cast_impl.h - kernel-side ray/hit views shared by the cast, bounce,
// batched and differentiable translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "vrt_internal.h"

namespace vrt {

struct RaysIn {
    const float* __restrict__ ox;
    const float* __restrict__ oy;
    const float* __restrict__ oz;
    const float* __restrict__ dx;
    const float* __restrict__ dy;
    const float* __restrict__ dz;
    int32_t W[3];
    uint32_t n;                        // capacity
    const uint32_t* __restrict__ count;  // optional device ray count (<= n)
};

// Rays to trace: the device count when given, clamped to the capacity.
__device__ __forceinline__ uint32_t ray_count(const RaysIn& r) {
    return r.count ? min(*r.count, r.n) : r.n;
}

struct HitsOut {
    float* __restrict__ t;
    int4* __restrict__ rec;
    float* __restrict__ pos_x;  // all three null when positions are not requested
    float* __restrict__ pos_y;
    float* __restrict__ pos_z;
    float* __restrict__ uv_u;   // both null when texture coordinates are not requested
    float* __restrict__ uv_v;
    uint32_t* __restrict__ material_data;  // null when not requested
    const uint64_t* __restrict__ palette;  // the world's palette (material_data)
};

// ---------------------------------------------------------------- work queues
constexpr uint32_t kQueues = 8;
constexpr uint32_t kQueueStride = 64;  // u32 per counter (256 bytes)

struct ChunkQueue {
    uint32_t* ctr;          // kQueues counters, kQueueStride apart, zero at launch
    uint32_t nchunks;       // total chunks
    uint32_t waves_per_q;   // waves with a static first chunk in each part
};

__device__ __forceinline__ uint32_t part_begin(uint32_t q, uint32_t nchunks) {
    return (uint32_t)(((uint64_t)nchunks * q) / kQueues);
}

// Next chunk for this wave, or ~0u when every part is drained.  Wave-uniform.
__device__ __forceinline__ uint32_t claim_chunk(const ChunkQueue& cq, uint32_t home, uint32_t& probe) {
    while (probe < kQueues) {
        uint32_t q = (home + probe) % kQueues;
        uint32_t b = part_begin(q, cq.nchunks), e = part_begin(q + 1, cq.nchunks);
        uint32_t k = 0;
        if (__lane_id() == 0) k = atomicAdd(cq.ctr + q * kQueueStride, 1u);
        k = __shfl(k, 0, 64);
        uint32_t c = b + cq.waves_per_q + k;
        if (c < e) return c;
        probe++;
    }
    return ~0u;
}


// Validates the ABI structs and converts them to kernel views; with extras = false the
// optional uv and material_data outputs are rejected (entry points that do not fill them).
int check_rays_hits(const vrt_rays* rays, const vrt_hits* hits, RaysIn& r, HitsOut& h, bool extras = false);
// Launches a cast with the given options on stream s.
int launch_cast(const vrt_world* world, const RaysIn& r, const HitsOut& h, uint32_t max_iters,
                const vrt_cast_opts& o, hipStream_t s);
// Compute units of device `dev` and, when queue is not null, the queue counters of persistent
// launches on stream s of that device (kQueues counters, kQueueStride apart).
int device_info(int dev, int* cus, uint32_t** queue, hipStream_t s);

}  // namespace vrt
