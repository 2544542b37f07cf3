// rays.hip - primary-ray generation and diffuse bounce generation on the device.
//
// vrt_camera_rays replaces GetPrimaryRay (src/VoxelRT/CpuRenderer.cpp:226-233,
// Shaders/VoxelRender.comp:20-25) with the pinhole camera of vrt_math.h.
// vrt_bounce_generate replaces the bounce construction of RenderRow
// (CpuRenderer.cpp:389-392) / VoxelRender.comp:63-71, with the PCG sample of
// RandomGen.glsl:6-20, and compacts the bounce rays in primary order
// (count -> scatter, each scatter block summing the counts before it; beyond 8.4 M
// primaries count -> scan -> scatter; deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "vrt_cast_impl.h"
#include "vrt_math.h"

namespace vrt {

__global__ __launch_bounds__(256) void camera_rays_kernel(vrt_camera cam, float inv_w, float inv_h, uint32_t n,
                                                          float* __restrict__ ox, float* __restrict__ oy,
                                                          float* __restrict__ oz, float* __restrict__ dx,
                                                          float* __restrict__ dy, float* __restrict__ dz) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t px, py;
    tile_pixel(i, cam.width, cam.tile_w, cam.tile_h, px, py);
    float d[3];
    camera_ray(cam.forward, cam.right, cam.up, inv_w, inv_h, px, py, d);
    ox[i] = cam.eye[0];
    oy[i] = cam.eye[1];
    oz[i] = cam.eye[2];
    dx[i] = d[0];
    dy[i] = d[1];
    dz[i] = d[2];
}

static const uint32_t kBounceBlock = 256;

// Number of hits in each block of kBounceBlock primaries.
__global__ __launch_bounds__(256) void bounce_count_kernel(const int4* __restrict__ rec, uint32_t n,
                                                           uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t wave_counts[4];
    uint32_t i = blockIdx.x * kBounceBlock + threadIdx.x;
    bool hit = i < n && (rec[i].w & 0xFF) != 0;
    uint64_t b = __ballot(hit);
    if (__lane_id() == 0) wave_counts[threadIdx.x >> 6] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = wave_counts[0] + wave_counts[1] + wave_counts[2] + wave_counts[3];
}

// Exclusive scan of the block counts in one workgroup; writes the total to *count.
// Up to kScanLds counts are staged through LDS with coalesced loads and stores;
// thread t scans a run of consecutive counts, runs are combined with wave
// shuffles and across the 16 waves through LDS.
static const uint32_t kScanLds = 15872;  // 62 KiB of LDS

__device__ __forceinline__ uint32_t scan_runs(uint32_t sum, uint32_t* wave_total, uint32_t& total) {
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t incl = sum;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < 16; w++) {
        uint32_t c = wave_total[w];
        run += w < wave ? c : 0u;
        total += c;
    }
    return run;  // exclusive prefix of this thread's run
}

__global__ __launch_bounds__(1024) void bounce_scan_kernel(uint32_t* __restrict__ block_counts, uint32_t nblocks,
                                                           uint32_t* __restrict__ count) {
    __shared__ uint32_t wave_total[16];
    __shared__ uint32_t stage[kScanLds];
    const uint32_t t = threadIdx.x;
    const bool staged = nblocks <= kScanLds;
    uint32_t* buf = staged ? stage : block_counts;
    if (staged) {
        for (uint32_t k = t; k < nblocks; k += 1024) stage[k] = block_counts[k];
        __syncthreads();
    }
    uint32_t per = (nblocks + 1023) / 1024;
    uint32_t begin = min(t * per, nblocks), end = min(begin + per, nblocks);
    uint32_t sum = 0;
    for (uint32_t k = begin; k < end; k++) sum += buf[k];
    uint32_t total;
    uint32_t run = scan_runs(sum, wave_total, total);
    for (uint32_t k = begin; k < end; k++) {
        uint32_t c = buf[k];
        buf[k] = run;
        run += c;
    }
    if (staged) {
        __syncthreads();
        for (uint32_t k = t; k < nblocks; k += 1024) block_counts[k] = stage[k];
    }
    if (t == 0) *count = total;
}

__global__ __launch_bounds__(256) void bounce_write_kernel(const int4* __restrict__ rec, const float* __restrict__ px,
                                                           const float* __restrict__ py, const float* __restrict__ pz,
                                                           uint32_t n, uint32_t width, uint32_t seed,
                                                           const uint32_t* __restrict__ block_offsets,
                                                           vrt_ray_buffers out, uint32_t* __restrict__ src) {
    __shared__ uint32_t wave_counts[4];
    uint32_t i = blockIdx.x * kBounceBlock + threadIdx.x;
    int4 r = i < n ? rec[i] : make_int4(0, 0, 0, 0);
    bool hit = i < n && (r.w & 0xFF) != 0;
    uint64_t b = __ballot(hit);
    uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) wave_counts[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    if (!hit) return;
    uint32_t base = block_offsets[blockIdx.x];
    for (uint32_t w = 0; w < wave; w++) base += wave_counts[w];
    uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
    uint32_t k = base + rank;
    float p[3] = {px[i], py[i], pz[i]};
    float o[3], d[3];
    bounce_ray((uint32_t)r.w, p, i, width, seed, o, d);
    out.ox[k] = o[0];
    out.oy[k] = o[1];
    out.oz[k] = o[2];
    out.dx[k] = d[0];
    out.dy[k] = d[1];
    out.dz[k] = d[2];
    src[k] = i;
}

// Two-launch generation for up to kGenMaxBlocks blocks of kGenBlock primaries (8.4 M):
// the count launch writes one hit count per block; each write block then sums the counts
// of the blocks before it (at most two loads per thread) instead of waiting for a scan
// launch, and block 0 also writes the total.  Thread t of block b handles primaries
// b * kGenBlock + j * 1024 + t, j = 0..3, so the output order is the primary order.
static const uint32_t kGenRows = 4;
static const uint32_t kGenBlock = 1024 * kGenRows;
static const uint32_t kGenMaxBlocks = 2048;
// Largest primary count: the block counts (n + kGenBlock - 1) / kGenBlock and
// (n + kBounceBlock - 1) / kBounceBlock are computed in 32 bits.
static const uint32_t kMaxBouncePrimaries = 0xFFFFF000u;

__global__ __launch_bounds__(1024) void bounce_count_rows_kernel(const int4* __restrict__ rec, uint32_t n,
                                                                 uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t wave_counts[16];
    const uint32_t t = threadIdx.x;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < kGenRows; j++) {
        uint32_t i = blockIdx.x * kGenBlock + j * 1024 + t;
        bool hit = i < n && (rec[i].w & 0xFF) != 0;
        c += (uint32_t)__popcll(__ballot(hit));
    }
    if (__lane_id() == 0) wave_counts[t >> 6] = c;
    __syncthreads();
    if (t < 64) {
        uint32_t v = t < 16 ? wave_counts[t] : 0u;
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (t == 0) block_counts[blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(1024) void bounce_write_rows_kernel(const int4* __restrict__ rec,
                                                                 const float* __restrict__ px,
                                                                 const float* __restrict__ py,
                                                                 const float* __restrict__ pz, uint32_t n,
                                                                 udiv32 width, uint32_t seed,
                                                                 const uint32_t* __restrict__ block_counts,
                                                                 uint32_t nblocks, vrt_ray_buffers out,
                                                                 uint32_t* __restrict__ src,
                                                                 uint32_t* __restrict__ count) {
    __shared__ uint32_t wave_sums[16];
    __shared__ uint32_t row_base[kGenRows * 16];  // per (row, wave): hits before it in this block
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t >> 6, b = blockIdx.x;
    // hits of the blocks before this one (block 0: of all blocks, the total)
    const uint32_t lim = b == 0 ? nblocks : b;
    uint32_t before = 0;
    for (uint32_t k = t; k < lim; k += 1024) before += block_counts[k];  // kGenMaxBlocks / 1024 loads at most
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) before += __shfl_xor(before, off, 64);
    if (lane == 0) wave_sums[wave] = before;
    int4 r[kGenRows];
    uint64_t ball[kGenRows];
#pragma unroll
    for (uint32_t j = 0; j < kGenRows; j++) {
        uint32_t i = b * kGenBlock + j * 1024 + t;
        r[j] = i < n ? rec[i] : make_int4(0, 0, 0, 0);
        ball[j] = __ballot(i < n && (r[j].w & 0xFF) != 0);
        if (lane == 0) row_base[j * 16 + wave] = (uint32_t)__popcll(ball[j]);
    }
    __syncthreads();
    if (t < 64) {
        // exclusive scan of the 64 (row, wave) counts in row-major order = primary order
        uint32_t c = t < kGenRows * 16 ? row_base[t] : 0u, incl = c;
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(incl, off, 64);
            if (t >= off) incl += y;
        }
        uint32_t base = 0;
#pragma unroll
        for (uint32_t w = 0; w < 16; w++) base += wave_sums[w];
        if (b == 0) {
            if (t == 0) *count = base;  // block 0 summed every block: the total
            base = 0;
        }
        if (t < kGenRows * 16) row_base[t] = base + incl - c;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kGenRows; j++) {
        uint32_t i = b * kGenBlock + j * 1024 + t;
        if (!(i < n && (r[j].w & 0xFF) != 0)) continue;
        uint32_t k = row_base[j * 16 + wave] +
                     __builtin_amdgcn_mbcnt_hi((uint32_t)(ball[j] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ball[j], 0u));
        float p[3] = {px[i], py[i], pz[i]};
        float o[3], d[3], u1, u2;
        bounce_samples(i, width, seed, u1, u2);
        bounce_ray_samples((uint32_t)r[j].w, p, u1, u2, o, d);
        out.ox[k] = o[0];
        out.oy[k] = o[1];
        out.oz[k] = o[2];
        out.dx[k] = d[0];
        out.dy[k] = d[1];
        out.dz[k] = d[2];
        src[k] = i;
    }
}

// ---------------------------------------------------------------- direction order

// Group key of hit i's bounce ray: bit a set when component a of N + s is >= 0, read off the
// terms' signs (d = normalize(N + s), cosine_dir): on the normal's axis the normal's sign,
// elsewhere the sign of s's component from the two PCG samples cosine_dir draws,
// s = (sin(2 pi u2) sy, 2 u1 - 1, cos(2 pi u2) sy) with sy >= 0.  It is d's octant except
// for rays with a component at a sign boundary; both passes compute it with this function,
// so the placement never depends on rounding.  8: no hit, no bounce ray.  u1, u2: the samples.
__device__ __forceinline__ uint32_t bounce_key(uint32_t info, uint32_t i, const udiv32& width, uint32_t seed,
                                               float& u1, float& u2) {
    if ((info & 0xFFu) == 0) return 8u;
    bounce_samples(i, width, seed, u1, u2);
    const uint32_t cx = (info >> 8) & 3u, cy = (info >> 10) & 3u, cz = (info >> 12) & 3u;  // 1: +, 3: -
    const bool bx = cx ? cx == 1u : u2 < 0.5f;
    const bool by = cy ? cy == 1u : u1 >= 0.5f;
    const bool bz = cz ? cz == 1u : (u2 < 0.25f || u2 >= 0.75f);
    return (bx ? 1u : 0u) | (by ? 2u : 0u) | (bz ? 4u : 0u);
}

// The lanes of the wave whose key equals this lane's, from one ballot per key bit (keys are
// 0..8; lanes with key 8 are left out of every mask); pops[o]: the count of key o (o < 8),
// the same for every lane.
__device__ __forceinline__ uint64_t same_key_lanes(uint32_t key, uint32_t pops[8]) {
    const uint64_t valid = __ballot(key < 8u);
    const uint64_t b0 = __ballot(key & 1u), b1 = __ballot(key & 2u), b2 = __ballot(key & 4u);
#pragma unroll
    for (uint32_t o = 0; o < 8; o++)
        pops[o] = (uint32_t)__popcll(valid & (o & 1u ? b0 : ~b0) & (o & 2u ? b1 : ~b1) & (o & 4u ? b2 : ~b2));
    return valid & (key & 1u ? b0 : ~b0) & (key & 2u ? b1 : ~b1) & (key & 4u ? b2 : ~b2);
}

// Per block of kGenBlock primaries, the hits of each key: counts[key * nblocks + block].
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void bounce_count_keys_kernel(const int4* __restrict__ rec, uint32_t n,
                                                                 udiv32 width, uint32_t seed, uint32_t nblocks,
                                                                 uint32_t* __restrict__ counts) {
    __shared__ uint32_t wave_counts[16][8];
    const uint32_t t = threadIdx.x, b = blockIdx.x;
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (uint32_t j = 0; j < kGenRows; j++) {
        const uint32_t i = b * kGenBlock + j * 1024 + t;
        float u1, u2;
        const uint32_t key = i < n ? bounce_key((uint32_t)rec[i].w, i, width, seed, u1, u2) : 8u;
        uint32_t pops[8];
        (void)same_key_lanes(key, pops);
#pragma unroll
        for (uint32_t o = 0; o < 8; o++) c[o] += pops[o];
    }
    if (__lane_id() == 0)
#pragma unroll
        for (uint32_t o = 0; o < 8; o++) wave_counts[t >> 6][o] = c[o];
    __syncthreads();
    if (t < 8) {
        uint32_t v = 0;
        for (uint32_t w = 0; w < 16; w++) v += wave_counts[w][t];
        counts[t * nblocks + b] = v;
    }
}

// Writes block b's bounce rays: key o's rays go after those of keys < o (all blocks) and of key
// o in the blocks before b, in primary order within the block.  Block 0 writes the total.
// Rows of 1024 primaries are staged in LDS two at a time, sorted by key, and copied out in 8
// contiguous runs, one per key, so the global stores coalesce (written straight from the
// lanes, a wave's stores would split into a run per key).
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void bounce_write_keys_kernel(const int4* __restrict__ rec,
                                                                 const float* __restrict__ px,
                                                                 const float* __restrict__ py,
                                                                 const float* __restrict__ pz, uint32_t n,
                                                                 udiv32 width, uint32_t seed,
                                                                 const uint32_t* __restrict__ counts,
                                                                 uint32_t nblocks, vrt_ray_buffers out,
                                                                 uint32_t* __restrict__ src,
                                                                 uint32_t* __restrict__ count) {
    __shared__ uint32_t key_before[8], key_total[8], key_base[8];
    __shared__ uint32_t local[8 * kGenRows * 16];  // per (key, row, wave): rays, then their exclusive prefix
    // rows are staged kStageRows at a time (a staging round)
    constexpr uint32_t kStageRows = 2, kRounds = kGenRows / kStageRows, kStage = kStageRows * 1024;
    __shared__ uint32_t round_start[kRounds][9];   // per round: where key o's rays start in its sorted list
    __shared__ uint32_t key_block[8];              // key o's rays in this block
    __shared__ float stage[6][kStage];             // one round's rays sorted by key: o, d
    __shared__ uint32_t stage_src[kStage];         // and their primary indices
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t >> 6, b = blockIdx.x;
    // key `wave` (waves 0..7): its rays in the blocks before b and in all blocks
    if (wave < 8) {
        const uint32_t* c = counts + wave * nblocks;
        uint32_t bef = 0, tot = 0;
        for (uint32_t k = lane; k < nblocks; k += 64) {  // kGenMaxBlocks / 64 loads at most
            const uint32_t v = c[k];
            tot += v;
            bef += k < b ? v : 0u;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            bef += __shfl_xor(bef, off, 64);
            tot += __shfl_xor(tot, off, 64);
        }
        if (lane == 0) {
            key_before[wave] = bef;
            key_total[wave] = tot;
        }
    }
    uint32_t info[kGenRows], key[kGenRows];
    float u1[kGenRows], u2[kGenRows];
    uint64_t same[kGenRows];  // the lanes of this wave and row with this lane's key
#pragma unroll
    for (uint32_t j = 0; j < kGenRows; j++) {
        const uint32_t i = b * kGenBlock + j * 1024 + t;
        info[j] = i < n ? (uint32_t)rec[i].w : 0u;
        key[j] = i < n ? bounce_key(info[j], i, width, seed, u1[j], u2[j]) : 8u;
        uint32_t pops[8];
        same[j] = same_key_lanes(key[j], pops);
        if (lane == 0)
#pragma unroll
            for (uint32_t o = 0; o < 8; o++) local[(o * kGenRows + j) * 16 + wave] = pops[o];
    }
    __syncthreads();
    if (t < 8) {
        uint32_t v = key_before[t];
        for (uint32_t o = 0; o < t; o++) v += key_total[o];
        key_base[t] = v;
    }
    if (b == 0 && t == 0) {
        uint32_t v = 0;
        for (uint32_t o = 0; o < 8; o++) v += key_total[o];
        *count = v;
    }
    __syncthreads();
    // exclusive scan of the 512 local counts in (key, row, wave) order within each key: wave o
    // scans key o's 64 entries (row-major, so primary order); its last lane holds the block's
    // count of key o
    if (t < 8 * kGenRows * 16) {
        const uint32_t c = local[t];
        uint32_t incl = c;
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        local[t] = incl - c;
        if (lane == 63) key_block[wave] = incl;
    }
    __syncthreads();
    // per staging round, where each key's rays start in the round's sorted list: key o has
    // (prefix at the round's end) - (prefix at its start) rays in the round
    if (t < kRounds) {
        uint32_t v = 0;
        for (uint32_t o = 0; o < 8; o++) {
            round_start[t][o] = v;
            const uint32_t first = local[(o * kGenRows + t * kStageRows) * 16];
            const uint32_t end = t + 1 < kRounds ? local[(o * kGenRows + (t + 1) * kStageRows) * 16] : key_block[o];
            v += end - first;
        }
        round_start[t][8] = v;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kRounds; q++) {
#pragma unroll
        for (uint32_t r = 0; r < kStageRows; r++) {
            const uint32_t j = q * kStageRows + r;
            const uint32_t i = b * kGenBlock + j * 1024 + t;
            if (key[j] < 8u) {
                const uint32_t kj = key[j];
                // this ray's place in the round's sorted list: the round's keys before kj, this
                // key's rays of the round's earlier rows and waves, then the lanes before this one
                const uint32_t slot = round_start[q][kj] + local[(kj * kGenRows + j) * 16 + wave] -
                                      local[(kj * kGenRows + q * kStageRows) * 16] +
                                      __builtin_amdgcn_mbcnt_hi((uint32_t)(same[j] >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)same[j], 0u));
                const float p[3] = {px[i], py[i], pz[i]};
                float o[3], d[3];
                bounce_ray_samples(info[j], p, u1[j], u2[j], o, d);
                stage[0][slot] = o[0];
                stage[1][slot] = o[1];
                stage[2][slot] = o[2];
                stage[3][slot] = d[0];
                stage[4][slot] = d[1];
                stage[5][slot] = d[2];
                stage_src[slot] = i;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = t; u < kStage; u += 1024) {
            if (u < round_start[q][8]) {
                // the key of sorted slot u, and where its run of this round goes
                uint32_t ku = 0;
#pragma unroll
                for (uint32_t o = 1; o < 8; o++) ku += u >= round_start[q][o] ? 1u : 0u;
                const uint32_t k = key_base[ku] + local[(ku * kGenRows + q * kStageRows) * 16] + (u - round_start[q][ku]);
                out.ox[k] = stage[0][u];
                out.oy[k] = stage[1][u];
                out.oz[k] = stage[2][u];
                out.dx[k] = stage[3][u];
                out.dy[k] = stage[4][u];
                out.dz[k] = stage[5][u];
                src[k] = stage_src[u];
            }
        }
        __syncthreads();
    }
}

}  // namespace vrt

using namespace vrt;

// ---------------------------------------------------------------- no-allocation scratch

// Scratch of vrt_bounce_generate_noalloc: the hit count of each block of kGenBlock primaries
// of the two-launch generation.  Every GPU the library runs on holds its own copy of this
// array, at its own address, until the process ends.  In host code the array's name is not a
// device address: kernels get the address hipGetSymbolAddress returns for the device.
static __device__ __align__(256) uint32_t scratch[kGenMaxBlocks];

// Primaries the no-allocation entry point takes: kGenMaxBlocks blocks of kGenBlock (8 388 608,
// e.g. a 3840x2160 frame).
static const uint32_t kNoAllocMaxPrimaries = kGenMaxBlocks * kGenBlock;

static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[64];  // per device: the device address of `scratch`

// The device address of `scratch` on the current device, looked up once per device.
static int scratch_of_current_device(uint32_t** out) {
    int dev = 0;
    VRT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(VRT_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return VRT_OK;
}

// ---------------------------------------------------------------- bounce generation

// Checks the arguments and launches the generation with `blocks` as the per-block counts
// (at least vrt_bounce_scratch_bytes(n) bytes, or kGenMaxBlocks counts when n <= max_n is
// the two-launch range).  Every block count the launches read is written by them first.
static int bounce_generate(const vrt_rays* primary, const vrt_hits* hits, uint32_t width, uint32_t seed,
                           vrt_ray_buffers* out, uint32_t* src, uint32_t* count, uint32_t* blocks, uint32_t max_n,
                           hipStream_t s, const char* entry, uint32_t order = VRT_BOUNCE_PRIMARY_ORDER) {
    const std::string name(entry);
    if (!primary || !hits || !out || !count) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": null argument");
    if (width == 0) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": width must be > 0");
    uint32_t n = primary->n;
    // block counts below are 32-bit: (n + 4095) / 4096 and (n + 255) / 256 must not wrap
    if (n > max_n)
        return fail(VRT_ERR_INVALID_ARGUMENT,
                    name + (max_n == kMaxBouncePrimaries ? ": at most 2^32 - 4096 primary rays"
                                                         : ": at most 8388608 primary rays (the size of its scratch)"));
    if (n && (!src || !blocks || !hits->rec)) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": null argument");
    if (n && !hits->pos) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": hits need positions");
    if (n && (!out->ox || !out->oy || !out->oz || !out->dx || !out->dy || !out->dz))
        return fail(VRT_ERR_INVALID_ARGUMENT, name + ": null ray buffer");
    if (order > VRT_BOUNCE_DIRECTION_ORDER) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": unknown order");
    if (order == VRT_BOUNCE_DIRECTION_ORDER && n > kGenMaxBlocks * kGenBlock)
        return fail(VRT_ERR_INVALID_ARGUMENT, name + ": direction order takes at most 8388608 primary rays");
    if (n == 0) {
        // writes *count = 0 and reads no block count
        hipLaunchKernelGGL(bounce_scan_kernel, dim3(1), dim3(1024), 0, s, blocks, 0u, count);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    const int4* rec = (const int4*)hits->rec;
    const uint32_t nrows = (n + kGenBlock - 1) / kGenBlock;
    if (order == VRT_BOUNCE_DIRECTION_ORDER) {  // nrows <= kGenMaxBlocks; blocks holds 8 * nrows counts
        const udiv32 wd = make_udiv32(width);
        bounce_count_keys_kernel<<<nrows, 1024, 0, s>>>(rec, n, wd, seed, nrows, blocks);
        bounce_write_keys_kernel<<<nrows, 1024, 0, s>>>(rec, hits->pos, hits->pos + n, hits->pos + 2 * (size_t)n, n,
                                                        wd, seed, blocks, nrows, *out, src, count);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    if (nrows <= kGenMaxBlocks) {
        bounce_count_rows_kernel<<<nrows, 1024, 0, s>>>(rec, n, blocks);
        bounce_write_rows_kernel<<<nrows, 1024, 0, s>>>(rec, hits->pos, hits->pos + n, hits->pos + 2 * (size_t)n, n,
                                                        make_udiv32(width), seed, blocks, nrows, *out, src, count);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    uint32_t nb = (n + kBounceBlock - 1) / kBounceBlock;
    bounce_count_kernel<<<nb, 256, 0, s>>>(rec, n, blocks);
    bounce_scan_kernel<<<1, 1024, 0, s>>>(blocks, nb, count);
    bounce_write_kernel<<<nb, 256, 0, s>>>(rec, hits->pos, hits->pos + n, hits->pos + 2 * (size_t)n, n, width, seed,
                                           blocks, *out, src);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

extern "C" {

int vrt_camera_rays(const vrt_camera* cam, vrt_ray_buffers* out, void* stream) {
    return guarded([&]() -> int {
        if (!cam || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_rays: null argument");
        if (!cam->tile_w || !cam->tile_h || !cam->width || !cam->height || cam->width % cam->tile_w ||
            cam->height % cam->tile_h)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_rays: size must be a positive multiple of the tile");
        if (!out->ox || !out->oy || !out->oz || !out->dx || !out->dy || !out->dz)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_rays: null ray buffer");
        uint64_t n64 = (uint64_t)cam->width * cam->height;
        // 256-thread blocks: the grid's lanes (blocks * 256) must fit in 32 bits
        if (n64 > 0xFFFFFF00ull) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_rays: too many pixels");
        uint32_t n = (uint32_t)n64;
        camera_rays_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(
            *cam, 1.0f / (float)cam->width, 1.0f / (float)cam->height, n, out->ox, out->oy, out->oz, out->dx, out->dy,
            out->dz);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}

size_t vrt_bounce_scratch_bytes(uint32_t n) {
    // one count per block of kBounceBlock primaries (+1), or 8 per block of kGenBlock (direction order)
    const uint64_t per_block = ((uint64_t)n + kBounceBlock - 1) / kBounceBlock + 1;
    const uint64_t per_key = 8 * (((uint64_t)n + kGenBlock - 1) / kGenBlock);
    return (size_t)std::max(per_block, per_key) * sizeof(uint32_t);
}

int vrt_bounce_generate(const vrt_rays* primary, const vrt_hits* hits, uint32_t width, uint32_t seed,
                        vrt_ray_buffers* out, uint32_t* src, uint32_t* count, void* workspace, void* stream) {
    return guarded([&]() -> int {
        if (primary && primary->n && !workspace)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_bounce_generate: null argument");
        return bounce_generate(primary, hits, width, seed, out, src, count, (uint32_t*)workspace, kMaxBouncePrimaries,
                               (hipStream_t)stream, "vrt_bounce_generate");
    });
}

int vrt_bounce_generate_ex(const vrt_rays* primary, const vrt_hits* hits, uint32_t width, uint32_t seed,
                           uint32_t order, vrt_ray_buffers* out, uint32_t* src, uint32_t* count, void* workspace,
                           void* stream) {
    return guarded([&]() -> int {
        if (primary && primary->n && !workspace)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_bounce_generate_ex: null argument");
        return bounce_generate(primary, hits, width, seed, out, src, count, (uint32_t*)workspace, kMaxBouncePrimaries,
                               (hipStream_t)stream, "vrt_bounce_generate_ex", order);
    });
}

int vrt_bounce_generate_noalloc(const vrt_rays* primary, const vrt_hits* hits, uint32_t width, uint32_t seed,
                                vrt_ray_buffers* out, uint32_t* src, uint32_t* count, void* stream) {
    return guarded([&]() -> int {
        uint32_t* blocks = nullptr;
        int rc = scratch_of_current_device(&blocks);
        if (rc != VRT_OK) return rc;
        return bounce_generate(primary, hits, width, seed, out, src, count, blocks, kNoAllocMaxPrimaries,
                               (hipStream_t)stream, "vrt_bounce_generate_noalloc");
    });
}

}  // extern "C"
