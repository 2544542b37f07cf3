// world.hip - device world lifecycle, slot arena, brick sync and occupancy build.
//
// Replaces GpuVoxelStorage (src/VoxelRT/GpuRenderer.cpp:17-167), its slot
// allocator (src/VoxelRT/BrickSlotAllocator.{h,cpp}) and the occupancy build
// shader (src/VoxelRT/Shaders/UpdateOccupancy.comp:9-35).  Layout in HBM:
// structure-of-arrays sector masks / base slots / per-brick occupancy /
// voxel bytes / palette, brick b of a sector at base + popcount(mask below b)
// (BrickSlotAllocator.h:30-41, VoxelMap.glsl:45-64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <bit>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "vrt_internal.h"

namespace vrt {

// ---------------------------------------------------------------- errors

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
int fail(int status, const std::string& msg) {
    g_error = msg;
    return status;
}
int hip_fail(hipError_t e, const char* what) {
    g_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    return VRT_ERR_HIP;
}
bool fault_injected(const char* entry) {
    const char* v = std::getenv("VRT_FAULT_INJECT");
    return v && std::strcmp(v, entry) == 0;
}

// ---------------------------------------------------------------- slot arena

// On failure the arena is left with the old run released; callers work on a
// copy and discard it when any sector does not fit.
bool SlotArena::realloc(uint32_t& base, uint32_t cur, uint32_t want) {
    if (want == 0) {
        if (cur != 0) release(base, cur);
        base = 0;
        return true;
    }
    if (want == cur) return true;
    if (want < cur) {
        release(base + want, cur - want);
        if (want == 0) base = 0;
        return true;
    }
    if (cur != 0) {
        // grow in place when the range right after the run is free and long enough
        auto it = free_.find(base + cur);
        if (it != free_.end() && it->second >= want - cur) {
            uint32_t extra = want - cur;
            uint32_t start = it->first, len = it->second;
            free_.erase(it);
            if (len > extra) free_[start + extra] = len - extra;
            allocated_ += extra;
            return true;
        }
        release(base, cur);  // every brick of the sector is rewritten, so the run may move
    }
    // best fit: the smallest free range that holds the run, lowest address on ties
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
        if (it->second >= want && (best == free_.end() || it->second < best->second)) best = it;
    if (best == free_.end()) return false;
    uint32_t start = best->first, len = best->second;
    free_.erase(best);
    if (len > want) free_[start + want] = len - want;
    allocated_ += want;
    base = start;
    return true;
}

void SlotArena::release(uint32_t base, uint32_t size) {
    if (size == 0) return;
    allocated_ -= size;
    auto it = free_.emplace(base, size).first;
    // merge with the following range
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_.erase(next);
    }
    // merge with the preceding range
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_.erase(it);
        }
    }
}

uint32_t SlotArena::high_water() const {
    // the free range that ends at capacity marks the end of the used prefix
    if (free_.empty()) return capacity_;
    auto last = std::prev(free_.end());
    if (last->first + last->second == capacity_) return last->first;
    return capacity_;
}

// ---------------------------------------------------------------- sync kernels

// Occupancy mask of cell `cell` (cx | cz << 1 | cy << 2) of a 512-byte brick: bit vx + 4 vz + 16 vy,
// as in CpuRenderer.cpp:63-83 (UpdateOccupancy.comp:9-35 writes the same 64-bit mask as two
// 4x2x4 halves).
__device__ __forceinline__ uint64_t cell_occupancy(const uint8_t* brick, uint32_t cell) {
    uint32_t cx = (cell & 1) * 4, cz = ((cell >> 1) & 1) * 4, cy = (cell >> 2) * 4;
    uint64_t m = 0;
#pragma unroll
    for (uint32_t vy = 0; vy < 4; vy++)
#pragma unroll
        for (uint32_t vz = 0; vz < 4; vz++) {
            uint32_t row = *(const uint32_t*)(brick + cx + (cz + vz) * 8 + (cy + vy) * 64);
#pragma unroll
            for (uint32_t vx = 0; vx < 4; vx++)
                if ((row >> (8 * vx)) & 0xFFu) m |= 1ull << (vx + 4 * vz + 16 * vy);
        }
    return m;
}

// Uploaded bricks: thread (k, part), 8 per brick, copies bytes [64 part, 64 part + 64) of staged
// brick k to its slot and builds the occupancy mask of cell `part` from the staged bytes.
__global__ __launch_bounds__(256) void sync_upload_kernel(uint8_t* __restrict__ vox, uint64_t* __restrict__ occ,
                                                          const uint8_t* __restrict__ bricks,
                                                          const uint32_t* __restrict__ slots, uint32_t n) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k = gid >> 3, part = gid & 7;
    if (k >= n) return;
    uint32_t slot = slots[k];
    const uint8_t* src = bricks + (size_t)k * 512;
    const uint4* s4 = (const uint4*)(src + part * 64);
    uint4* d4 = (uint4*)(vox + (size_t)slot * 512 + part * 64);
#pragma unroll
    for (int j = 0; j < 4; j++) d4[j] = s4[j];
    occ[(size_t)slot * 8 + part] = cell_occupancy(src, part);
}

// Bricks that keep their contents but change slot (their sector's allocation changed), in two
// launches so that no move reads a slot another move has written: thread (k, part), 9 per
// brick, copies 64 bytes of slot src[k] (parts 0-7: voxels, part 8: the 8 occupancy masks)
// into the move area, then back out into slot dst[k].
__global__ __launch_bounds__(256) void sync_gather_kernel(const uint8_t* __restrict__ vox,
                                                          const uint64_t* __restrict__ occ,
                                                          const uint32_t* __restrict__ src, uint32_t n,
                                                          uint8_t* __restrict__ area) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k = gid / 9, part = gid % 9;
    if (k >= n) return;
    size_t slot = src[k];
    const uint4* s4 = part < 8 ? (const uint4*)(vox + slot * 512 + part * 64) : (const uint4*)(occ + slot * 8);
    uint4* d4 = (uint4*)(area + (size_t)k * 576 + part * 64);
#pragma unroll
    for (int j = 0; j < 4; j++) d4[j] = s4[j];
}

__global__ __launch_bounds__(256) void sync_place_kernel(uint8_t* __restrict__ vox, uint64_t* __restrict__ occ,
                                                         const uint32_t* __restrict__ dst, uint32_t n,
                                                         const uint8_t* __restrict__ area) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k = gid / 9, part = gid % 9;
    if (k >= n) return;
    size_t slot = dst[k];
    const uint4* s4 = (const uint4*)(area + (size_t)k * 576 + part * 64);
    uint4* d4 = part < 8 ? (uint4*)(vox + slot * 512 + part * 64) : (uint4*)(occ + slot * 8);
#pragma unroll
    for (int j = 0; j < 4; j++) d4[j] = s4[j];
}

// Records of the synced sectors (indices unique) and the sector-of-sectors words they changed.
__global__ __launch_bounds__(256) void sync_records_kernel(uint4* __restrict__ sec, const uint4* __restrict__ rec,
                                                           const uint32_t* __restrict__ rec_idx, uint32_t nrec,
                                                           uint64_t* __restrict__ groups,
                                                           const uint64_t* __restrict__ gval,
                                                           const uint32_t* __restrict__ gidx, uint32_t ngrp) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrec)
        sec[rec_idx[i]] = rec[i];
    else if (i - nrec < ngrp)
        groups[gidx[i - nrec]] = gval[i - nrec];
}

}  // namespace vrt

using namespace vrt;

// ---------------------------------------------------------------- helpers

static int grow_slots(vrt_world* w, uint32_t need, hipStream_t s) {
    if (need <= w->slot_cap) return VRT_OK;
    uint32_t cap = std::max<uint32_t>(need, 1024);
    cap = std::bit_ceil(cap);
    cap = std::min<uint32_t>(std::max(cap, need), w->arena.capacity());
    uint64_t* occ = nullptr;
    uint8_t* vox = nullptr;
    VRT_HIP(hipMalloc(&occ, (size_t)cap * 8 * sizeof(uint64_t)));
    hipError_t e = hipMalloc(&vox, (size_t)cap * 512);
    if (e != hipSuccess) {
        (void)hipFree(occ);
        return hip_fail(e, "hipMalloc(voxels)");
    }
    if (w->slot_cap) {
        if ((e = hipMemcpyAsync(occ, w->d_occ, (size_t)w->slot_cap * 64, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(vox, w->d_vox, (size_t)w->slot_cap * 512, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess) {
            (void)hipFree(occ);  // the world keeps its old buffers
            (void)hipFree(vox);
            return hip_fail(e, "growing the slot buffers");
        }
    }
    (void)hipFree(w->d_occ);
    (void)hipFree(w->d_vox);
    w->d_occ = occ;
    w->d_vox = vox;
    w->slot_cap = cap;
    w->generation++;
    return VRT_OK;
}

static void free_world(vrt_world* w) {
    (void)hipFree(w->d_sec);
    (void)hipFree(w->d_occ);
    (void)hipFree(w->d_vox);
    (void)hipFree(w->d_palette);
    (void)hipFree(w->d_groups);
    (void)hipFree(w->d_lut);
    (void)hipFree(w->d_stage);
    (void)hipFree(w->d_moves);
    if (w->h_stage) (void)hipHostFree(w->h_stage);
    for (hipEvent_t e : w->stage_ev)
        if (e) (void)hipEventDestroy(e);
    if (w->moves_ev) (void)hipEventDestroy(w->moves_ev);
    delete w;
}

int vrt::check_world(const vrt_world* w, const char* entry) {
    if (w && w->failed)
        return fail(VRT_ERR_STATE, std::string(entry) +
                                       ": a vrt_world_sync failed after its device work began, so the world's "
                                       "device tables are inconsistent; destroy the world");
    // the world's buffers, and the state kept per (device, stream), live on the device that
    // was current at vrt_world_create: a call with another one current would allocate on it
    int cur = -1;
    if (w && hipGetDevice(&cur) == hipSuccess && cur != w->device)
        return fail(VRT_ERR_STATE, std::string(entry) + ": device " + std::to_string(cur) +
                                       " is current, the world lives on device " + std::to_string(w->device));
    return VRT_OK;
}

// ---------------------------------------------------------------- sync staging

// The staging halves and their events, allocated by the first sync and kept.
static int ensure_staging(vrt_world* w) {
    if (w->h_stage) return VRT_OK;
    if (fault_injected("vrt_world_sync:staging")) return fail(VRT_ERR_HIP, "vrt_world_sync: injected staging failure");
    uint8_t* h = nullptr;
    uint8_t* d = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipHostMalloc((void**)&h, 2 * kSyncHalf, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&d, 2 * kSyncHalf);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    if (e != hipSuccess) {
        if (h) (void)hipHostFree(h);
        (void)hipFree(d);
        for (hipEvent_t x : ev)
            if (x) (void)hipEventDestroy(x);
        return hip_fail(e, "vrt_world_sync: staging buffers");
    }
    w->h_stage = h;
    w->d_stage = d;
    w->stage_ev[0] = ev[0];
    w->stage_ev[1] = ev[1];
    return VRT_OK;
}

// The move area for n moved bricks: the source and destination slot lists (256-byte
// aligned), then 576 bytes per brick.
static size_t moves_bytes(uint32_t n) { return (((size_t)n * 8 + 255) & ~(size_t)255) + (size_t)n * 576; }

static int ensure_moves(vrt_world* w, uint32_t n) {
    if (n == 0) return VRT_OK;
    if (!w->moves_ev) VRT_HIP(hipEventCreateWithFlags(&w->moves_ev, hipEventDisableTiming));
    size_t need = moves_bytes(n);
    if (w->moves_cap >= need) return VRT_OK;
    if (w->moves_pending) VRT_HIP(hipEventSynchronize(w->moves_ev));  // the old area's last reader
    w->moves_pending = false;
    (void)hipFree(w->d_moves);
    w->d_moves = nullptr;
    w->moves_cap = 0;
    size_t cap = std::bit_ceil(std::max<size_t>(need, (size_t)1 << 20));
    VRT_HIP(hipMalloc(&w->d_moves, cap));
    w->moves_cap = cap;
    return VRT_OK;
}

// The next staging half, once the host may write its pinned part and the stream its device
// part (the last kernel or copy that read it has finished).
static hipError_t acquire_half(vrt_world* w, uint32_t& h) {
    h = w->stage_next;
    w->stage_next ^= 1u;
    if (w->stage_pending[h]) {
        hipError_t e = hipEventSynchronize(w->stage_ev[h]);
        if (e != hipSuccess) return e;
        w->stage_pending[h] = false;
    }
    return hipSuccess;
}

static hipError_t release_half(vrt_world* w, uint32_t h, hipStream_t s) {
    hipError_t e = hipEventRecord(w->stage_ev[h], s);
    if (e == hipSuccess) w->stage_pending[h] = true;
    return e;
}

// Host bytes to device memory through the pinned halves, kSyncHalf at a time.
static hipError_t stream_to_device(vrt_world* w, hipStream_t s, uint8_t* dst, const uint8_t* src, size_t bytes) {
    for (size_t off = 0; off < bytes; off += kSyncHalf) {
        size_t len = std::min(kSyncHalf, bytes - off);
        uint32_t h;
        hipError_t e = acquire_half(w, h);
        if (e != hipSuccess) return e;
        std::memcpy(w->h_stage + h * kSyncHalf, src + off, len);
        e = hipMemcpyAsync(dst + off, w->h_stage + h * kSyncHalf, len, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = release_half(w, h, s);
        if (e != hipSuccess) return e;
        w->upload_bytes += len;
    }
    return hipSuccess;
}

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

void vrt::ray_cell_lut(uint64_t table[512]) {
    for (uint32_t oct = 0; oct < 8; oct++)
        for (uint32_t c = 0; c < 64; c++) {
            uint32_t cx = c & 3u, cz = (c >> 2) & 3u, cy = (c >> 4) & 3u;
            // reachable coordinates along each axis: [c, 3] going up, [0, c] going down
            auto range = [&](uint32_t v, bool up) -> uint32_t { return up ? (0xFu << v) & 0xFu : 0xFu >> (3u - v); };
            uint32_t rx = range(cx, oct & 1u), ry = range(cy, oct & 2u), rz = range(cz, oct & 4u);
            uint64_t m = 0;
            for (uint32_t b = 0; b < 64; b++)
                if ((rx >> (b & 3u) & 1u) && (rz >> ((b >> 2) & 3u) & 1u) && (ry >> (b >> 4) & 1u)) m |= 1ull << b;
            table[oct * 64 + c] = m;
        }
}

// ---------------------------------------------------------------- sync

// The DirtyLocs protocol of GpuRenderer.cpp:45-167 for a set of sectors: sector sector_idx[i]
// holds the bricks of alloc_mask[i] afterwards, and the bytes of those in dirty_mask[i] (all of
// them when dirty_mask is NULL) follow in `bricks`, sector after sector, in increasing brick
// index.  Bricks that stay allocated and are not dirty keep their device contents; when their
// sector's slot run moves or their popcount position shifts (the reference then re-uploads the
// whole sector, BrickSlotAllocator.cpp:7-23), they are moved on the device.  Every allocation
// is made before any device work, and the host mirror is committed only after all device work
// has been issued: a failure before the device work leaves the world unchanged, a failure
// during it marks the world failed.
static int world_sync(vrt_world* w, uint32_t nsec, const uint32_t* sector_idx, const uint64_t* alloc_mask,
                      const uint64_t* dirty_mask, const uint8_t* bricks, hipStream_t s, const char* entry) {
    const std::string name(entry);
    if (!w) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": null world");
    int rc = check_world(w, entry);
    if (rc != VRT_OK) return rc;
    if (nsec == 0) return VRT_OK;
    if (!sector_idx || !alloc_mask) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": null sector arrays");
    if (fault_injected("vrt_world_sync")) return fail(VRT_ERR_HIP, name + ": injected HIP fault");
    {
        std::vector<uint32_t> sorted(sector_idx, sector_idx + nsec);
        std::sort(sorted.begin(), sorted.end());
        if (sorted.back() >= w->nsec)
            return fail(VRT_ERR_INVALID_ARGUMENT, name + ": sector index " + std::to_string(sorted.back()) +
                                                      " out of range (" + std::to_string(w->nsec) + " sectors)");
        for (uint32_t i = 1; i < nsec; i++)
            if (sorted[i] == sorted[i - 1])
                return fail(VRT_ERR_INVALID_ARGUMENT, name + ": sector " + std::to_string(sorted[i]) + " listed twice");
    }
    // allocation on a copy of the arena, committed only when every sector fits
    SlotArena arena = w->arena;
    std::vector<uint32_t> new_base(nsec);
    uint64_t nup = 0;
    for (uint32_t i = 0; i < nsec; i++) {
        uint32_t idx = sector_idx[i];
        uint64_t old_m = w->h_mask[idx], new_m = alloc_mask[i];
        uint64_t dirty = dirty_mask ? dirty_mask[i] & new_m : new_m;
        if (uint64_t fresh = new_m & ~old_m & ~dirty)
            return fail(VRT_ERR_INVALID_ARGUMENT, name + ": brick " + std::to_string(std::countr_zero(fresh)) +
                                                      " of sector " + std::to_string(idx) +
                                                      " is newly allocated but not dirty (its bytes are needed)");
        uint32_t base = w->h_base[idx];
        if (!arena.realloc(base, (uint32_t)std::popcount(old_m), (uint32_t)std::popcount(new_m)))
            return fail(VRT_ERR_CAPACITY, "Could not allocate brick slots (sector " + std::to_string(idx) + ", " +
                                              std::to_string(std::popcount(new_m)) + " bricks, " +
                                              std::to_string(arena.allocated()) + "/" +
                                              std::to_string(arena.capacity()) + " slots in use)");
        new_base[i] = new_m ? base : 0u;
        nup += (uint64_t)std::popcount(dirty);
    }
    if (nup && !bricks) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": null brick data");
    if (nup > 0xFFFFFFFFull) return fail(VRT_ERR_INVALID_ARGUMENT, name + ": too many bricks in one call");
    // destination slot of every uploaded brick (in the order of `bricks`), and the kept bricks
    // whose slot changes
    std::vector<uint32_t> up_slot;
    up_slot.reserve((size_t)nup);
    std::vector<uint32_t> mv_src, mv_dst;
    for (uint32_t i = 0; i < nsec; i++) {
        uint32_t idx = sector_idx[i];
        uint64_t old_m = w->h_mask[idx], new_m = alloc_mask[i];
        uint64_t dirty = dirty_mask ? dirty_mask[i] & new_m : new_m;
        for (uint64_t m = new_m; m; m &= m - 1) {
            uint32_t b = (uint32_t)std::countr_zero(m);
            uint64_t below = (1ull << b) - 1ull;
            uint32_t slot = new_base[i] + (uint32_t)std::popcount(new_m & below);
            if ((dirty >> b) & 1ull) {
                up_slot.push_back(slot);
            } else {
                uint32_t old_slot = w->h_base[idx] + (uint32_t)std::popcount(old_m & below);
                if (old_slot != slot) {
                    mv_src.push_back(old_slot);
                    mv_dst.push_back(slot);
                }
            }
        }
    }
    const uint32_t nmv = (uint32_t)mv_src.size();
    // sector records and the sector-of-sectors words they change
    std::vector<uint4> recs(nsec);
    std::vector<uint32_t> rec_idx(sector_idx, sector_idx + nsec);
    std::map<uint32_t, uint64_t> groups;
    {
        uint32_t nxz = 1u << w->shift_xz;
        uint32_t gm_xz = (1u << w->gshift_xz) - 1u, gm_y = (1u << w->gshift_y) - 1u;
        for (uint32_t i = 0; i < nsec; i++) {
            uint32_t idx = sector_idx[i];
            recs[i] = sector_record(alloc_mask[i], new_base[i]);
            uint32_t sx = idx & (nxz - 1u), sz = (idx >> w->shift_xz) & (nxz - 1u), sy = idx >> (2 * w->shift_xz);
            uint32_t g = ((sx >> 2) & gm_xz) | ((sz >> 2) & gm_xz) << w->gshift_xz |
                         ((sy >> 2) & gm_y) << (2 * w->gshift_xz);
            uint64_t bit = 1ull << ((sx & 3u) | (sz & 3u) << 2 | (sy & 3u) << 4);
            auto it = groups.emplace(g, w->h_groups[g]).first;
            it->second = alloc_mask[i] ? (it->second | bit) : (it->second & ~bit);
        }
    }
    // every allocation before any device work: slot growth, staging, the move area
    rc = grow_slots(w, arena.high_water(), s);
    if (rc == VRT_OK) rc = ensure_staging(w);
    if (rc == VRT_OK) rc = ensure_moves(w, nmv);
    if (rc != VRT_OK) return rc;

    // device work; a failure from here on leaves the device tables part-written
    auto device_fail = [&](hipError_t e, const char* what) -> int {
        w->failed = true;
        hip_fail(e, what);
        return fail(VRT_ERR_HIP, name + ": " + vrt_last_error() + "; the world is unusable");
    };
    hipError_t e = hipSuccess;
    if (nmv) {
        // gather every moved brick before any slot is written, then place them
        if (w->moves_pending) e = hipStreamWaitEvent(s, w->moves_ev, 0);
        std::vector<uint32_t> lists(mv_src);
        lists.insert(lists.end(), mv_dst.begin(), mv_dst.end());
        if (e == hipSuccess) e = stream_to_device(w, s, w->d_moves, (const uint8_t*)lists.data(), lists.size() * 4);
        if (e != hipSuccess) return device_fail(e, "moving kept bricks");
        const uint32_t* d_src = (const uint32_t*)w->d_moves;
        const uint32_t* d_dst = d_src + nmv;
        uint8_t* area = w->d_moves + (((size_t)nmv * 8 + 255) & ~(size_t)255);
        uint32_t blocks = (uint32_t)(((uint64_t)nmv * 9 + 255) / 256);
        sync_gather_kernel<<<blocks, 256, 0, s>>>(w->d_vox, w->d_occ, d_src, nmv, area);
        sync_place_kernel<<<blocks, 256, 0, s>>>(w->d_vox, w->d_occ, d_dst, nmv, area);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(w->moves_ev, s);
        if (e != hipSuccess) return device_fail(e, "moving kept bricks");
        w->moves_pending = true;
    }
    // uploaded bricks, a staging half at a time: the bytes, then their slots
    const uint64_t per_chunk = kSyncHalf / (512 + 4);
    for (uint64_t k0 = 0; k0 < nup; k0 += per_chunk) {
        uint32_t cnt = (uint32_t)std::min<uint64_t>(per_chunk, nup - k0);
        uint32_t h;
        e = acquire_half(w, h);
        if (e != hipSuccess) return device_fail(e, "uploading bricks");
        uint8_t* hp = w->h_stage + h * kSyncHalf;
        uint8_t* dp = w->d_stage + h * kSyncHalf;
        size_t bytes_b = (size_t)cnt * 512;
        std::memcpy(hp, bricks + k0 * 512, bytes_b);
        std::memcpy(hp + bytes_b, up_slot.data() + k0, (size_t)cnt * 4);
        size_t bytes = bytes_b + (size_t)cnt * 4;
        e = hipMemcpyAsync(dp, hp, bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            sync_upload_kernel<<<(uint32_t)(((uint64_t)cnt * 8 + 255) / 256), 256, 0, s>>>(
                w->d_vox, w->d_occ, dp, (const uint32_t*)(dp + bytes_b), cnt);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = release_half(w, h, s);
        if (e != hipSuccess) return device_fail(e, "uploading bricks");
        w->upload_bytes += bytes;
        w->upload_brick_bytes += bytes_b;
    }
    if (fault_injected("vrt_world_sync:device")) return device_fail(hipErrorUnknown, "injected device failure");
    // sector records, then the changed sector-of-sectors words, a staging half at a time
    std::vector<uint64_t> gval;
    std::vector<uint32_t> gidx;
    for (auto& [g, v] : groups) {
        gidx.push_back(g);
        gval.push_back(v);
    }
    const uint64_t nrec = nsec, ngrp = gidx.size();
    for (uint64_t r0 = 0, g0 = 0; r0 < nrec || g0 < ngrp;) {
        uint32_t nr = (uint32_t)std::min<uint64_t>(nrec - r0, (kSyncHalf - 256) / 20);
        uint32_t ng = nr ? 0u : (uint32_t)std::min<uint64_t>(ngrp - g0, (kSyncHalf - 256) / 12);
        uint32_t h;
        e = acquire_half(w, h);
        if (e != hipSuccess) return device_fail(e, "sector records");
        uint8_t* hp = w->h_stage + h * kSyncHalf;
        uint8_t* dp = w->d_stage + h * kSyncHalf;
        size_t o_idx = align16((size_t)nr * 16), o_gval = align16(o_idx + (size_t)nr * 4);
        size_t o_gidx = o_gval + (size_t)ng * 8, bytes = o_gidx + (size_t)ng * 4;
        std::memcpy(hp, recs.data() + r0, (size_t)nr * 16);
        std::memcpy(hp + o_idx, rec_idx.data() + r0, (size_t)nr * 4);
        std::memcpy(hp + o_gval, gval.data() + g0, (size_t)ng * 8);
        std::memcpy(hp + o_gidx, gidx.data() + g0, (size_t)ng * 4);
        e = hipMemcpyAsync(dp, hp, bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            sync_records_kernel<<<(nr + ng + 255) / 256, 256, 0, s>>>(
                w->d_sec, (const uint4*)dp, (const uint32_t*)(dp + o_idx), nr, w->d_groups,
                (const uint64_t*)(dp + o_gval), (const uint32_t*)(dp + o_gidx), ng);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = release_half(w, h, s);
        if (e != hipSuccess) return device_fail(e, "sector records");
        w->upload_bytes += bytes;
        r0 += nr;
        g0 += ng;
    }
    // commit the host mirror
    w->arena = arena;
    for (uint32_t i = 0; i < nsec; i++) {
        uint32_t idx = sector_idx[i];
        w->num_bricks -= (uint32_t)std::popcount(w->h_mask[idx]);
        w->num_bricks += (uint32_t)std::popcount(alloc_mask[i]);
        w->h_mask[idx] = alloc_mask[i];
        w->h_base[idx] = new_base[i];
    }
    for (size_t j = 0; j < gidx.size(); j++) w->h_groups[gidx[j]] = gval[j];
    w->moved_bricks += nmv;
    return VRT_OK;
}

// ---------------------------------------------------------------- C ABI

extern "C" {

const char* vrt_last_error(void) { return g_error.c_str(); }
int vrt_abi_version(void) { return VRT_ABI_VERSION; }

int vrt_world_create(const vrt_world_desc* desc, void* stream, vrt_world** out) {
    return guarded([&]() -> int {
        if (!desc || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_create: null argument");
        *out = nullptr;
        if (desc->log2_size_xz < 5 || desc->log2_size_xz > 14 || desc->log2_size_y < 5 || desc->log2_size_y > 14)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_create: log2 extents must be in 5..14");
        if (fault_injected("vrt_world_create")) return fail(VRT_ERR_HIP, "vrt_world_create: injected HIP fault");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            return fail(VRT_ERR_NO_DEVICE, "vrt_world_create: no HIP device");
        hipStream_t s = (hipStream_t)stream;
        int dev = 0;
        VRT_HIP(hipGetDevice(&dev));
        auto* w = new vrt_world();
        w->device = dev;
        w->log2_xz = desc->log2_size_xz;
        w->log2_y = desc->log2_size_y;
        w->shift_xz = w->log2_xz - 5;
        w->shift_y = w->log2_y - 5;
        w->nsec = 1u << (2 * w->shift_xz + w->shift_y);
        // the arena bounds slot addresses only (memory follows the high-water mark): by default
        // twice the bricks the grid can hold, so that a sector whose run must move finds a
        // contiguous free range even when the world is nearly full
        uint64_t max_slots = (uint64_t)w->nsec * 128;
        if (desc->max_slots) max_slots = std::min<uint64_t>(max_slots, desc->max_slots);
        // occupancy indices slot * 8 + cell are 32-bit in the traversal: at most 2^28 slots
        // (154 GB of brick data)
        max_slots = std::min<uint64_t>(max_slots, kMaxSlots);
        w->arena = SlotArena((uint32_t)max_slots);
        w->h_mask.assign(w->nsec, 0);
        w->h_base.assign(w->nsec, 0);
        w->gshift_xz = w->shift_xz > 2 ? w->shift_xz - 2 : 0;
        w->gshift_y = w->shift_y > 2 ? w->shift_y - 2 : 0;
        w->h_groups.assign((size_t)1 << (2 * w->gshift_xz + w->gshift_y), 0);
        static const std::array<uint64_t, 512> lut = [] {
            std::array<uint64_t, 512> t;
            ray_cell_lut(t.data());
            return t;
        }();
        int rc = VRT_OK;
        auto check = [&](hipError_t e, const char* what) {
            if (e != hipSuccess && rc == VRT_OK) rc = hip_fail(e, what);
        };
        check(hipMalloc(&w->d_sec, (size_t)w->nsec * 16), "hipMalloc(sector records)");
        check(hipMalloc(&w->d_palette, 256 * 8), "hipMalloc(palette)");
        check(hipMalloc(&w->d_groups, w->h_groups.size() * 8), "hipMalloc(sector groups)");
        check(hipMalloc(&w->d_lut, 512 * 8), "hipMalloc(ray-cell table)");
        if (rc == VRT_OK) {
            std::vector<uint4> empty(w->nsec, sector_record(0, 0));
            check(hipMemcpyAsync(w->d_sec, empty.data(), (size_t)w->nsec * 16, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(sector records)");
            check(hipStreamSynchronize(s), "hipStreamSynchronize(sector records)");
            check(hipMemsetAsync(w->d_palette, 0, 256 * 8, s), "hipMemsetAsync(palette)");
            check(hipMemsetAsync(w->d_groups, 0, w->h_groups.size() * 8, s), "hipMemsetAsync(sector groups)");
            check(hipMemcpyAsync(w->d_lut, lut.data(), 512 * 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync(ray-cell table)");
        }
        if (rc == VRT_OK && desc->initial_slots) rc = grow_slots(w, std::min(desc->initial_slots, w->arena.capacity()), s);
        if (rc == VRT_OK) check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (rc != VRT_OK) {
            free_world(w);
            return rc;
        }
        *out = w;
        return VRT_OK;
    });
}

int vrt_world_destroy(vrt_world* world) {
    return guarded([&]() -> int {
        if (!world) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_destroy: null world");
        (void)hipDeviceSynchronize();
        free_world(world);
        return VRT_OK;
    });
}

int vrt_world_get_info(const vrt_world* w, vrt_world_info* out) {
    return guarded([&]() -> int {
        if (!w || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_get_info: null argument");
        out->size_xz = 1u << w->log2_xz;
        out->size_y = 1u << w->log2_y;
        out->sectors_xz = 1u << w->shift_xz;
        out->sectors_y = 1u << w->shift_y;
        out->num_sectors = w->nsec;
        out->num_bricks = w->num_bricks;
        out->slot_capacity = w->slot_cap;
        out->device_bytes = (uint64_t)w->nsec * 16 + 256 * 8 + (uint64_t)w->slot_cap * (512 + 64) +
                            (uint64_t)w->h_groups.size() * 8 + 512 * 8 + (w->d_stage ? 2 * kSyncHalf : 0) +
                            w->moves_cap;
        out->slot_high_water = w->arena.high_water();
        out->generation = w->generation;
        out->upload_bytes = w->upload_bytes;
        out->upload_brick_bytes = w->upload_brick_bytes;
        out->moved_bricks = w->moved_bricks;
        return VRT_OK;
    });
}

int vrt_world_get_buffers(const vrt_world* w, vrt_world_buffers* out) {
    return guarded([&]() -> int {
        if (!w || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_get_buffers: null argument");
        out->sectors = (const vrt_sector_record*)w->d_sec;
        out->occupancy = w->d_occ;
        out->voxels = w->d_vox;
        out->palette = w->d_palette;
        return VRT_OK;
    });
}

int vrt_world_set_palette(vrt_world* w, const uint64_t* palette, void* stream) {
    return guarded([&]() -> int {
        if (!w || !palette) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_set_palette: null argument");
        int rc = check_world(w, "vrt_world_set_palette");
        if (rc != VRT_OK) return rc;
        hipStream_t s = (hipStream_t)stream;
        VRT_HIP(hipMemcpyAsync(w->d_palette, palette, 256 * 8, hipMemcpyHostToDevice, s));
        VRT_HIP(hipStreamSynchronize(s));
        return VRT_OK;
    });
}

int vrt_world_sync(vrt_world* w, uint32_t nsec, const uint32_t* sector_idx, const uint64_t* alloc_mask,
                   const uint8_t* bricks, void* stream) {
    return guarded([&]() -> int {
        return world_sync(w, nsec, sector_idx, alloc_mask, nullptr, bricks, (hipStream_t)stream, "vrt_world_sync");
    });
}

int vrt_world_sync_bricks(vrt_world* w, uint32_t nsec, const uint32_t* sector_idx, const uint64_t* alloc_mask,
                          const uint64_t* dirty_mask, const uint8_t* bricks, void* stream) {
    return guarded([&]() -> int {
        if (nsec && !dirty_mask) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync_bricks: null dirty masks");
        return world_sync(w, nsec, sector_idx, alloc_mask, dirty_mask, bricks, (hipStream_t)stream,
                          "vrt_world_sync_bricks");
    });
}

}  // extern "C"
