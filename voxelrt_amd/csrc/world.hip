// world.hip - device world lifecycle, slot arena, sector sync and occupancy build.
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

// ---------------------------------------------------------------- occupancy build

// One thread per (brick, 4x4x4 cell): bit vx + 4 vz + 16 vy of cell
// cx | cz << 1 | cy << 2, as in CpuRenderer.cpp:63-83 (UpdateOccupancy.comp
// writes the same 64-bit mask as two 4x2x4 halves).
__global__ __launch_bounds__(256) void build_occupancy_kernel(const uint8_t* __restrict__ vox,
                                                              uint64_t* __restrict__ occ,
                                                              const uint32_t* __restrict__ slots,
                                                              uint32_t nslots) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t b = gid >> 3, cell = gid & 7;
    if (b >= nslots) return;
    uint32_t slot = slots[b];
    const uint8_t* brick = vox + (size_t)slot * 512;
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
    occ[(size_t)slot * 8 + cell] = m;
}

hipError_t launch_build_occupancy(vrt_world* w, const uint32_t* d_slots, uint32_t nslots, hipStream_t s) {
    if (nslots == 0) return hipSuccess;
    uint64_t threads = (uint64_t)nslots * 8;
    uint32_t blocks = (uint32_t)((threads + 255) / 256);
    build_occupancy_kernel<<<blocks, 256, 0, s>>>(w->d_vox, w->d_occ, d_slots, nslots);
    return hipGetLastError();
}

// The records of the synced sectors (indices unique) into the sector table.
__global__ __launch_bounds__(256) void scatter_records_kernel(uint4* __restrict__ sec, const uint4* __restrict__ rec,
                                                              const uint32_t* __restrict__ idx, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sec[idx[i]] = rec[i];
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
    return VRT_OK;
}

static void free_world(vrt_world* w) {
    (void)hipFree(w->d_sec);
    (void)hipFree(w->d_occ);
    (void)hipFree(w->d_vox);
    (void)hipFree(w->d_palette);
    (void)hipFree(w->d_groups);
    (void)hipFree(w->d_lut);
    delete w;
}

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

// Sector-of-sectors bit of sector idx, from its current mask.
static void update_group_bit(vrt_world* w, uint32_t idx) {
    uint32_t nxz = 1u << w->shift_xz;
    uint32_t sx = idx & (nxz - 1u), sz = (idx >> w->shift_xz) & (nxz - 1u), sy = idx >> (2 * w->shift_xz);
    uint32_t gm_xz = (1u << w->gshift_xz) - 1u, gm_y = (1u << w->gshift_y) - 1u;
    uint32_t g = ((sx >> 2) & gm_xz) | ((sz >> 2) & gm_xz) << w->gshift_xz | ((sy >> 2) & gm_y) << (2 * w->gshift_xz);
    uint64_t bit = 1ull << ((sx & 3u) | (sz & 3u) << 2 | (sy & 3u) << 4);
    if (w->h_mask[idx])
        w->h_groups[g] |= bit;
    else
        w->h_groups[g] &= ~bit;
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
        uint64_t max_slots = (uint64_t)w->nsec * 64;
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
                            (uint64_t)w->h_groups.size() * 8 + 512 * 8;
        out->slot_high_water = w->arena.high_water();
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
        hipStream_t s = (hipStream_t)stream;
        VRT_HIP(hipMemcpyAsync(w->d_palette, palette, 256 * 8, hipMemcpyHostToDevice, s));
        VRT_HIP(hipStreamSynchronize(s));
        return VRT_OK;
    });
}

int vrt_world_sync(vrt_world* w, uint32_t nsec, const uint32_t* sector_idx, const uint64_t* alloc_mask,
                   const uint8_t* bricks, void* stream) {
    return guarded([&]() -> int {
        if (!w) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync: null world");
        if (nsec == 0) return VRT_OK;
        if (!sector_idx || !alloc_mask) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync: null sector arrays");
        if (fault_injected("vrt_world_sync")) return fail(VRT_ERR_HIP, "vrt_world_sync: injected HIP fault");
        hipStream_t s = (hipStream_t)stream;
        {
            std::vector<uint32_t> sorted(sector_idx, sector_idx + nsec);
            std::sort(sorted.begin(), sorted.end());
            if (sorted.back() >= w->nsec)
                return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync: sector index " + std::to_string(sorted.back()) +
                                                          " out of range (" + std::to_string(w->nsec) + " sectors)");
            for (uint32_t i = 1; i < nsec; i++)
                if (sorted[i] == sorted[i - 1])
                    return fail(VRT_ERR_INVALID_ARGUMENT,
                                "vrt_world_sync: sector " + std::to_string(sorted[i]) + " listed twice");
        }
        // Allocation pass on copies, committed only when every sector fits.
        SlotArena arena = w->arena;
        struct Run {
            uint32_t dst, src, count;
        };
        std::vector<Run> runs;
        std::vector<uint32_t> new_base(nsec);
        uint64_t total = 0;
        for (uint32_t i = 0; i < nsec; i++) {
            uint32_t idx = sector_idx[i];
            uint32_t cnt_old = (uint32_t)std::popcount(w->h_mask[idx]);
            uint32_t cnt_new = (uint32_t)std::popcount(alloc_mask[i]);
            uint32_t base = w->h_base[idx];
            if (!arena.realloc(base, cnt_old, cnt_new))
                return fail(VRT_ERR_CAPACITY, "Could not allocate brick slots (sector " + std::to_string(idx) + ", " +
                                                  std::to_string(cnt_new) + " bricks, " +
                                                  std::to_string(arena.allocated()) + "/" +
                                                  std::to_string(arena.capacity()) + " slots in use)");
            if (cnt_new == 0) base = 0;
            new_base[i] = base;
            if (cnt_new) {
                if (!runs.empty() && runs.back().dst + runs.back().count == base &&
                    runs.back().src + runs.back().count == total)
                    runs.back().count += cnt_new;
                else
                    runs.push_back({base, (uint32_t)total, cnt_new});
            }
            total += cnt_new;
        }
        if (total && !bricks) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync: null brick data");
        if (total > 0xFFFFFFFFull) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync: too many bricks in one call");
        int rc = grow_slots(w, arena.high_water(), s);
        if (rc != VRT_OK) return rc;
        // commit host state
        w->arena = arena;
        std::vector<uint32_t> slots;
        slots.reserve((size_t)total);
        for (uint32_t i = 0; i < nsec; i++) {
            uint32_t idx = sector_idx[i];
            w->num_bricks -= (uint32_t)std::popcount(w->h_mask[idx]);
            w->num_bricks += (uint32_t)std::popcount(alloc_mask[i]);
            w->h_mask[idx] = alloc_mask[i];
            w->h_base[idx] = new_base[i];
            update_group_bit(w, idx);
        }
        for (const Run& r : runs)
            for (uint32_t k = 0; k < r.count; k++) slots.push_back(r.dst + k);
        // device: bricks, occupancy, headers (in stream order).  One scratch allocation holds
        // the slots whose occupancy is rebuilt and, when the synced sectors are few against the
        // grid, their records for a scatter launch (the largest grid's table is 2 GiB).
        const bool scatter = (uint64_t)nsec * 64 <= w->nsec;
        const size_t slots_bytes = (slots.size() * 4 + 15) & ~(size_t)15;
        const size_t tmp_bytes = slots_bytes + (scatter ? (size_t)nsec * 20 : 0);
        std::vector<uint8_t> h_tmp(tmp_bytes);
        std::memcpy(h_tmp.data(), slots.data(), slots.size() * 4);
        if (scatter) {
            uint4* rec = (uint4*)(h_tmp.data() + slots_bytes);
            uint32_t* idx = (uint32_t*)(h_tmp.data() + slots_bytes + (size_t)nsec * 16);
            for (uint32_t i = 0; i < nsec; i++) {
                idx[i] = sector_idx[i];
                rec[i] = sector_record(w->h_mask[sector_idx[i]], w->h_base[sector_idx[i]]);
            }
        }
        uint8_t* d_tmp = nullptr;
        if (tmp_bytes) VRT_HIP(hipMalloc(&d_tmp, tmp_bytes));
        auto done = [&](hipError_t e, const char* what) -> int {
            hipError_t e2 = hipStreamSynchronize(s);  // the scratch is in use until the stream gets here
            (void)hipFree(d_tmp);
            if (e == hipSuccess) e = e2;
            return e == hipSuccess ? VRT_OK : hip_fail(e, what);
        };
        std::vector<uint4> recs;  // the whole table, when it is uploaded: alive until done()
        hipError_t e = hipSuccess;
        for (const Run& r : runs)
            if (e == hipSuccess)
                e = hipMemcpyAsync(w->d_vox + (size_t)r.dst * 512, bricks + (size_t)r.src * 512, (size_t)r.count * 512,
                                   hipMemcpyHostToDevice, s);
        if (e == hipSuccess && tmp_bytes) e = hipMemcpyAsync(d_tmp, h_tmp.data(), tmp_bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && !slots.empty())
            e = launch_build_occupancy(w, (const uint32_t*)d_tmp, (uint32_t)slots.size(), s);
        if (e != hipSuccess) return done(e, "vrt_world_sync: bricks and occupancy");
        if (scatter) {
            scatter_records_kernel<<<(nsec + 255) / 256, 256, 0, s>>>(
                w->d_sec, (const uint4*)(d_tmp + slots_bytes), (const uint32_t*)(d_tmp + slots_bytes + (size_t)nsec * 16),
                nsec);
            e = hipGetLastError();
        } else {
            recs.resize(w->nsec);
            for (uint32_t i = 0; i < w->nsec; i++)
                recs[i] = sector_record(w->h_mask[i], w->h_base[i]);
            e = hipMemcpyAsync(w->d_sec, recs.data(), (size_t)w->nsec * 16, hipMemcpyHostToDevice, s);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(w->d_groups, w->h_groups.data(), w->h_groups.size() * 8, hipMemcpyHostToDevice, s);
        return done(e, "vrt_world_sync: sector records");
    });
}

}  // extern "C"
