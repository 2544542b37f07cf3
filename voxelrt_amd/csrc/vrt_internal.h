// vrt_internal.h - shared internals of libvoxelrt_hip.so (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "vrt.h"

namespace vrt {

// ---------------------------------------------------------------- errors

void set_error(const std::string& msg);
void clear_error();

// Returns a vrt_status after recording the message.
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define VRT_HIP(call)                                     \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return ::vrt::hip_fail(e_, #call); \
    } while (0)

// Wraps a body so that no C++ exception crosses the ABI.
template <typename F>
int guarded(F&& f) {
    try {
        clear_error();
        return f();
    } catch (const std::bad_alloc&) {
        return fail(VRT_ERR_CAPACITY, "host allocation failed");
    } catch (const std::exception& ex) {
        return fail(VRT_ERR_STATE, ex.what());
    } catch (...) {
        return fail(VRT_ERR_STATE, "unknown exception");
    }
}

// Test hook: VRT_FAULT_INJECT=<entry name> makes that entry report a HIP error.
bool fault_injected(const char* entry);

// ---------------------------------------------------------------- slot arena

// Largest slot arena of a world.
constexpr uint64_t kMaxSlots = 1ull << 28;


// Free-list arena of brick slots.  Each sector owns one contiguous run of
// popcount(mask) slots.  Slot 0 is valid (no sentinel).  Best-fit allocation,
// growth in place when the following range is free, coalescing on free.
class SlotArena {
public:
    explicit SlotArena(uint32_t capacity = 0) : capacity_(capacity) {
        if (capacity) free_[0] = capacity;
    }
    // Moves/grows a run from (base, cur) to a run of `want` slots. Returns false when full.
    bool realloc(uint32_t& base, uint32_t cur, uint32_t want);
    void release(uint32_t base, uint32_t size);
    uint32_t allocated() const { return allocated_; }
    uint32_t capacity() const { return capacity_; }
    size_t free_ranges() const { return free_.size(); }
    // One past the highest allocated slot.
    uint32_t high_water() const;

private:
    std::map<uint32_t, uint32_t> free_;  // start -> length
    uint32_t capacity_ = 0;
    uint32_t allocated_ = 0;
};

// ---------------------------------------------------------------- device world view

// Sector record in HBM (vrt_sector_record in vrt.h): allocation mask (lo, hi), base
// slot, and the cell mask of a step at the sector level whose 2x2x2 group of bricks
// is empty: 15 (a 16-voxel cell), or 31 (the whole sector) when the mask is 0, the
// reference's +2 LOD of an empty mask; a step whose group holds a brick uses half.
inline uint4 sector_record(uint64_t mask, uint32_t base) {
    return make_uint4((uint32_t)mask, (uint32_t)(mask >> 32), base, mask ? 15u : 31u);
}

struct WorldView {
    const uint4* __restrict__ sec;       // [nsec]
    const uint64_t* __restrict__ occ;    // [slots * 8]
    const uint8_t* __restrict__ vox;     // [slots * 512]
    uint32_t size_xz, size_y;            // voxel extents
    uint32_t shift_xz, shift_y;          // log2 sector grid extents
};

// What the shader method reads beyond WorldView: the sector-of-sectors masks
// (one u64 per 4x4x4 group of sectors, bit = that sector has a brick; group grid
// of max(1, sectors / 4) per axis) and the ray-cell interaction table.
struct ShaderView {
    WorldView w;
    const uint64_t* __restrict__ groups;  // [1 << (2 * gshift_xz + gshift_y)]
    const uint64_t* __restrict__ lut;     // [8 octants * 64 cells]
    uint32_t gshift_xz, gshift_y;
};

// Ray-cell interaction table: for direction octant oct (bit a set = positive
// along axis a) and cell c of a 4x4x4 mask, the cells c + j * dir (j >= 0 per
// axis) inside the mask, i.e. the cells a ray in that octant can reach from c.
void ray_cell_lut(uint64_t table[512]);

}  // namespace vrt

struct vrt_world {
    int device = 0;
    uint32_t log2_xz = 0, log2_y = 0;
    uint32_t shift_xz = 0, shift_y = 0;
    uint32_t nsec = 0;
    uint4* d_sec = nullptr;           // packed sector records (mask, base slot)
    uint64_t* d_occ = nullptr;
    uint8_t* d_vox = nullptr;
    uint64_t* d_palette = nullptr;
    uint32_t slot_cap = 0;            // device slots backed by memory
    std::vector<uint64_t> h_mask;     // host mirror of the sector headers
    std::vector<uint32_t> h_base;
    vrt::SlotArena arena;
    uint32_t num_bricks = 0;
    // shader method: sector-of-sectors masks and the ray-cell table
    uint32_t gshift_xz = 0, gshift_y = 0;
    std::vector<uint64_t> h_groups;
    uint64_t* d_groups = nullptr;
    uint64_t* d_lut = nullptr;
    // bumped whenever d_occ / d_vox move (slot growth): device pointers taken earlier, e.g. in a
    // captured graph, are stale once it changes
    uint32_t generation = 0;
    // set when a sync fails after its device work began: the device tables may then disagree
    // with the host mirror, and every later call on the world returns VRT_ERR_STATE
    bool failed = false;
    // sync staging, owned by the world and reused by every sync: two pinned host halves and two
    // device halves of kSyncHalf bytes (a chunk is packed into one half while the other's copy
    // and apply kernel run); stage_ev[h] is recorded after the last kernel that reads half h
    uint8_t* h_stage = nullptr;
    uint8_t* d_stage = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_pending[2] = {false, false};
    uint32_t stage_next = 0;
    // bricks whose slot changes while their contents stay: gathered here, then written back
    uint8_t* d_moves = nullptr;
    size_t moves_cap = 0;
    hipEvent_t moves_ev = nullptr;
    bool moves_pending = false;
    // totals since creation
    uint64_t upload_bytes = 0, upload_brick_bytes = 0, moved_bricks = 0;

    vrt::WorldView view() const {
        vrt::WorldView v;
        v.sec = d_sec;
        v.occ = d_occ;
        v.vox = d_vox;
        v.size_xz = 1u << log2_xz;
        v.size_y = 1u << log2_y;
        v.shift_xz = shift_xz;
        v.shift_y = shift_y;
        return v;
    }
    vrt::ShaderView shader_view() const {
        vrt::ShaderView v;
        v.w = view();
        v.groups = d_groups;
        v.lut = d_lut;
        v.gshift_xz = gshift_xz;
        v.gshift_y = gshift_y;
        return v;
    }
};

namespace vrt {
// Bytes of one sync staging half (host pinned and device).
constexpr size_t kSyncHalf = (size_t)8 << 20;

// VRT_OK, or VRT_ERR_STATE (with the message) when a failed sync left the world unusable.
int check_world(const vrt_world* w, const char* entry);
}  // namespace vrt
