// vrt_scene_impl.h - the host scene object behind vrt_scene.h (not part of the ABI).
#pragma once

#include <cstdint>
#include <map>
#include <unordered_map>
#include <vector>

// A sector being edited: its allocation mask and all 64 bricks dense by brick index (32 KiB),
// so a voxel write allocates a brick without moving the packed stream's bytes.
struct vrt_edit_sector {
    uint64_t mask = 0;
    std::vector<uint8_t> bricks;  // 64 * 512 bytes
};

struct vrt_scene {
    uint32_t log2_xz = 0, log2_y = 0;
    uint32_t shift_xz = 0, shift_y = 0;
    // the packed sector stream: sectors in increasing index, popcount(mask) bricks each
    std::vector<uint32_t> sector_idx;
    std::vector<uint64_t> alloc_mask;
    std::vector<uint8_t> bricks;
    std::vector<uint64_t> brick_offset;   // first brick of each listed sector
    std::vector<int32_t> sector_lookup;   // sector index -> position in the lists, -1 if absent
    uint64_t solid = 0;
    bool has_palette = false;             // palette below set (a loaded file), else placeholder colours
    uint64_t palette[256] = {};           // encoded materials (Material::GetEncoded)
    // sectors written since the stream was last packed; they shadow the stream's entries
    std::unordered_map<uint32_t, vrt_edit_sector> edits;
    // VoxelMap::DirtyLocs (VoxelMap.h:185): sector index -> bricks changed since the last sync
    std::map<uint32_t, uint64_t> dirty;
};

// Folds the edited sectors back into the packed stream (sectors left without bricks are
// dropped) and clears the edit overlay; the dirty set is kept.
void scene_pack_edits(vrt_scene* sc);
