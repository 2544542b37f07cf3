// vrt_scene_impl.h - the host scene object behind vrt_scene.h (not part of the ABI).
#pragma once

#include <cstdint>
#include <vector>

struct vrt_scene {
    uint32_t log2_xz = 0, log2_y = 0;
    uint32_t shift_xz = 0, shift_y = 0;
    std::vector<uint32_t> sector_idx;
    std::vector<uint64_t> alloc_mask;
    std::vector<uint8_t> bricks;
    std::vector<uint64_t> brick_offset;   // first brick of each listed sector
    std::vector<int32_t> sector_lookup;   // sector index -> position in the lists, -1 if absent
    uint64_t solid = 0;
    bool has_palette = false;             // palette below set (a loaded file), else placeholder colours
    uint64_t palette[256] = {};           // encoded materials (Material::GetEncoded)
};
