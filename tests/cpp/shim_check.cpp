// shim_check.cpp - drives the C++ shim (include/vrt.hpp) the way the reference's
// callers use VoxelMap::RayCast (Main.cpp:248, Brush.cpp:47).
//   shim_check nodevice              expects vrt::Error(VRT_ERR_NO_DEVICE) from a World
//   shim_check pick VOX N RAYS       VOX: N^3 uint8 voxels [y][z][x]; RAYS: doubles
//                                    ox oy oz dx dy dz per ray; prints one line per ray:
//                                    distance vx vy vz nx ny nz u v
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vrt.hpp"
#include "vrt_scene.h"

struct D3 {
    double x, y, z;  // the shape of glm::dvec3
};

static std::vector<char> read_file(const char* path) {
    std::vector<char> buf;
    FILE* f = std::fopen(path, "rb");
    if (!f) return buf;
    char tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    std::fclose(f);
    return buf;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "nodevice") == 0) {
        vrt::HitResult h;
        if (!h.IsMiss()) return 3;
        try {
            vrt::World w(6, 6);
            std::printf("unexpected: a world was created\n");
            return 1;
        } catch (const vrt::Error& e) {
            std::printf("status %d: %s\n", e.status, e.what());
            return e.status == VRT_ERR_NO_DEVICE ? 0 : 2;
        }
    }
    if (argc == 5 && std::strcmp(argv[1], "pick") == 0) {
        uint32_t n = (uint32_t)std::atoi(argv[3]);
        std::vector<char> vox = read_file(argv[2]), rays = read_file(argv[4]);
        if (vox.size() != (size_t)n * n * n || rays.size() % 48 != 0) return 4;
        uint32_t log2 = 0;
        while ((1u << log2) < n) log2++;
        vrt_scene* sc = nullptr;
        vrt::check(vrt_scene_from_dense(log2, log2, (const uint8_t*)vox.data(), &sc));
        vrt_world* raw = nullptr;
        vrt::check(vrt_world_from_scene(sc, nullptr, &raw));
        vrt_scene_destroy(sc);
        vrt::World world(raw);
        vrt::VoxelMapCaster caster(world);
        const double* r = (const double*)rays.data();
        for (size_t i = 0; i < rays.size() / 48; i++) {
            D3 o{r[6 * i + 0], r[6 * i + 1], r[6 * i + 2]}, d{r[6 * i + 3], r[6 * i + 4], r[6 * i + 5]};
            vrt::HitResult h = caster.RayCast(o, d, 1024);
            std::printf("%.17g %d %d %d %.9g %.9g %.9g %.9g %.9g\n", h.Distance, h.VoxelPos.x, h.VoxelPos.y,
                        h.VoxelPos.z, h.Normal.x, h.Normal.y, h.Normal.z, h.UV.x, h.UV.y);
        }
        return 0;
    }
    std::fprintf(stderr, "usage: shim_check nodevice | pick VOX N RAYS\n");
    return 5;
}
