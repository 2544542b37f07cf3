// scene.hip - host-side synthetic worlds and rays (include/vrt_scene.h).
//
// Replaces the reference's FastNoise2 terrain producer
// (src/VoxelRT/TerrainGenerator.cpp:5-146: one job per 32^3 sector, a worker
// pool) with seeded value-noise generators, and its primary-ray generator
// (GetPrimaryRay, src/VoxelRT/CpuRenderer.cpp:226-233) with the pinhole camera
// of vrt_math.h.  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vrt_internal.h"
#include "vrt_math.h"
#include "vrt_scene.h"
#include "vrt_scene_impl.h"

namespace {

using namespace vrt;

// ---------------------------------------------------------------- hashing and noise

inline uint32_t fmix32(uint32_t x) {  // MurmurHash3 finaliser
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
inline uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return fmix32((uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^ (uint32_t)z * 0xcb1ab31fu ^ fmix32(seed));
}
inline double lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return (double)hash3(x, y, z, seed) * (1.0 / 4294967296.0);
}
inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

// Bilinear value noise on the integer lattice, smoothstep weights, in [0, 1).
double noise2(double x, double z, uint32_t seed) {
    double fx = std::floor(x), fz = std::floor(z);
    int32_t ix = (int32_t)fx, iz = (int32_t)fz;
    double tx = smooth(x - fx), tz = smooth(z - fz);
    double a = lattice(ix, 0, iz, seed), b = lattice(ix + 1, 0, iz, seed);
    double c = lattice(ix, 0, iz + 1, seed), d = lattice(ix + 1, 0, iz + 1, seed);
    double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    return ab + (cd - ab) * tz;
}
// Trilinear value noise, smoothstep weights, in [0, 1).
double noise3(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double v[2][2];
    for (int j = 0; j < 2; j++)
        for (int k = 0; k < 2; k++) {
            double a = lattice(ix, iy + j, iz + k, seed), b = lattice(ix + 1, iy + j, iz + k, seed);
            v[j][k] = a + (b - a) * tx;
        }
    double y0 = v[0][0] + (v[1][0] - v[0][0]) * ty;
    double y1 = v[0][1] + (v[1][1] - v[0][1]) * ty;
    return y0 + (y1 - y0) * tz;
}
// fBm with lacunarity 2, normalised by the sum of octave amplitudes.
double fbm2(double x, double z, uint32_t seed, uint32_t oct, double freq, double gain) {
    double sum = 0, norm = 0, amp = 1, f = freq;
    for (uint32_t o = 0; o < oct; o++) {
        sum += amp * noise2(x * f, z * f, seed + 1013u * o);
        norm += amp;
        amp *= gain;
        f *= 2.0;
    }
    return sum / norm;
}
double fbm3(double x, double y, double z, uint32_t seed, uint32_t oct, double freq, double gain) {
    double sum = 0, norm = 0, amp = 1, f = freq;
    for (uint32_t o = 0; o < oct; o++) {
        sum += amp * noise3(x * f, y * f, z * f, seed + 7919u * o);
        norm += amp;
        amp *= gain;
        f *= 2.0;
    }
    return sum / norm;
}

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline double unit(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

uint32_t pick_threads(uint32_t t) {
    if (t) return t;
    unsigned h = std::thread::hardware_concurrency();
    return h ? std::min(h, 64u) : 4u;
}

template <typename F>
void parallel_for(uint32_t n, uint32_t threads, F&& f) {
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

}  // namespace

// ---------------------------------------------------------------- scene object


namespace {

// Dense 32^3 sector buffer [y][z][x] -> allocation mask + packed bricks.
uint64_t pack_sector(const uint8_t* dense, std::vector<uint8_t>& out, uint64_t& solid) {
    uint64_t mask = 0;
    uint8_t brick[512];
    for (uint32_t b = 0; b < 64; b++) {
        uint32_t bx = b & 3, bz = (b >> 2) & 3, by = (b >> 4) & 3;
        bool any = false;
        for (uint32_t y = 0; y < 8; y++)
            for (uint32_t z = 0; z < 8; z++)
                for (uint32_t x = 0; x < 8; x++) {
                    uint8_t v = dense[((by * 8 + y) * 32 + (bz * 8 + z)) * 32 + (bx * 8 + x)];
                    brick[x | z << 3 | y << 6] = v;
                    if (v) {
                        any = true;
                        solid++;
                    }
                }
        if (any) {
            mask |= 1ull << b;
            out.insert(out.end(), brick, brick + 512);
        }
    }
    return mask;
}

struct SectorOut {
    uint64_t mask = 0;
    std::vector<uint8_t> bricks;
    uint64_t solid = 0;
};

void assemble(vrt_scene* sc, std::vector<SectorOut>& secs) {
    uint32_t nsec = (uint32_t)secs.size();
    sc->sector_lookup.assign(nsec, -1);
    uint64_t total = 0;
    for (auto& s : secs) total += s.bricks.size();
    sc->bricks.reserve(total);
    for (uint32_t i = 0; i < nsec; i++) {
        SectorOut& s = secs[i];
        if (!s.mask) continue;
        sc->sector_lookup[i] = (int32_t)sc->sector_idx.size();
        sc->sector_idx.push_back(i);
        sc->alloc_mask.push_back(s.mask);
        sc->brick_offset.push_back(sc->bricks.size() / 512);
        sc->bricks.insert(sc->bricks.end(), s.bricks.begin(), s.bricks.end());
        sc->solid += s.solid;
        std::vector<uint8_t>().swap(s.bricks);
    }
}

vrt_scene* new_scene(uint32_t lxz, uint32_t ly) {
    auto* sc = new vrt_scene();
    sc->log2_xz = lxz;
    sc->log2_y = ly;
    sc->shift_xz = lxz - 5;
    sc->shift_y = ly - 5;
    return sc;
}

// Brick index of voxel (x, y, z) within its sector, and the voxel's index within the brick.
inline uint32_t brick_of(uint32_t x, uint32_t y, uint32_t z) { return ((x >> 3) & 3u) | ((z >> 3) & 3u) << 2 | ((y >> 3) & 3u) << 4; }
inline uint32_t voxel_of(uint32_t x, uint32_t y, uint32_t z) { return (x & 7u) | (z & 7u) << 3 | (y & 7u) << 6; }

uint32_t sector_of(const vrt_scene* sc, uint32_t x, uint32_t y, uint32_t z) {
    return (x >> 5) | (z >> 5) << sc->shift_xz | (y >> 5) << (2 * sc->shift_xz);
}

// The sector's mask in the scene (the edit overlay first, then the packed stream).
uint64_t scene_mask(const vrt_scene* sc, uint32_t si) {
    if (auto it = sc->edits.find(si); it != sc->edits.end()) return it->second.mask;
    int32_t k = sc->sector_lookup[si];
    return k < 0 ? 0ull : sc->alloc_mask[k];
}

// Brick b of sector si (allocated), from the overlay or the packed stream.
const uint8_t* scene_brick(const vrt_scene* sc, uint32_t si, uint32_t b) {
    if (auto it = sc->edits.find(si); it != sc->edits.end()) return it->second.bricks.data() + (size_t)b * 512;
    int32_t k = sc->sector_lookup[si];
    uint64_t m = sc->alloc_mask[k];
    return sc->bricks.data() + (sc->brick_offset[k] + (uint64_t)std::popcount(m & ((1ull << b) - 1))) * 512;
}

// The overlay entry of sector si, copied out of the packed stream on first use.
vrt_edit_sector& edit_sector(vrt_scene* sc, uint32_t si) {
    auto [it, fresh] = sc->edits.try_emplace(si);
    vrt_edit_sector& e = it->second;
    if (fresh) {
        e.bricks.assign(64 * 512, 0);
        int32_t k = sc->sector_lookup[si];
        if (k >= 0) {
            e.mask = sc->alloc_mask[k];
            const uint8_t* src = sc->bricks.data() + sc->brick_offset[k] * 512;
            for (uint64_t m = e.mask; m; m &= m - 1, src += 512)
                std::memcpy(e.bricks.data() + (size_t)std::countr_zero(m) * 512, src, 512);
        }
    }
    return e;
}

}  // namespace

void scene_pack_edits(vrt_scene* sc) {
    if (sc->edits.empty()) return;
    std::vector<uint32_t> idx;
    std::vector<uint64_t> mask, offset;
    std::vector<uint8_t> bricks;
    bricks.reserve(sc->bricks.size());
    std::vector<uint32_t> edited;
    for (auto& [si, e] : sc->edits) edited.push_back(si);
    std::sort(edited.begin(), edited.end());
    // merge the packed sectors with the edited ones, in increasing sector index
    size_t a = 0, b = 0;
    auto emit = [&](uint32_t si, uint64_t m, const uint8_t* packed, const vrt_edit_sector* e) {
        if (!m) return;
        idx.push_back(si);
        mask.push_back(m);
        offset.push_back(bricks.size() / 512);
        if (e) {
            for (uint64_t r = m; r; r &= r - 1) {
                const uint8_t* src = e->bricks.data() + (size_t)std::countr_zero(r) * 512;
                bricks.insert(bricks.end(), src, src + 512);
            }
        } else {
            bricks.insert(bricks.end(), packed, packed + (size_t)std::popcount(m) * 512);
        }
    };
    while (a < sc->sector_idx.size() || b < edited.size()) {
        uint32_t sa = a < sc->sector_idx.size() ? sc->sector_idx[a] : UINT32_MAX;
        uint32_t sb = b < edited.size() ? edited[b] : UINT32_MAX;
        if (sb <= sa) {
            const vrt_edit_sector& e = sc->edits.at(sb);
            emit(sb, e.mask, nullptr, &e);
            b++;
            if (sa == sb) a++;
        } else {
            emit(sa, sc->alloc_mask[a], sc->bricks.data() + sc->brick_offset[a] * 512, nullptr);
            a++;
        }
    }
    for (int32_t& k : sc->sector_lookup) k = -1;
    for (size_t k = 0; k < idx.size(); k++) sc->sector_lookup[idx[k]] = (int32_t)k;
    sc->sector_idx.swap(idx);
    sc->alloc_mask.swap(mask);
    sc->brick_offset.swap(offset);
    sc->bricks.swap(bricks);
    sc->edits.clear();
}

extern "C" {

int vrt_scene_terrain(const vrt_terrain_params* p, uint32_t threads, vrt_scene** out) {
    return guarded([&]() -> int {
        if (!p || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_terrain: null argument");
        if (p->log2_size_xz < 5 || p->log2_size_xz > 14 || p->log2_size_y < 5 || p->log2_size_y > 14)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_terrain: log2 extents must be in 5..14");
        if (p->octaves == 0 || p->band == 0) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_terrain: octaves and band must be > 0");
        vrt_scene* sc = new_scene(p->log2_size_xz, p->log2_size_y);
        uint32_t sxz = 1u << sc->shift_xz, sy = 1u << sc->shift_y;
        uint32_t nsec = sxz * sxz * sy;
        std::vector<SectorOut> secs(nsec);
        uint32_t seed = (uint32_t)(p->seed ^ (p->seed >> 32));
        uint32_t cave_seed = fmix32(seed ^ 0x5bd1e995u);
        parallel_for(sxz * sxz, pick_threads(threads), [&](uint32_t col) {
            uint32_t cx = col % sxz, cz = col / sxz;
            int32_t H[32 * 32];
            int32_t hmax = 0;
            for (uint32_t z = 0; z < 32; z++)
                for (uint32_t x = 0; x < 32; x++) {
                    double wx = (double)(cx * 32 + x), wz = (double)(cz * 32 + z);
                    double h = p->base + p->amp * fbm2(wx, wz, seed, p->octaves, p->frequency, p->gain);
                    int32_t hi = (int32_t)std::ceil(h);  // solid where y < h, i.e. y < ceil(h)
                    H[z * 32 + x] = hi;
                    hmax = std::max(hmax, hi);
                }
            std::vector<uint8_t> dense(32 * 32 * 32);
            for (uint32_t cy = 0; cy < sy; cy++) {
                int32_t y0 = (int32_t)(cy * 32);
                if (y0 >= hmax) break;
                std::fill(dense.begin(), dense.end(), 0);
                for (uint32_t y = 0; y < 32; y++) {
                    int32_t wy = y0 + (int32_t)y;
                    uint8_t mat = (uint8_t)(1 + ((uint32_t)wy / p->band) % 255u);
                    for (uint32_t z = 0; z < 32; z++)
                        for (uint32_t x = 0; x < 32; x++) {
                            if (wy >= H[z * 32 + x]) continue;
                            if (p->cave_threshold > 0) {
                                double n = fbm3((double)(cx * 32 + x) + 0.5, (double)wy + 0.5, (double)(cz * 32 + z) + 0.5,
                                                cave_seed, p->cave_octaves, p->cave_frequency, 0.5);
                                if (!(n > p->cave_threshold)) continue;
                            }
                            dense[(y * 32 + z) * 32 + x] = mat;
                        }
                }
                uint32_t si = cx | cz << sc->shift_xz | cy << (2 * sc->shift_xz);
                SectorOut& so = secs[si];
                so.mask = pack_sector(dense.data(), so.bricks, so.solid);
            }
        });
        assemble(sc, secs);
        *out = sc;
        return VRT_OK;
    });
}

int vrt_scene_spheres(uint32_t lxz, uint32_t ly, uint32_t count, float rmin, float rmax, uint64_t seed, uint32_t threads,
                      vrt_scene** out) {
    return guarded([&]() -> int {
        if (!out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_spheres: null argument");
        if (lxz < 5 || lxz > 14 || ly < 5 || ly > 14) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_spheres: log2 extents must be in 5..14");
        if (!(rmin > 0 && rmax >= rmin)) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_spheres: need 0 < rmin <= rmax");
        vrt_scene* sc = new_scene(lxz, ly);
        uint32_t sxz = 1u << sc->shift_xz, sy = 1u << sc->shift_y;
        double S = (double)(1u << lxz), SY = (double)(1u << ly);
        struct Sphere {
            double c[3], r;
            uint8_t mat;
        };
        std::vector<Sphere> sph(count);
        uint64_t st = seed;
        for (uint32_t i = 0; i < count; i++) {
            sph[i].c[0] = unit(st) * S;
            sph[i].c[1] = unit(st) * SY;
            sph[i].c[2] = unit(st) * S;
            sph[i].r = (double)rmin + unit(st) * ((double)rmax - (double)rmin);
            sph[i].mat = (uint8_t)(1 + splitmix64(st) % 255);
        }
        // bin spheres into the sectors their bounding boxes touch (in sphere order)
        uint32_t nsec = sxz * sxz * sy;
        std::vector<std::vector<uint32_t>> bins(nsec);
        for (uint32_t i = 0; i < count; i++) {
            int32_t lo[3], hi[3];
            uint32_t lim[3] = {sxz, sy, sxz};
            for (int a = 0; a < 3; a++) {
                lo[a] = std::max(0, (int32_t)std::floor((sph[i].c[a] - sph[i].r) / 32.0));
                hi[a] = std::min((int32_t)lim[a] - 1, (int32_t)std::floor((sph[i].c[a] + sph[i].r) / 32.0));
            }
            for (int32_t y = lo[1]; y <= hi[1]; y++)
                for (int32_t z = lo[2]; z <= hi[2]; z++)
                    for (int32_t x = lo[0]; x <= hi[0]; x++)
                        bins[(uint32_t)x | (uint32_t)z << sc->shift_xz | (uint32_t)y << (2 * sc->shift_xz)].push_back(i);
        }
        std::vector<SectorOut> secs(nsec);
        parallel_for(nsec, pick_threads(threads), [&](uint32_t si) {
            if (bins[si].empty()) return;
            uint32_t cx = si & (sxz - 1), cz = (si >> sc->shift_xz) & (sxz - 1), cy = si >> (2 * sc->shift_xz);
            std::vector<uint8_t> dense(32 * 32 * 32, 0);
            double ox = cx * 32.0, oy = cy * 32.0, oz = cz * 32.0;
            for (uint32_t k : bins[si]) {
                const Sphere& s = sph[k];
                double r2 = s.r * s.r;
                for (int32_t y = 0; y < 32; y++) {
                    double dy = oy + y + 0.5 - s.c[1];
                    if (dy * dy > r2) continue;
                    for (int32_t z = 0; z < 32; z++) {
                        double dz = oz + z + 0.5 - s.c[2];
                        double rem = r2 - dy * dy - dz * dz;
                        if (rem < 0) continue;
                        // voxel centres x + 0.5 with |x + 0.5 - c| <= sqrt(rem)
                        double half = std::sqrt(rem);
                        double xl = std::ceil(s.c[0] - half - 0.5 - ox), xh = std::floor(s.c[0] + half - 0.5 - ox);
                        int32_t x0 = (int32_t)std::max(0.0, xl), x1 = (int32_t)std::min(31.0, xh);
                        for (int32_t x = x0; x <= x1; x++) dense[(y * 32 + z) * 32 + x] = s.mat;
                    }
                }
            }
            SectorOut& so = secs[si];
            so.mask = pack_sector(dense.data(), so.bricks, so.solid);
        });
        std::vector<std::vector<uint32_t>>().swap(bins);
        assemble(sc, secs);
        *out = sc;
        return VRT_OK;
    });
}

int vrt_scene_from_dense(uint32_t lxz, uint32_t ly, const uint8_t* voxels, vrt_scene** out) {
    return guarded([&]() -> int {
        if (!voxels || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_from_dense: null argument");
        if (lxz < 5 || lxz > 12 || ly < 5 || ly > 12) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_from_dense: log2 extents must be in 5..12");
        vrt_scene* sc = new_scene(lxz, ly);
        uint32_t sxz = 1u << sc->shift_xz, sy = 1u << sc->shift_y;
        size_t SX = (size_t)1 << lxz;
        std::vector<SectorOut> secs((size_t)sxz * sxz * sy);
        std::vector<uint8_t> dense(32 * 32 * 32);
        for (uint32_t si = 0; si < secs.size(); si++) {
            uint32_t cx = si & (sxz - 1), cz = (si >> sc->shift_xz) & (sxz - 1), cy = si >> (2 * sc->shift_xz);
            for (uint32_t y = 0; y < 32; y++)
                for (uint32_t z = 0; z < 32; z++)
                    std::memcpy(&dense[(y * 32 + z) * 32], voxels + ((size_t)(cy * 32 + y) * SX + (cz * 32 + z)) * SX + cx * 32, 32);
            secs[si].mask = pack_sector(dense.data(), secs[si].bricks, secs[si].solid);
        }
        assemble(sc, secs);
        *out = sc;
        return VRT_OK;
    });
}

int vrt_scene_get(const vrt_scene* sc, vrt_sector_stream* out) {
    return guarded([&]() -> int {
        if (!sc || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_get: null argument");
        scene_pack_edits(const_cast<vrt_scene*>(sc));
        out->log2_size_xz = sc->log2_xz;
        out->log2_size_y = sc->log2_y;
        out->nsec = (uint32_t)sc->sector_idx.size();
        out->sector_idx = sc->sector_idx.data();
        out->alloc_mask = sc->alloc_mask.data();
        out->bricks = sc->bricks.data();
        out->nbricks = sc->bricks.size() / 512;
        out->solid_voxels = sc->solid;
        return VRT_OK;
    });
}

uint8_t vrt_scene_voxel(const vrt_scene* sc, int32_t x, int32_t y, int32_t z) {
    if (!sc) return 0;
    if (!((uint32_t)(x | z) < (1u << sc->log2_xz) && (uint32_t)y < (1u << sc->log2_y))) return 0;
    uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
    uint32_t si = sector_of(sc, ux, uy, uz), b = brick_of(ux, uy, uz);
    if (!((scene_mask(sc, si) >> b) & 1)) return 0;
    return scene_brick(sc, si, b)[voxel_of(ux, uy, uz)];
}

int vrt_scene_set_voxel(vrt_scene* sc, int32_t x, int32_t y, int32_t z, uint8_t value) {
    return guarded([&]() -> int {
        if (!sc) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_set_voxel: null scene");
        // outside the scene: ignored, as VoxelMap::Set ignores positions outside its sector grid
        if (!((uint32_t)(x | z) < (1u << sc->log2_xz) && (uint32_t)y < (1u << sc->log2_y))) return VRT_OK;
        uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
        uint32_t si = sector_of(sc, ux, uy, uz), b = brick_of(ux, uy, uz);
        vrt_edit_sector& e = edit_sector(sc, si);
        e.mask |= 1ull << b;  // GetBrick(create = true): a new brick starts empty
        uint8_t& v = e.bricks[(size_t)b * 512 + voxel_of(ux, uy, uz)];
        sc->solid += (value != 0) - (v != 0);
        v = value;
        sc->dirty[si] |= 1ull << b;  // markAsDirty (VoxelMap.cpp:118-121)
        return VRT_OK;
    });
}

int vrt_scene_set_voxels(vrt_scene* sc, uint64_t n, const int32_t* xyz, const uint8_t* values) {
    return guarded([&]() -> int {
        if (!sc) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_set_voxels: null scene");
        if (n && (!xyz || !values)) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_set_voxels: null argument");
        // the writes in order, each as vrt_scene_set_voxel; consecutive writes to one sector reuse
        // its overlay entry
        uint32_t last_si = UINT32_MAX;
        vrt_edit_sector* e = nullptr;
        uint64_t* dirty = nullptr;
        for (uint64_t k = 0; k < n; k++) {
            int32_t x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
            if (!((uint32_t)(x | z) < (1u << sc->log2_xz) && (uint32_t)y < (1u << sc->log2_y))) continue;
            uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
            uint32_t si = sector_of(sc, ux, uy, uz), b = brick_of(ux, uy, uz);
            if (si != last_si) {
                e = &edit_sector(sc, si);
                dirty = &sc->dirty[si];
                last_si = si;
            }
            e->mask |= 1ull << b;
            uint8_t& v = e->bricks[(size_t)b * 512 + voxel_of(ux, uy, uz)];
            sc->solid += (values[k] != 0) - (v != 0);
            v = values[k];
            *dirty |= 1ull << b;
        }
        return VRT_OK;
    });
}

int vrt_scene_mark_all_dirty(vrt_scene* sc) {
    return guarded([&]() -> int {
        if (!sc) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_mark_all_dirty: null scene");
        for (size_t k = 0; k < sc->sector_idx.size(); k++)
            if (!sc->edits.count(sc->sector_idx[k])) sc->dirty[sc->sector_idx[k]] = sc->alloc_mask[k];
        for (auto& [si, e] : sc->edits)
            if (e.mask) sc->dirty[si] = e.mask;
        return VRT_OK;
    });
}

int vrt_scene_delete_empty_bricks(vrt_scene* sc, uint64_t* deleted) {
    return guarded([&]() -> int {
        if (!sc) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_delete_empty_bricks: null scene");
        uint64_t count = 0;
        auto empty_brick = [](const uint8_t* p) {
            for (uint32_t i = 0; i < 512; i++)
                if (p[i]) return false;
            return true;
        };
        // the packed sectors that hold an empty brick move into the overlay first
        std::vector<uint32_t> todo;
        for (size_t k = 0; k < sc->sector_idx.size(); k++) {
            uint32_t si = sc->sector_idx[k];
            if (sc->edits.count(si)) continue;
            const uint8_t* src = sc->bricks.data() + sc->brick_offset[k] * 512;
            for (uint64_t m = sc->alloc_mask[k]; m; m &= m - 1, src += 512)
                if (empty_brick(src)) {
                    todo.push_back(si);
                    break;
                }
        }
        for (uint32_t si : todo) edit_sector(sc, si);
        for (auto& [si, e] : sc->edits) {
            uint64_t gone = 0;
            for (uint64_t m = e.mask; m; m &= m - 1) {
                uint32_t b = (uint32_t)std::countr_zero(m);
                if (empty_brick(e.bricks.data() + (size_t)b * 512)) gone |= 1ull << b;
            }
            if (!gone) continue;
            e.mask &= ~gone;
            sc->dirty[si] |= gone;  // the reference marks the emptied bricks dirty before collecting them
            count += (uint64_t)std::popcount(gone);
        }
        if (deleted) *deleted = count;
        return VRT_OK;
    });
}

uint32_t vrt_scene_dirty_sectors(const vrt_scene* sc) { return sc ? (uint32_t)sc->dirty.size() : 0u; }

int vrt_world_sync_scene(vrt_world* world, vrt_scene* sc, void* stream) {
    return guarded([&]() -> int {
        if (!world || !sc) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync_scene: null argument");
        vrt_world_info info;
        int rc = vrt_world_get_info(world, &info);
        if (rc != VRT_OK) return rc;
        if (info.size_xz != (1u << sc->log2_xz) || info.size_y != (1u << sc->log2_y))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_sync_scene: the world's extents differ from the scene's");
        // SyncBuffers (GpuRenderer.cpp:51-79): each dirty sector's mask now (0 for a sector the
        // scene no longer holds) and its dirty bricks that are still allocated
        std::vector<uint32_t> idx;
        std::vector<uint64_t> alloc, dirty;
        std::vector<uint8_t> bytes;
        for (auto& [si, d] : sc->dirty) {
            uint64_t m = scene_mask(sc, si), dm = d & m;
            idx.push_back(si);
            alloc.push_back(m);
            dirty.push_back(dm);
            for (uint64_t r = dm; r; r &= r - 1) {
                const uint8_t* src = scene_brick(sc, si, (uint32_t)std::countr_zero(r));
                bytes.insert(bytes.end(), src, src + 512);
            }
        }
        if (!idx.empty()) {
            rc = vrt_world_sync_bricks(world, (uint32_t)idx.size(), idx.data(), alloc.data(), dirty.data(),
                                       bytes.data(), stream);
            if (rc != VRT_OK) return rc;  // the dirty set is kept for the next attempt
            sc->dirty.clear();
        }
        // the scene's palette goes up with every sync, as VoxelMap::Palette does in SyncBuffers
        // (GpuRenderer.cpp:100-102); a scene without one leaves the world's palette as it is
        if (!sc->has_palette) return VRT_OK;
        return vrt_world_set_palette(world, sc->palette, stream);
    });
}

void vrt_scene_destroy(vrt_scene* sc) { delete sc; }

int vrt_world_from_scene(const vrt_scene* sc, void* stream, vrt_world** out) {
    if (!sc || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_world_from_scene: null argument");
    scene_pack_edits(const_cast<vrt_scene*>(sc));
    vrt_world_desc desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.log2_size_xz = sc->log2_xz;
    desc.log2_size_y = sc->log2_y;
    desc.initial_slots = (uint32_t)(sc->bricks.size() / 512);
    int rc = vrt_world_create(&desc, stream, out);
    if (rc != VRT_OK) return rc;
    rc = vrt_world_sync(*out, (uint32_t)sc->sector_idx.size(), sc->sector_idx.data(), sc->alloc_mask.data(),
                        sc->bricks.data(), stream);
    if (rc != VRT_OK) {
        std::string msg = vrt_last_error();
        vrt_world_destroy(*out);
        *out = nullptr;
        return fail(rc, msg);
    }
    uint64_t pal[256];
    if (sc->has_palette)
        std::memcpy(pal, sc->palette, sizeof(pal));
    else
        for (uint32_t i = 0; i < 256; i++) pal[i] = (uint64_t)i * 0x01010101ull;  // placeholder colours by id
    return vrt_world_set_palette(*out, pal, stream);
}

int vrt_camera_look(const double pos[3], double yaw_deg, double pitch_deg, double fov_deg, uint32_t width,
                    uint32_t height, uint32_t tile_w, uint32_t tile_h, vrt_camera* out) {
    return guarded([&]() -> int {
        if (!pos || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_look: null argument");
        if (!width || !height || !tile_w || !tile_h || width % tile_w || height % tile_h)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_look: width/height must be positive multiples of the tile");
        const double rad = 3.14159265358979323846 / 180.0;
        double yaw = yaw_deg * rad, pitch = pitch_deg * rad;
        double f[3] = {std::sin(yaw) * std::cos(pitch), std::sin(pitch), std::cos(yaw) * std::cos(pitch)};
        // horizontal right vector and up = forward x right (+y at pitch 0)
        double r[3] = {std::cos(yaw), 0.0, -std::sin(yaw)};
        double u[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
        double th = std::tan(fov_deg * rad * 0.5), aspect = (double)width / (double)height;
        for (int a = 0; a < 3; a++) {
            out->world_origin[a] = (int32_t)std::floor(pos[a]);
            out->eye[a] = (float)(pos[a] - std::floor(pos[a]));
            out->forward[a] = (float)f[a];
            out->right[a] = (float)(r[a] * th * aspect);
            out->up[a] = (float)(u[a] * th);
        }
        out->width = width;
        out->height = height;
        out->tile_w = tile_w;
        out->tile_h = tile_h;
        return VRT_OK;
    });
}

int vrt_camera_rays_host(const vrt_camera* cam, float* const out[6]) {
    return guarded([&]() -> int {
        if (!cam || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_rays_host: null argument");
        if (!cam->tile_w || !cam->tile_h || cam->width % cam->tile_w || cam->height % cam->tile_h)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_camera_rays_host: size must be a multiple of the tile");
        uint32_t n = cam->width * cam->height;
        float inv_w = 1.0f / (float)cam->width, inv_h = 1.0f / (float)cam->height;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t px, py;
            tile_pixel(i, cam->width, cam->tile_w, cam->tile_h, px, py);
            float d[3];
            camera_ray(cam->forward, cam->right, cam->up, inv_w, inv_h, px, py, d);
            for (int a = 0; a < 3; a++) {
                out[a][i] = cam->eye[a];
                out[3 + a][i] = d[a];
            }
        }
        return VRT_OK;
    });
}

int vrt_incoherent_rays_host(uint32_t n, uint32_t size_xz, uint32_t size_y, uint64_t seed, float* const out[6]) {
    return guarded([&]() -> int {
        if (!out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_incoherent_rays_host: null argument");
        uint64_t st = seed;
        double ext[3] = {(double)size_xz, (double)size_y, (double)size_xz};
        for (uint32_t i = 0; i < n; i++) {
            for (int a = 0; a < 3; a++) {
                float o = (float)(unit(st) * ext[a]);
                if (o >= (float)ext[a]) o = std::nextafter((float)ext[a], 0.0f);
                out[a][i] = o;
            }
            // uniform direction on the sphere: z = 2u - 1, phi = 2 pi v
            double z = 2.0 * unit(st) - 1.0, phi = 2.0 * 3.14159265358979323846 * unit(st);
            double s = std::sqrt(std::max(0.0, 1.0 - z * z));
            out[3][i] = nudge_zero((float)(s * std::cos(phi)));
            out[4][i] = nudge_zero((float)z);
            out[5][i] = nudge_zero((float)(s * std::sin(phi)));
        }
        return VRT_OK;
    });
}

int vrt_bounce_host(uint32_t n, const vrt_hit* rec, const float* pos, uint32_t width, uint32_t seed, float* const out[6],
                    uint32_t* src, uint32_t* count) {
    return guarded([&]() -> int {
        if (!rec || !pos || !out || !src || !count || !width) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_bounce_host: bad argument");
        uint32_t k = 0;
        for (uint32_t i = 0; i < n; i++) {
            if ((rec[i].info & 0xFFu) == 0) continue;
            float p[3] = {pos[i], pos[n + i], pos[2 * (size_t)n + i]};
            float o[3], d[3];
            bounce_ray(rec[i].info, p, i, width, seed, o, d);
            for (int a = 0; a < 3; a++) {
                out[a][k] = o[a];
                out[3 + a][k] = d[a];
            }
            src[k] = i;
            k++;
        }
        *count = k;
        return VRT_OK;
    });
}

}  // extern "C"
