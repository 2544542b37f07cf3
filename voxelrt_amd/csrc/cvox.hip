// cvox.hip - the reference's world file ("cvox 0004", VoxelMap::Deserialize /
// Serialize, src/VoxelRT/VoxelMap.cpp:207-274) as host-side scene I/O.
//
// Layout (little endian):
//   u64 magic 0x0000000478766f63 ("cvox", version 4)
//   u32 number of sectors
//   blob(palette): 256 Material records of 16 bytes (VoxelMap.h:22-26:
//        u8 colour[3], u8 metal fuzziness, f32 emission, 8 bytes padding)
//   packs until every sector is read:
//        u32 decompressed size, blob(records), records = per sector
//        u32 world sector index (WorldSectorIndexer<12,8,signed>, VoxelMap.h:100),
//        u64 allocation mask, popcount(mask) bricks of 512 bytes in increasing index
// where blob(x) = u32 byte count + one zstd frame of x (BinaryIO.cpp:8-60).
// zstd (libzstd.so.1) is opened at run time, so the library loads without it.
#include <dlfcn.h>

#include <algorithm>
#include <bit>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vrt_internal.h"
#include "vrt_scene.h"
#include "vrt_scene_impl.h"

namespace {

using namespace vrt;

constexpr uint64_t kMagic = 0x0000000478766f63ull;
constexpr size_t kPackBytes = 16u << 20;  // a pack is flushed once it reaches this size

// The part of zstd's public streaming API used here (zstd.h, v1.4).
struct ZInBuf {
    const void* src;
    size_t size, pos;
};
struct ZOutBuf {
    void* dst;
    size_t size, pos;
};
struct Zstd {
    void* h = nullptr;
    void* (*createDCtx)() = nullptr;
    size_t (*freeDCtx)(void*) = nullptr;
    size_t (*decompressStream)(void*, ZOutBuf*, ZInBuf*) = nullptr;
    void* (*createCCtx)() = nullptr;
    size_t (*freeCCtx)(void*) = nullptr;
    size_t (*setParameter)(void*, int, int) = nullptr;
    size_t (*compressStream2)(void*, ZOutBuf*, ZInBuf*, int) = nullptr;
    unsigned (*isError)(size_t) = nullptr;
    const char* (*errorName)(size_t) = nullptr;
    std::string error;
};
constexpr int kZstdCompressionLevel = 100, kZstdChecksumFlag = 201, kZstdEndFrame = 2, kZstdDefaultLevel = 3;

const Zstd& zstd() {
    static Zstd z;
    static std::once_flag once;
    std::call_once(once, [] {
        z.h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!z.h) {
            z.error = std::string("libzstd.so.1 not found: ") + dlerror();
            return;
        }
        bool ok = true;
        auto sym = [&](auto& fn, const char* name) {
            fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(z.h, name));
            ok = ok && fn != nullptr;
        };
        sym(z.createDCtx, "ZSTD_createDCtx");
        sym(z.freeDCtx, "ZSTD_freeDCtx");
        sym(z.decompressStream, "ZSTD_decompressStream");
        sym(z.createCCtx, "ZSTD_createCCtx");
        sym(z.freeCCtx, "ZSTD_freeCCtx");
        sym(z.setParameter, "ZSTD_CCtx_setParameter");
        sym(z.compressStream2, "ZSTD_compressStream2");
        sym(z.isError, "ZSTD_isError");
        sym(z.errorName, "ZSTD_getErrorName");
        if (!ok) z.error = "libzstd.so.1 lacks the streaming API";
    });
    return z;
}

struct Reader {
    FILE* f;
    std::string err;
    bool read(void* dst, size_t n) {
        if (n && std::fread(dst, 1, n, f) != n) {
            err = "end of file";
            return false;
        }
        return true;
    }
    template <typename T>
    bool get(T& v) { return read(&v, sizeof(T)); }
    // u32 byte count + one zstd frame that must decompress to exactly `size` bytes
    bool blob(void* dst, size_t size) {
        const Zstd& z = zstd();
        uint32_t clen;
        if (!get(clen)) return false;
        std::vector<char> in(clen);
        if (!read(in.data(), clen)) return false;
        void* dctx = z.createDCtx();
        if (!dctx) {
            err = "ZSTD_createDCtx failed";
            return false;
        }
        ZInBuf ib{in.data(), in.size(), 0};
        ZOutBuf ob{dst, size, 0};
        size_t ret = 1;
        while (ib.pos < ib.size) {
            size_t before_in = ib.pos, before_out = ob.pos;
            ret = z.decompressStream(dctx, &ob, &ib);
            if (z.isError(ret)) {
                err = std::string("zstd: ") + z.errorName(ret);
                break;
            }
            if (ib.pos == before_in && ob.pos == before_out) {
                err = "compressed blob decompresses to more than expected";
                break;
            }
        }
        z.freeDCtx(dctx);
        if (!err.empty()) return false;
        if (ob.pos != size) {
            err = "compressed blob decompresses to fewer bytes than expected";
            return false;
        }
        return true;
    }
};

struct Writer {
    FILE* f;
    std::string err;
    bool write(const void* src, size_t n) {
        if (n && std::fwrite(src, 1, n, f) != n) {
            err = "write failed";
            return false;
        }
        return true;
    }
    template <typename T>
    bool put(const T& v) { return write(&v, sizeof(T)); }
    bool blob(const void* src, size_t size) {
        const Zstd& z = zstd();
        void* cctx = z.createCCtx();
        if (!cctx) {
            err = "ZSTD_createCCtx failed";
            return false;
        }
        z.setParameter(cctx, kZstdCompressionLevel, kZstdDefaultLevel);
        z.setParameter(cctx, kZstdChecksumFlag, 1);
        std::vector<char> out;
        char buf[1 << 16];
        ZInBuf ib{src, size, 0};
        for (;;) {
            ZOutBuf ob{buf, sizeof(buf), 0};
            size_t rem = z.compressStream2(cctx, &ob, &ib, kZstdEndFrame);
            if (z.isError(rem)) {
                err = std::string("zstd: ") + z.errorName(rem);
                break;
            }
            out.insert(out.end(), buf, buf + ob.pos);
            if (rem == 0) break;
        }
        z.freeCCtx(cctx);
        if (!err.empty()) return false;
        if (out.size() > 0xFFFFFFFFull) {
            err = "compressed blob over 4 GiB";
            return false;
        }
        return put((uint32_t)out.size()) && write(out.data(), out.size());
    }
};

// IEEE binary16 of x, round to nearest even (glm::packHalf2x16's component).
uint16_t float_to_half(float x) {
    uint32_t b = std::bit_cast<uint32_t>(x);
    uint32_t sign = (b >> 16) & 0x8000u, e = (b >> 23) & 0xFFu, m = b & 0x7FFFFFu;
    if (e == 0xFF) return (uint16_t)(sign | 0x7C00u | (m ? 0x200u | (m >> 13) : 0u));
    int32_t he = (int32_t)e - 127 + 15;
    if (he >= 31) return (uint16_t)(sign | 0x7C00u);
    if (he <= 0) {
        if (he < -10) return (uint16_t)sign;
        uint32_t mant = m | 0x800000u;
        uint32_t shift = (uint32_t)(14 - he);
        uint32_t h = mant >> shift, rest = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rest > half || (rest == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)he << 10) | (m >> 13), rest = m & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h++;  // may carry into the exponent: correct
    return (uint16_t)(sign | h);
}

// Material::GetEncoded (VoxelMap.h:27-41) of one 16-byte Material record.
uint64_t encode_material(const uint8_t rec[16]) {
    float emission;
    std::memcpy(&emission, rec + 4, 4);
    uint64_t p = (uint64_t)(rec[0] >> 3) << 11 | (uint64_t)(rec[1] >> 2) << 5 | (uint64_t)(rec[2] >> 3);
    p |= (uint64_t)float_to_half(emission) << 16;
    p |= (uint64_t)rec[3] << 32;
    return p;
}

int32_t sext(uint32_t v, int bits) { return (int32_t)(v << (32 - bits)) >> (32 - bits); }

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int vrt_scene_load_cvox(const char* path, uint32_t lxz, uint32_t ly, const int32_t sector_offset[3],
                        vrt_scene** out, vrt_cvox_info* info) {
    return guarded([&]() -> int {
        if (!path || !out) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_load_cvox: null argument");
        *out = nullptr;
        if (lxz < 5 || lxz > 14 || ly < 5 || ly > 14)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_load_cvox: log2 extents must be in 5..14");
        const Zstd& z = zstd();
        if (!z.error.empty()) return fail(VRT_ERR_STATE, "vrt_scene_load_cvox: " + z.error);
        FileCloser fc{std::fopen(path, "rb")};
        if (!fc.f) return fail(VRT_ERR_INVALID_ARGUMENT, std::string("vrt_scene_load_cvox: cannot open ") + path);
        Reader r{fc.f, {}};
        auto bad = [&](const std::string& what) {
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_load_cvox: " + what + (r.err.empty() ? "" : ": " + r.err));
        };
        uint64_t magic;
        uint32_t nsec;
        if (!r.get(magic)) return bad("truncated header");
        if (magic != kMagic) return bad("not a cvox 0004 file");
        if (!r.get(nsec)) return bad("truncated header");
        uint8_t mats[256 * 16];
        if (!r.blob(mats, sizeof(mats))) return bad("palette");
        int32_t off[3] = {0, 0, 0};
        if (sector_offset) std::memcpy(off, sector_offset, sizeof(off));
        uint32_t shift_xz = lxz - 5, shift_y = ly - 5;
        int32_t nxz = 1 << shift_xz, ny = 1 << shift_y;
        struct Sec {
            uint32_t idx;
            uint64_t mask;
            size_t first;  // offset in `data`
        };
        std::vector<Sec> secs;
        std::vector<uint8_t> data;
        std::vector<uint8_t> pack;
        size_t pos = 0;
        vrt_cvox_info inf{};
        inf.sectors_in_file = nsec;
        for (uint32_t i = 0; i < nsec; i++) {
            if (pos == pack.size()) {
                uint32_t size;
                if (!r.get(size)) return bad("truncated pack header");
                pack.assign(size, 0);
                if (!r.blob(pack.data(), size)) return bad("sector pack");
                pos = 0;
            }
            if (pack.size() - pos < 12) return bad("truncated sector record");
            uint32_t widx;
            uint64_t mask;
            std::memcpy(&widx, pack.data() + pos, 4);
            std::memcpy(&mask, pack.data() + pos + 4, 8);
            pos += 12;
            size_t nb = (size_t)std::popcount(mask) * 512;
            if (pack.size() - pos < nb) return bad("truncated brick data");
            int32_t x = sext(widx & 0xFFFu, 12) - off[0], zz = sext((widx >> 12) & 0xFFFu, 12) - off[2],
                    y = sext((widx >> 24) & 0xFFu, 8) - off[1];
            if (x < 0 || x >= nxz || zz < 0 || zz >= nxz || y < 0 || y >= ny || mask == 0) {
                if (mask) inf.sectors_outside++;
                pos += nb;
                continue;
            }
            uint32_t si = (uint32_t)x | (uint32_t)zz << shift_xz | (uint32_t)y << (2 * shift_xz);
            secs.push_back({si, mask, data.size()});
            data.insert(data.end(), pack.begin() + (ptrdiff_t)pos, pack.begin() + (ptrdiff_t)(pos + nb));
            pos += nb;
        }
        std::sort(secs.begin(), secs.end(), [](const Sec& a, const Sec& b) { return a.idx < b.idx; });
        for (size_t k = 1; k < secs.size(); k++)
            if (secs[k].idx == secs[k - 1].idx) return bad("sector " + std::to_string(secs[k].idx) + " listed twice");
        auto* sc = new vrt_scene();
        sc->log2_xz = lxz;
        sc->log2_y = ly;
        sc->shift_xz = shift_xz;
        sc->shift_y = shift_y;
        sc->sector_lookup.assign((size_t)1 << (2 * shift_xz + shift_y), -1);
        sc->bricks.reserve(data.size());
        for (const Sec& s : secs) {
            sc->sector_lookup[s.idx] = (int32_t)sc->sector_idx.size();
            sc->sector_idx.push_back(s.idx);
            sc->alloc_mask.push_back(s.mask);
            sc->brick_offset.push_back(sc->bricks.size() / 512);
            size_t nb = (size_t)std::popcount(s.mask) * 512;
            const uint8_t* src = data.data() + s.first;
            for (size_t b = 0; b < nb; b++) sc->solid += src[b] != 0;
            sc->bricks.insert(sc->bricks.end(), src, src + nb);
            inf.bricks_loaded += (uint64_t)std::popcount(s.mask);
        }
        inf.sectors_loaded = (uint32_t)secs.size();
        for (uint32_t m = 0; m < 256; m++) sc->palette[m] = encode_material(mats + m * 16);
        sc->has_palette = true;
        if (info) *info = inf;
        *out = sc;
        return VRT_OK;
    });
}

int vrt_scene_save_cvox(const vrt_scene* sc, const int32_t sector_offset[3], const uint8_t* materials,
                        const char* path) {
    return guarded([&]() -> int {
        if (!sc || !path) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_save_cvox: null argument");
        scene_pack_edits(const_cast<vrt_scene*>(sc));
        const Zstd& z = zstd();
        if (!z.error.empty()) return fail(VRT_ERR_STATE, "vrt_scene_save_cvox: " + z.error);
        int32_t off[3] = {0, 0, 0};
        if (sector_offset) std::memcpy(off, sector_offset, sizeof(off));
        uint32_t nxz = 1u << sc->shift_xz;
        FileCloser fc{std::fopen(path, "wb")};
        if (!fc.f) return fail(VRT_ERR_INVALID_ARGUMENT, std::string("vrt_scene_save_cvox: cannot create ") + path);
        Writer w{fc.f, {}};
        auto bad = [&](const std::string& what) {
            return fail(VRT_ERR_STATE, "vrt_scene_save_cvox: " + what + ": " + w.err);
        };
        uint8_t mats[256 * 16] = {};
        if (materials) std::memcpy(mats, materials, sizeof(mats));
        if (!w.put(kMagic) || !w.put((uint32_t)sc->sector_idx.size()) || !w.blob(mats, sizeof(mats)))
            return bad("header");
        std::vector<uint8_t> pack;
        auto flush = [&](bool last) {
            // as VoxelMap::Serialize: a pack per 16 MiB, and a final one even when empty
            if (pack.size() < kPackBytes && !last) return true;
            bool ok = w.put((uint32_t)pack.size()) && w.blob(pack.data(), pack.size());
            pack.clear();
            return ok;
        };
        for (size_t k = 0; k < sc->sector_idx.size(); k++) {
            uint32_t si = sc->sector_idx[k];
            int32_t x = (int32_t)(si & (nxz - 1u)) + off[0], zz = (int32_t)((si >> sc->shift_xz) & (nxz - 1u)) + off[2],
                    y = (int32_t)(si >> (2 * sc->shift_xz)) + off[1];
            if (x < -2048 || x > 2047 || zz < -2048 || zz > 2047 || y < -128 || y > 127)
                return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_save_cvox: sector outside the world index range");
            uint32_t widx = ((uint32_t)x & 0xFFFu) | ((uint32_t)zz & 0xFFFu) << 12 | ((uint32_t)y & 0xFFu) << 24;
            uint64_t mask = sc->alloc_mask[k];
            size_t nb = (size_t)std::popcount(mask) * 512;
            const uint8_t* src = sc->bricks.data() + sc->brick_offset[k] * 512;
            uint8_t hdr[12];
            std::memcpy(hdr, &widx, 4);
            std::memcpy(hdr + 4, &mask, 8);
            pack.insert(pack.end(), hdr, hdr + 12);
            pack.insert(pack.end(), src, src + nb);
            if (!flush(false)) return bad("sector pack");
        }
        if (!flush(true)) return bad("sector pack");
        return VRT_OK;
    });
}

int vrt_scene_get_palette(const vrt_scene* sc, uint64_t palette[256]) {
    return guarded([&]() -> int {
        if (!sc || !palette) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_get_palette: null argument");
        if (!sc->has_palette) return fail(VRT_ERR_STATE, "vrt_scene_get_palette: the scene has no palette");
        std::memcpy(palette, sc->palette, sizeof(sc->palette));
        return VRT_OK;
    });
}

int vrt_scene_set_palette(vrt_scene* sc, const uint64_t palette[256]) {
    return guarded([&]() -> int {
        if (!sc || !palette) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_scene_set_palette: null argument");
        std::memcpy(sc->palette, palette, sizeof(sc->palette));
        sc->has_palette = true;
        return VRT_OK;
    });
}

}  // extern "C"
