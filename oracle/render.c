/*
 * render.c - scalar restatement of the GPU renderer's per-pixel shading
 * (src/VoxelRT/Shaders/VoxelRender.comp:28-94) over the shader method.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 * Per pixel: rayCast of the primary ray; on a hit the palette colour
 * (getMaterialColor, VoxelMap.glsl:78-82: (RGB565 / (31,63,31))^2) and emission
 * (the half in bits 16..31, :83-85), the projected depth of pos / 16 (:43-44),
 * a coarse sun shadow ray (:54-61), then max_bounces coarse diffuse bounces
 * (:63-86), the first two testing the sun; a miss sees the sky (:87-90).  The
 * G-buffer packing follows GBuffer.glsl:15-21 (packGNormal) and the RGBA8 unorm
 * rule round(clamp(x, 0, 1) * 255).  Bounce directions are normalize(N + s), s
 * from the PCG stream random_init(uvec3(px, py, frame)) (RandomGen.glsl:6-20).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static float half_to_float(uint32_t h) {
    uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu, bits;
    float mag;
    if (e == 0x1Fu) {
        bits = 0x7F800000u | (m << 13);
        memcpy(&mag, &bits, 4);
    } else if (e == 0) {
        mag = (float)m * 5.9604644775390625e-8f;
    } else {
        bits = ((e + 112u) << 23) | (m << 13);
        memcpy(&mag, &bits, 4);
    }
    memcpy(&bits, &mag, 4);
    bits |= sign;
    memcpy(&mag, &bits, 4);
    return mag;
}

static void material_color(uint64_t entry, float rgb[3]) {
    uint32_t x = (uint32_t)entry;
    float r = (float)((x >> 11) & 31u) * (1.0f / 31.0f);
    float g = (float)((x >> 5) & 63u) * (1.0f / 63.0f);
    float b = (float)(x & 31u) * (1.0f / 31.0f);
    rgb[0] = r * r;
    rgb[1] = g * g;
    rgb[2] = b * b;
}

static void normal_of(uint32_t info, float N[3]) {
    for (int a = 0; a < 3; a++) {
        int32_t c = (int32_t)((info >> (8 + 2 * a)) & 3u);
        N[a] = (float)((c ^ 2) - 2);
    }
}

/* normalize(N + s), s uniform on the sphere from the next two PCG samples; N + s = 0 -> N. */
static void cosine_dir(const float N[3], uint32_t* st, float dd[3]) {
    float u1 = (float)oracle_pcg_next(st) * (1.0f / 4294967296.0f);
    float u2 = (float)oracle_pcg_next(st) * (1.0f / 4294967296.0f);
    float y = u1 * 2.0f - 1.0f;
    float sn, cs;
    oracle_sincos_turns(u2, &sn, &cs);
    float sy = sqrtf(1.0f - y * y);
    float v0 = N[0] + sn * sy, v1 = N[1] + y, v2 = N[2] + cs * sy;
    float len2 = v0 * v0 + v1 * v1 + v2 * v2;
    if (len2 == 0.0f) {
        dd[0] = N[0]; dd[1] = N[1]; dd[2] = N[2];
    } else {
        float il = 1.0f / sqrtf(len2);
        dd[0] = v0 * il; dd[1] = v1 * il; dd[2] = v2 * il;
    }
}

static inline float nudge(float x) { return x == 0.0f ? 5.9604644775390625e-8f : x; }
static inline uint32_t unorm8(float x) { return (uint32_t)(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f + 0.5f); }

void oracle_render_frame(const oracle_world* w, uint32_t width, uint32_t height, const float* const o[3],
                         const float* const d[3], const int32_t W[3], const float fwd[3], const float sun[3],
                         float pa, float pb, uint32_t max_bounces, uint32_t frame, uint32_t lut_on,
                         const float sky[3], uint32_t* albedo_normal, float* irradiance, float* depth) {
    uint64_t table[512];
    oracle_ray_cell_lut(table);
    const uint64_t* lut = lut_on ? table : NULL;
    const float sun_rgb[3] = {1.2f * 5.0f, 1.1f * 5.0f, 1.0f * 5.0f};
    for (uint32_t py = 0; py < height; py++)
        for (uint32_t px = 0; px < width; px++) {
            size_t i = (size_t)py * width + px;
            uint32_t st = oracle_pcg_init(px, py, frame);
            float ro[3] = {o[0][i], o[1][i], o[2][i]}, rd[3] = {d[0][i], d[1][i], d[2][i]};
            float albedo[3] = {1.0f, 1.0f, 1.0f}, irr[3], N[3] = {0.0f, 0.0f, 0.0f}, dep = -1.0f;
            float t, pos[3];
            oracle_hit rec;
            oracle_shader_ray(w, ro, rd, W, 0, lut, 256, &t, &rec, pos);
            uint32_t iters = rec.info >> 16;
            if (rec.info & 0xFFu) {
                uint64_t mat = w->palette[rec.info & 0xFFu];
                material_color(mat, albedo);
                normal_of(rec.info, N);
                float q = fwd[0] * (pos[0] / 16.0f) + fwd[1] * (pos[1] / 16.0f) + fwd[2] * (pos[2] / 16.0f);
                dep = (pa * -q + pb) / q;
                float em = half_to_float((uint32_t)mat >> 16);
                float thr[3] = {1.0f, 1.0f, 1.0f};
                for (int c = 0; c < 3; c++) irr[c] = albedo[c] * em;
                if (max_bounces == 0) {
                    for (int c = 0; c < 3; c++) irr[c] = albedo[c];
                } else {
                    float so[3], sp[3], stt;
                    oracle_hit sr;
                    for (int a = 0; a < 3; a++) so[a] = pos[a] + N[a] * 0.01f;
                    oracle_shader_ray(w, so, sun, W, 1, lut, 96, &stt, &sr, sp);
                    if (!(sr.info & 0xFFu)) {
                        for (int c = 0; c < 3; c++) irr[c] = irr[c] + sun_rgb[c];
                    } else {
                        for (int c = 0; c < 3; c++) thr[c] = thr[c] * 0.5f;
                    }
                }
                for (uint32_t b = 0; b < max_bounces; b++) {
                    float bo[3], bd[3], dd[3];
                    for (int a = 0; a < 3; a++) bo[a] = pos[a] + N[a] * 0.01f;
                    cosine_dir(N, &st, dd);
                    for (int a = 0; a < 3; a++) bd[a] = nudge(dd[a]);
                    oracle_shader_ray(w, bo, bd, W, 1, lut, 96, &t, &rec, pos);
                    iters = rec.info >> 16;
                    if (!(rec.info & 0xFFu)) {
                        for (int c = 0; c < 3; c++) irr[c] = irr[c] + thr[c] * sky[c];
                        break;
                    }
                    normal_of(rec.info, N);
                    uint64_t bm = w->palette[rec.info & 0xFFu];
                    float col[3];
                    material_color(bm, col);
                    for (int c = 0; c < 3; c++) thr[c] = thr[c] * col[c];
                    float bem = half_to_float((uint32_t)bm >> 16);
                    if (b < 2) {
                        float so[3], sp[3], stt;
                        oracle_hit sr;
                        for (int a = 0; a < 3; a++) so[a] = pos[a] + N[a] * 0.01f;
                        oracle_shader_ray(w, so, sun, W, 1, lut, 96, &stt, &sr, sp);
                        if (!(sr.info & 0xFFu)) {
                            for (int c = 0; c < 3; c++) thr[c] = thr[c] * sun_rgb[c];
                            bem = bem + 5.0f;
                        }
                    }
                    for (int c = 0; c < 3; c++) irr[c] = irr[c] + thr[c] * bem;
                }
            } else {
                for (int c = 0; c < 3; c++) irr[c] = sky[c];
            }
            uint32_t ncode = 0;
            for (int a = 0; a < 3; a++) ncode |= (uint32_t)(int32_t)(N[a] + 1.0f) << (2 * a);
            albedo_normal[i] = unorm8(albedo[0]) | unorm8(albedo[1]) << 8 | unorm8(albedo[2]) << 16 | ncode << 24;
            for (int c = 0; c < 3; c++) irradiance[i * 4 + c] = irr[c];
            irradiance[i * 4 + 3] = (float)iters;
            depth[i] = dep;
        }
}
