// vrt_math.h - arithmetic shared by host generators and device kernels.
//
// Everything here is built from IEEE-754 operations whose results are fixed
// by the standard (add, mul, correctly-rounded div and sqrt, fmaf, floor,
// int<->float conversion), so host and device produce identical bits when the
// translation unit is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace vrt {

// x86 cvtps2dq after floor (SIMD_AVX2.h:115): NaN or out of int32 range -> INT_MIN.
__host__ __device__ inline int32_t floor2i(float x) {
    float f = floorf(x);
    return fabsf(f) < 2147483648.0f ? (int32_t)f : INT32_MIN;
}

// x86 minps (SIMD_AVX2.h:129): a < b ? a : b, i.e. the second operand on ties and NaN.
__host__ __device__ inline float min_x86(float a, float b) { return a < b ? a : b; }

// A zero direction component becomes 2^-24 so 1/d stays finite.
__host__ __device__ inline float nudge_zero(float x) { return x == 0.0f ? 5.9604644775390625e-8f : x; }

// PCG hash of RandomGen.glsl:6-20 (random_init(uvec3) / random_u32()).
__host__ __device__ inline uint32_t pcg_init(uint32_t px, uint32_t py, uint32_t seed) {
    uint32_t qx = 1103515245u * px, qy = 1103515245u * py;
    return (qx ^ (qy >> 3)) + seed * 12345u;
}
__host__ __device__ inline uint32_t pcg_next(uint32_t& state) {
    uint32_t s = state;
    state = s * 747796405u + 2891336453u;
    uint32_t word = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    return (word >> 22) ^ word;
}
__host__ __device__ inline float pcg_unit(uint32_t& state) {
    return (float)pcg_next(state) * (1.0f / 4294967296.0f);
}

// sin and cos of 2*pi*u.  The nearest quarter turn k is split off exactly, the
// remainder (|x| <= pi/4 rad) goes through Taylor polynomials of degree 9
// (sin) and 10 (cos) in Horner form with fmaf (truncation error < 4e-10),
// and the pair is rotated by k quarter turns.
__host__ __device__ inline void sincos_turns(float u, float& s_out, float& c_out) {
    float k = floorf(u * 4.0f + 0.5f);
    float r = u - k * 0.25f;
    float x = r * 6.283185307179586f;
    float x2 = x * x;
    float ps = fmaf(x2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = fmaf(x2, ps, 1.0f / 120.0f);
    ps = fmaf(x2, ps, -1.0f / 6.0f);
    float s = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = fmaf(x2, pc, -1.0f / 720.0f);
    pc = fmaf(x2, pc, 1.0f / 24.0f);
    pc = fmaf(x2, pc, -0.5f);
    float c = fmaf(x2, pc, 1.0f);
    switch ((int)k & 3) {
        case 0: s_out = s; c_out = c; break;
        case 1: s_out = c; c_out = -s; break;
        case 2: s_out = -s; c_out = -c; break;
        default: s_out = -c; c_out = s; break;
    }
}

// Normal code of a hit record: 2 bits per axis in two's complement.
__host__ __device__ inline float normal_component(uint32_t info, int axis) {
    int32_t c = (int32_t)((info >> (8 + 2 * axis)) & 3u);
    return (float)((c ^ 2) - 2);
}

// Pinhole camera ray for pixel (px, py) (vrt_camera in vrt.h).
__host__ __device__ inline void camera_ray(const float fwd[3], const float right[3], const float up[3],
                                           float inv_w, float inv_h, uint32_t px, uint32_t py, float d[3]) {
    float u = ((float)px + 0.5f) * inv_w * 2.0f - 1.0f;
    float v = 1.0f - ((float)py + 0.5f) * inv_h * 2.0f;
    float x = fwd[0] + u * right[0] + v * up[0];
    float y = fwd[1] + u * right[1] + v * up[1];
    float z = fwd[2] + u * right[2] + v * up[2];
    float il = 1.0f / sqrtf(x * x + y * y + z * z);
    d[0] = nudge_zero(x * il);
    d[1] = nudge_zero(y * il);
    d[2] = nudge_zero(z * il);
}

// Ray index -> pixel for tiles of tw x th pixels, row-major tiles, row-major inside a tile.
__host__ __device__ inline void tile_pixel(uint32_t i, uint32_t width, uint32_t tw, uint32_t th,
                                           uint32_t& px, uint32_t& py) {
    uint32_t per = tw * th;
    uint32_t tile = i / per, in = i - tile * per;
    uint32_t tiles_x = width / tw;
    uint32_t ty = tile / tiles_x, tx = tile - ty * tiles_x;
    px = tx * tw + in % tw;
    py = ty * th + in / tw;
}

// normalize(N + s) for s uniform on the unit sphere from the next two PCG
// samples of st (random_dir, RandomGen.glsl:37-48, with PCG samples in place of
// the blue-noise texture); N + s = 0 falls back to N.
__host__ __device__ inline void cosine_dir_samples(const float N[3], float u1, float u2, float dd[3]) {
    float y = u1 * 2.0f - 1.0f;
    float sn, cs;
    sincos_turns(u2, sn, cs);
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
__host__ __device__ inline void cosine_dir(const float N[3], uint32_t& st, float dd[3]) {
    float u1 = pcg_unit(st);
    float u2 = pcg_unit(st);
    cosine_dir_samples(N, u1, u2, dd);
}

// Diffuse bounce of one hit (CpuRenderer.cpp:389-392, VoxelRender.comp:63-71):
// origin = pos + N * 0.01, direction = normalize(N + s), s uniform on the unit
// sphere from two PCG samples; N + s = 0 falls back to N.
// n / d for a divisor fixed per launch (d >= 1): a multiply-high and two shifts (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI 1994: the round-up
// multiplier m = floor(2^32 (2^l - d) / d) + 1, l = ceil(log2 d), with its add-and-shift step),
// exact for every 32-bit n.
struct udiv32 {
    uint32_t d, m, s1, s2;
};
inline udiv32 make_udiv32(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) l++;
    udiv32 v;
    v.d = d;
    v.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    v.s1 = l < 1 ? l : 1;
    v.s2 = l > 1 ? l - 1 : 0;
    return v;
}
__host__ __device__ inline uint32_t udiv(const udiv32& v, uint32_t n) {
    const uint32_t t = (uint32_t)(((uint64_t)v.m * n) >> 32);
    return (t + ((n - t) >> v.s1)) >> v.s2;
}

// Hit i's two PCG samples: the first two of pcg_init(i % width, i / width, seed), with the
// division by the width as udiv.
__host__ __device__ inline void bounce_samples(uint32_t i, const udiv32& width, uint32_t seed, float& u1, float& u2) {
    const uint32_t py = udiv(width, i), px = i - py * width.d;
    uint32_t st = pcg_init(px, py, seed);
    u1 = pcg_unit(st);
    u2 = pcg_unit(st);
}

// The same from hit i's two PCG samples (u1, u2: the first two of pcg_init(i % width, i / width,
// seed)), for callers that already drew them.
__host__ __device__ inline void bounce_ray_samples(uint32_t info, const float pos[3], float u1, float u2, float o[3],
                                                   float d[3]) {
    float N[3] = {normal_component(info, 0), normal_component(info, 1), normal_component(info, 2)};
    float dd[3];
    cosine_dir_samples(N, u1, u2, dd);
    for (int a = 0; a < 3; a++) {
        o[a] = pos[a] + N[a] * 0.01f;
        d[a] = nudge_zero(dd[a]);
    }
}
__host__ __device__ inline void bounce_ray(uint32_t info, const float pos[3], uint32_t i, uint32_t width,
                                           uint32_t seed, float o[3], float d[3]) {
    uint32_t st = pcg_init(i % width, i / width, seed);
    float u1 = pcg_unit(st);
    float u2 = pcg_unit(st);
    bounce_ray_samples(info, pos, u1, u2, o, d);
}

// IEEE binary16 bits -> float (exact).
__host__ __device__ inline float half_to_float(uint32_t h) {
    uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    float mag;
    if (e == 0x1Fu) {
        uint32_t bits = 0x7F800000u | (m << 13);
        mag = __builtin_bit_cast(float, bits);
    } else if (e == 0) {
        mag = (float)m * 5.9604644775390625e-8f;  // m * 2^-24
    } else {
        uint32_t bits = ((e + 112u) << 23) | (m << 13);
        mag = __builtin_bit_cast(float, bits);
    }
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, mag) | sign);
}

// Palette entry decode (Material::GetEncoded, VoxelMap.h:27-41; VoxelMap.glsl:78-85):
// colour = (RGB565 / (31, 63, 31))^2, emission = the half in bits 16..31.
__host__ __device__ inline void material_color(uint64_t entry, float rgb[3]) {
    uint32_t x = (uint32_t)entry;
    float r = (float)((x >> 11) & 31u) * (1.0f / 31.0f);
    float g = (float)((x >> 5) & 63u) * (1.0f / 63.0f);
    float b = (float)(x & 31u) * (1.0f / 31.0f);
    rgb[0] = r * r;
    rgb[1] = g * g;
    rgb[2] = b * b;
}
__host__ __device__ inline float material_emission(uint64_t entry) { return half_to_float((uint32_t)entry >> 16); }

}  // namespace vrt
