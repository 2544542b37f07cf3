/*
 * extras.c - FP64 picking cast, brute-force DDA, PCG, bounce construction and
 * the reference gradient of the Lambertian render.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * VoxelMap::RayCast (src/VoxelRT/VoxelMap.cpp:140-170) and GetStepLevel (:125-139).
 * Sector lookup goes through WorldSectorIndexer<12,8,signed> (VoxelMap.h:100):
 * sector coordinates wrap at 4096 (x, z) and 256 (y); a sector is present when
 * its wrapped coordinates fall inside this world's grid and its mask is not 0.
 * The reference's -ffast-math build contracts tStart + pos*invDir and
 * origin + tmin*dir; they are explicit fma here.  glm::min(a, b) = b < a ? b : a.
 */
static inline int32_t d2i_floor(double x) {
    double f = floor(x);
    if (!(f >= -2147483648.0 && f < 2147483648.0)) return INT32_MIN;
    return (int32_t)f;
}
static inline double gmin(double a, double b) { return b < a ? b : a; }

static int step_level(const oracle_world* w, const int32_t p[3]) {
    uint32_t sx = (uint32_t)(p[0] >> 5) & 4095u, sy = (uint32_t)(p[1] >> 5) & 255u,
             sz = (uint32_t)(p[2] >> 5) & 4095u;
    uint32_t nxz = 1u << w->shift_xz, ny = 1u << w->shift_y;
    if (sx >= nxz || sz >= nxz || sy >= ny) return 5;
    uint32_t si = sx | sz << w->shift_xz | sy << (2 * w->shift_xz);
    uint64_t m = w->sector_masks[si];
    if (m == 0) return 5;
    uint32_t bi = ((uint32_t)(p[0] >> 3) & 3u) | ((uint32_t)(p[2] >> 3) & 3u) << 2 | ((uint32_t)(p[1] >> 3) & 3u) << 4;
    if (!((m >> bi) & 1)) return 3;
    uint32_t vi = ((uint32_t)p[0] & 7u) | ((uint32_t)p[2] & 7u) << 3 | ((uint32_t)p[1] & 7u) << 6;
    return w->storage[(size_t)si * 32768 + (size_t)bi * 512 + vi] == 0 ? 0 : -1;
}

void oracle_cast_f64(const oracle_world* w, const double origin[3], const double dir[3],
                     uint32_t max_iters, oracle_hit_f64* out) {
    double inv[3], ts[3];
    int32_t pos[3];
    for (int a = 0; a < 3; a++) {
        inv[a] = 1.0 / dir[a];
        ts[a] = ((dir[a] < 0.0 ? 0.0 : 1.0) - origin[a]) * inv[a]; /* glm::step(0, dir) */
        pos[a] = d2i_floor(origin[a]);
    }
    memset(out, 0, sizeof(*out));
    out->distance = -1.0;
    for (uint32_t i = 0; i < max_iters; i++) {
        double side[3], hp[3];
        for (int a = 0; a < 3; a++) side[a] = fma((double)pos[a], inv[a], ts[a]);
        double tmin = gmin(gmin(side[0], side[1]), side[2]) + 0.0001;
        for (int a = 0; a < 3; a++) {
            hp[a] = fma(tmin, dir[a], origin[a]);
            pos[a] = d2i_floor(hp[a]);
        }
        int k = step_level(w, pos);
        if (k < 0) {
            int sm[3];
            for (int a = 0; a < 3; a++) {
                sm[a] = tmin >= side[a];
                double sg = dir[a] > 0 ? 1.0 : (dir[a] < 0 ? -1.0 : 0.0);
                out->normal[a] = sm[a] ? (float)(-sg) : 0.0f;
                out->voxel[a] = pos[a];
            }
            float u = sm[0] ? (float)hp[1] : (float)hp[0];
            float v = sm[2] ? (float)hp[1] : (float)hp[2];
            out->uv[0] = u - floorf(u);
            out->uv[1] = v - floorf(v);
            out->distance = tmin;
            return;
        }
        int32_t m = (int32_t)((1u << k) - 1u);
        for (int a = 0; a < 3; a++) pos[a] = dir[a] < 0 ? (pos[a] & ~m) : (pos[a] | m);
    }
}

/* ---------------------------------------------------------------------------
 * Amanatides & Woo (1987) voxel walk in FP64, no LOD, from p = world_origin + o.
 */
int oracle_dda_f64(const oracle_world* w, const double o[3], const double d[3],
                   const int32_t world_origin[3], uint32_t max_steps,
                   int32_t voxel[3], double* t, double* min_gap) {
    double p[3], tmax[3], tdelta[3];
    int32_t v[3], step[3];
    for (int a = 0; a < 3; a++) {
        p[a] = (double)world_origin[a] + o[a];
        v[a] = (int32_t)floor(p[a]);
        if (d[a] > 0) {
            step[a] = 1;
            tmax[a] = ((double)v[a] + 1.0 - p[a]) / d[a];
            tdelta[a] = 1.0 / d[a];
        } else if (d[a] < 0) {
            step[a] = -1;
            tmax[a] = ((double)v[a] - p[a]) / d[a];
            tdelta[a] = -1.0 / d[a];
        } else {
            step[a] = 0;
            tmax[a] = INFINITY;
            tdelta[a] = INFINITY;
        }
    }
    double tcur = 0.0, gap = INFINITY;
    for (uint32_t s = 0; s <= max_steps; s++) {
        if (!((uint32_t)(v[0] | v[2]) < w->size_xz && (uint32_t)v[1] < w->size_y)) break;
        if (oracle_world_get_voxel(w, v[0], v[1], v[2]) != 0) {
            voxel[0] = v[0];
            voxel[1] = v[1];
            voxel[2] = v[2];
            *t = tcur;
            *min_gap = gap;
            return 1;
        }
        int a = 0;
        if (tmax[1] < tmax[a]) a = 1;
        if (tmax[2] < tmax[a]) a = 2;
        /* length of this voxel's segment, and how close the exit is to a second face */
        double seg = tmax[a] - tcur;
        if (seg < gap) gap = seg;
        for (int b = 0; b < 3; b++)
            if (b != a && tmax[b] - tmax[a] < gap) gap = tmax[b] - tmax[a];
        tcur = tmax[a];
        tmax[a] += tdelta[a];
        v[a] += step[a];
    }
    *min_gap = gap;
    return 0;
}

/* ---------------------------------------------------------------------------
 * PCG hash of RandomGen.glsl:6-20 (Jarzynski & Olano, "Hash Functions for GPU
 * Rendering", JCGT 2020): random_init(uvec3) and random_u32().
 */
uint32_t oracle_pcg_init(uint32_t px, uint32_t py, uint32_t seed) {
    uint32_t qx = 1103515245u * px, qy = 1103515245u * py;
    uint32_t n = qx ^ (qy >> 3);
    return n + seed * 12345u;
}
uint32_t oracle_pcg_next(uint32_t* state) {
    uint32_t s = *state;
    *state = s * 747796405u + 2891336453u;
    uint32_t word = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    return (word >> 22) ^ word;
}

/* sin/cos of 2*pi*u: reduce to the nearest quarter turn k (exact), map the
 * remainder r in [-1/8, 1/8] turns to x = r * 2pi and evaluate the Taylor
 * polynomials of sin (to x^9) and cos (to x^10) in Horner form with fmaf; the
 * truncation error on |x| <= pi/4 is below 4e-10.  Rotate by k quarter turns. */
void oracle_sincos_turns(float u, float* s_out, float* c_out) {
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
    int q = (int)k & 3;
    switch (q) {
        case 0: *s_out = s; *c_out = c; break;
        case 1: *s_out = c; *c_out = -s; break;
        case 2: *s_out = -s; *c_out = -c; break;
        default: *s_out = -c; *c_out = s; break;
    }
}

static inline float nudge(float x) { return x == 0.0f ? 5.9604644775390625e-8f : x; }

uint32_t oracle_bounce(uint32_t n, const float* t, const oracle_hit* rec, const float* const pos[3],
                       const float* const d[3], uint32_t width, uint32_t seed,
                       float* const out_o[3], float* const out_d[3], uint32_t* src) {
    (void)t;
    (void)d;
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t info = rec[i].info;
        if ((info & 0xFFu) == 0) continue;
        float N[3];
        for (int a = 0; a < 3; a++) {
            int32_t c = (int32_t)((info >> (8 + 2 * a)) & 3u);
            c = (c ^ 2) - 2; /* sign-extend 2 bits */
            N[a] = (float)c;
        }
        uint32_t st = oracle_pcg_init(i % width, i / width, seed);
        float u1 = (float)oracle_pcg_next(&st) * (1.0f / 4294967296.0f);
        float u2 = (float)oracle_pcg_next(&st) * (1.0f / 4294967296.0f);
        float y = u1 * 2.0f - 1.0f;
        float sn, cs;
        oracle_sincos_turns(u2, &sn, &cs);
        float sy = sqrtf(1.0f - y * y);
        float s[3] = {sn * sy, y, cs * sy};
        float v[3] = {N[0] + s[0], N[1] + s[1], N[2] + s[2]};
        float len2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        float dir[3];
        if (len2 == 0.0f) {
            dir[0] = N[0]; dir[1] = N[1]; dir[2] = N[2];
        } else {
            float il = 1.0f / sqrtf(len2);
            dir[0] = v[0] * il; dir[1] = v[1] * il; dir[2] = v[2] * il;
        }
        for (int a = 0; a < 3; a++) {
            out_o[a][k] = pos[a][i] + N[a] * 0.01f;
            out_d[a][k] = nudge(dir[a]);
        }
        src[k] = i;
        k++;
    }
    return k;
}

/* ---------------------------------------------------------------------------
 * Backward of the Lambertian render: radiance_pc = albedo[v_p][c] * w_p, so
 * dL/dalbedo[v][c] = sum over pixels p with v_p = v of dL/dradiance_pc * w_p.
 * FP64 sums; abssum[v][c] = sum of |terms|, the scale of the rounding error of
 * any fp32 evaluation of the same sum.
 */
void oracle_render_backward(uint32_t npix, const uint32_t* saved_voxel, const float* saved_weight,
                            const float* grad_image, double* grad, double* abssum) {
    for (uint32_t p = 0; p < npix; p++) {
        uint32_t v = saved_voxel[p];
        if (v == 0xFFFFFFFFu) continue;
        double wgt = (double)saved_weight[p];
        for (int c = 0; c < 3; c++) {
            double term = (double)grad_image[(size_t)p * 3 + c] * wgt;
            grad[(size_t)v * 3 + c] += term;
            abssum[(size_t)v * 3 + c] += fabs(term);
        }
    }
}

/* L = sum_p sum_c (albedo[v_p][c] * w_p - target_pc)^2 in FP64 (0 radiance on a miss). */
double oracle_render_loss(uint32_t npix, const uint32_t* saved_voxel, const float* saved_weight,
                          const double* albedo, const float* target) {
    double loss = 0.0;
    for (uint32_t p = 0; p < npix; p++) {
        uint32_t v = saved_voxel[p];
        for (int c = 0; c < 3; c++) {
            double rad = v == 0xFFFFFFFFu ? 0.0 : albedo[(size_t)v * 3 + c] * (double)saved_weight[p];
            double e = rad - (double)target[(size_t)p * 3 + c];
            loss += e * e;
        }
    }
    return loss;
}
