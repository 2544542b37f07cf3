/*
 * simd.c - AVX-512 (16 lanes) and AVX2 (8 lanes) instantiations of the
 * reference's vector traversal, and the multi-threaded driver that mirrors
 * the row-parallel loop of CpuRenderer::RenderFrame (CpuRenderer.cpp:455-462).
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline and a second checker.
 * Lane semantics follow src/LibGlimpsw/SwRast/SIMD_AVX512.h and SIMD_AVX2.h.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ============================ AVX-512, 16 lanes ============================ */
#define LANES 16
#define FN(name) name##_avx512
#define TARGET __attribute__((target("avx512f,avx512bw,avx512dq,avx512vl,fma")))
#define VF __m512
#define VI __m512i
#define VM __mmask16
#define VF_SET1(x) _mm512_set1_ps(x)
#define VI_SET1(x) _mm512_set1_epi32(x)
#define VF_LOADN(p, n) _mm512_maskz_loadu_ps(M_FIRSTN(n), (p))
#define VF_STORE(p, v) _mm512_storeu_ps((p), (v))
#define VI_STORE(p, v) _mm512_storeu_si512((void*)(p), (v))
#define VF_ADD _mm512_add_ps
#define VF_SUB _mm512_sub_ps
#define VF_MUL _mm512_mul_ps
#define VF_DIV _mm512_div_ps
#define VF_MIN _mm512_min_ps
#define VF_FMA _mm512_fmadd_ps
#define VF_FLOOR2I(x) _mm512_cvt_roundps_epi32((x), _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC)
#define VI_CVTF _mm512_cvtepi32_ps
#define VF_BLEND(m, a, b) _mm512_mask_blend_ps((m), (a), (b))
#define VI_BLEND(m, a, b) _mm512_mask_blend_epi32((m), (a), (b))
#define VI_ADD _mm512_add_epi32
#define VI_SUB _mm512_sub_epi32
#define VI_AND _mm512_and_si512
#define VI_OR _mm512_or_si512
#define VI_ANDNOT _mm512_andnot_si512
#define VI_SRAI(a, n) _mm512_srai_epi32((a), (n))
#define VI_SLLI(a, n) _mm512_slli_epi32((a), (n))
#define VI_SRAV _mm512_srav_epi32
#define VI_SLLV _mm512_sllv_epi32
#define VI_GATHER(src, base, idx, m, scale) _mm512_mask_i32gather_epi32((src), (m), (idx), (base), (scale))
#define M_FIRSTN(n) ((__mmask16)((n) >= 16 ? 0xFFFFu : ((1u << (n)) - 1u)))
#define M_NONE() ((__mmask16)0)
#define M_FLT(a, b) _mm512_cmp_ps_mask((a), (b), _CMP_LT_OQ)
#define M_ULT(a, b) _mm512_cmplt_epu32_mask((a), (b))
#define M_ILT(a, b) _mm512_cmplt_epi32_mask((a), (b))
#define M_NEZ(a) _mm512_test_epi32_mask((a), (a))
#define M_EQZ(a) _mm512_testn_epi32_mask((a), (a))
#define M_AND(a, b) ((__mmask16)((a) & (b)))
#define M_OR(a, b) ((__mmask16)((a) | (b)))
#define M_ANDNOT(a, b) ((__mmask16)(~(a) & (b)))
#define M_ANY(m) ((m) != 0)
#define M_BITS(m) ((uint32_t)(m))
#include "simd_body.inc"
#undef LANES
#undef FN
#undef TARGET
#undef VF
#undef VI
#undef VM
#undef VF_SET1
#undef VI_SET1
#undef VF_LOADN
#undef VF_STORE
#undef VI_STORE
#undef VF_ADD
#undef VF_SUB
#undef VF_MUL
#undef VF_DIV
#undef VF_MIN
#undef VF_FMA
#undef VF_FLOOR2I
#undef VI_CVTF
#undef VF_BLEND
#undef VI_BLEND
#undef VI_ADD
#undef VI_SUB
#undef VI_AND
#undef VI_OR
#undef VI_ANDNOT
#undef VI_SRAI
#undef VI_SLLI
#undef VI_SRAV
#undef VI_SLLV
#undef VI_GATHER
#undef M_FIRSTN
#undef M_NONE
#undef M_FLT
#undef M_ULT
#undef M_ILT
#undef M_NEZ
#undef M_EQZ
#undef M_AND
#undef M_OR
#undef M_ANDNOT
#undef M_ANY
#undef M_BITS

/* ============================= AVX2, 8 lanes ============================== */
#define LANES 8
#define FN(name) name##_avx2
#define TARGET __attribute__((target("avx2,fma")))
#define VF __m256
#define VI __m256i
#define VM __m256i
TARGET static inline __m256i m8_firstn(uint32_t n) {
    __m256i lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
    return _mm256_cmpgt_epi32(_mm256_set1_epi32((int)n), lane);
}
TARGET static inline __m256i m8_ult(__m256i a, __m256i b) {
    __m256i s = _mm256_set1_epi32((int)0x80000000u);
    return _mm256_cmpgt_epi32(_mm256_xor_si256(b, s), _mm256_xor_si256(a, s));
}
TARGET static inline __m256i m8_floor2i(__m256 x) { return _mm256_cvtps_epi32(_mm256_floor_ps(x)); }
#define VF_SET1(x) _mm256_set1_ps(x)
#define VI_SET1(x) _mm256_set1_epi32(x)
#define VF_LOADN(p, n) _mm256_maskload_ps((p), m8_firstn(n))
#define VF_STORE(p, v) _mm256_storeu_ps((p), (v))
#define VI_STORE(p, v) _mm256_storeu_si256((__m256i*)(p), (v))
#define VF_ADD _mm256_add_ps
#define VF_SUB _mm256_sub_ps
#define VF_MUL _mm256_mul_ps
#define VF_DIV _mm256_div_ps
#define VF_MIN _mm256_min_ps
#define VF_FMA _mm256_fmadd_ps
#define VF_FLOOR2I(x) m8_floor2i(x)
#define VI_CVTF _mm256_cvtepi32_ps
#define VF_BLEND(m, a, b) _mm256_blendv_ps((a), (b), _mm256_castsi256_ps(m))
#define VI_BLEND(m, a, b) _mm256_blendv_epi8((a), (b), (m))
#define VI_ADD _mm256_add_epi32
#define VI_SUB _mm256_sub_epi32
#define VI_AND _mm256_and_si256
#define VI_OR _mm256_or_si256
#define VI_ANDNOT _mm256_andnot_si256
#define VI_SRAI(a, n) _mm256_srai_epi32((a), (n))
#define VI_SLLI(a, n) _mm256_slli_epi32((a), (n))
#define VI_SRAV _mm256_srav_epi32
#define VI_SLLV _mm256_sllv_epi32
#define VI_GATHER(src, base, idx, m, scale) _mm256_mask_i32gather_epi32((src), (base), (idx), (m), (scale))
#define M_FIRSTN(n) m8_firstn(n)
#define M_NONE() _mm256_setzero_si256()
#define M_FLT(a, b) _mm256_castps_si256(_mm256_cmp_ps((a), (b), _CMP_LT_OQ))
#define M_ULT(a, b) m8_ult((a), (b))
#define M_ILT(a, b) _mm256_cmpgt_epi32((b), (a))
#define M_NEZ(a) _mm256_xor_si256(_mm256_cmpeq_epi32((a), _mm256_setzero_si256()), _mm256_set1_epi32(-1))
#define M_EQZ(a) _mm256_cmpeq_epi32((a), _mm256_setzero_si256())
#define M_AND(a, b) _mm256_and_si256((a), (b))
#define M_OR(a, b) _mm256_or_si256((a), (b))
#define M_ANDNOT(a, b) _mm256_andnot_si256((a), (b))
#define M_ANY(m) (_mm256_movemask_epi8(m) != 0)
#define M_BITS(m) ((uint32_t)_mm256_movemask_ps(_mm256_castsi256_ps(m)))
#include "simd_body.inc"

/* ================================ driver ================================== */

int oracle_simd_width(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
        __builtin_cpu_supports("avx512dq") && __builtin_cpu_supports("avx512vl"))
        return 16;
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return 8;
    return 0;
}

int oracle_num_cores(void) {
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

typedef struct job {
    const oracle_world* w;
    const float* const* o;
    const float* const* d;
    const int32_t* W;
    uint32_t n, max_iters, lanes;
    float* t;
    oracle_hit* rec;
    float* const* pos;
    atomic_uint next;   /* next block of 64 vectors to claim */
} job;

static void* worker(void* arg) {
    job* j = (job*)arg;
    const uint32_t block = j->lanes * 64;
    for (;;) {
        uint32_t b = atomic_fetch_add(&j->next, 1);
        uint64_t start = (uint64_t)b * block;
        if (start >= j->n) break;
        uint32_t end = (uint32_t)(start + block < j->n ? start + block : j->n);
        for (uint32_t base = (uint32_t)start; base < end; base += j->lanes) {
            uint32_t cnt = end - base < j->lanes ? end - base : j->lanes;
            if (j->lanes == 16)
                cast_block_avx512(j->w, j->o, j->d, j->W, base, cnt, j->max_iters, j->t, j->rec, j->pos);
            else
                cast_block_avx2(j->w, j->o, j->d, j->W, base, cnt, j->max_iters, j->t, j->rec, j->pos);
        }
    }
    return NULL;
}

int oracle_cast_simd(const oracle_world* w, const float* const o[3], const float* const d[3],
                     const int32_t world_origin[3], uint32_t n, uint32_t max_iters,
                     uint32_t threads, float* t, oracle_hit* rec, float* const pos[3]) {
    int lanes = oracle_simd_width();
    if (lanes == 0) return 0;
    if (threads == 0) threads = (uint32_t)oracle_num_cores();
    job j;
    j.w = w;
    j.o = o;
    j.d = d;
    j.W = world_origin;
    j.n = n;
    j.max_iters = max_iters;
    j.lanes = (uint32_t)lanes;
    j.t = t;
    j.rec = rec;
    j.pos = pos;
    atomic_init(&j.next, 0);
    if (threads == 1) {
        worker(&j);
        return lanes;
    }
    pthread_t* tid = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    uint32_t started = 0;
    for (uint32_t i = 0; i < threads; i++)
        if (pthread_create(&tid[i], NULL, worker, &j) == 0) started++;
        else break;
    if (started == 0) worker(&j);
    for (uint32_t i = 0; i < started; i++) pthread_join(tid[i], NULL);
    free(tid);
    return lanes;
}
