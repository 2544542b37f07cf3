// diff.hip - differentiable Lambertian render: forward (cast + shade, saving
// the hit voxel and shading weight of every pixel) and backward (scatter-add
// of dL/dpixel x weight into per-voxel RGB albedo gradients).
//
// No reference counterpart (the reference renders, it does not differentiate);
// the shading term is the reference's sun light: direction
// normalize(0.3, 0.9, -0.28), intensity 5 (Shaders/VoxelRender.comp:133-134)
// with a Lambert N.L factor.  The traversal stays discrete and exact (the same
// cast as vrt_cast), so radiance is linear in the albedo and
//   d radiance_pc / d albedo[v][c] = weight_p  if pixel p hit voxel v.
#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>
#include <string>

#include "traverse.h"
#include "vrt_cast_impl.h"

namespace vrt {

// Up to kFwdViews views per launch, their cameras passed in the kernel arguments: one
// launch over many small views fills the chip where one launch per view (a 512x512
// view is half a round of resident waves) leaves it draining after every view.
constexpr uint32_t kFwdViews = 16;
struct FwdViews {
    vrt_camera cam[kFwdViews];
    uint32_t block_begin[kFwdViews + 1];  // first block of 64 pixels of each view; [nv] = all blocks
    uint32_t pixel_base[kFwdViews];       // first output pixel of each view
    uint32_t nv;
};

__global__ __launch_bounds__(64) void render_fwd_kernel(WorldView w, FwdViews views, uint32_t max_iters,
                                                        float lx, float ly, float lz, float intensity,
                                                        const float* __restrict__ albedo, float* __restrict__ image,
                                                        uint32_t* __restrict__ saved_voxel,
                                                        float* __restrict__ saved_weight) {
    // the wave's view (blocks never straddle views)
    const uint32_t b = blockIdx.x;
    uint32_t k = 0;
    while (k + 1 < views.nv && b >= views.block_begin[k + 1]) k++;
    const vrt_camera& cam = views.cam[k];
    const uint32_t n = cam.width * cam.height;
    const uint32_t li = (b - views.block_begin[k]) * 64u + threadIdx.x;  // pixel within the view
    const bool valid = li < n;
    const uint32_t i = views.pixel_base[k] + li;                          // output pixel
    RayState st;
    int32_t v[3] = {0, 0, 0};
    bool tame = true;
    if (valid) {
        uint32_t px, py;
        tile_pixel(li, cam.width, cam.tile_w, cam.tile_h, px, py);
        float d[3];
        camera_ray(cam.forward, cam.right, cam.up, 1.0f / (float)cam.width, 1.0f / (float)cam.height, px, py, d);
        ray_setup(st, cam.eye[0], cam.eye[1], cam.eye[2], d[0], d[1], d[2], cam.world_origin[0], cam.world_origin[1],
                  cam.world_origin[2]);
        tame = ray_is_tame(st);
    }
    bool wave_tame = __all(tame);
    if (valid) {
        uint32_t slot = 0;
        bool hit = ray_trace_wave(w, st, max_iters, v, slot, wave_tame,
                                  origin_kind(cam.world_origin[0], cam.world_origin[1], cam.world_origin[2]));
        float wgt = 0.0f;
        uint32_t gv = 0xFFFFFFFFu;
        if (hit) {
            float t;
            int4 rec;
            ray_hit_record(w, st, v, slot, t, rec);
            float nx = normal_component((uint32_t)rec.w, 0), ny = normal_component((uint32_t)rec.w, 1),
                  nz = normal_component((uint32_t)rec.w, 2);
            float ndl = nx * lx + ny * ly + nz * lz;
            wgt = fmaxf(ndl, 0.0f) * intensity;
            uint32_t ux = (uint32_t)v[0], uy = (uint32_t)v[1], uz = (uint32_t)v[2];
            gv = slot * 512u + ((ux & 7u) | (uz & 7u) << 3 | (uy & 7u) << 6);
        }
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (gv != 0xFFFFFFFFu) {
            r = albedo[(size_t)gv * 3 + 0] * wgt;
            g = albedo[(size_t)gv * 3 + 1] * wgt;
            b = albedo[(size_t)gv * 3 + 2] * wgt;
        }
        image[(size_t)i * 3 + 0] = r;
        image[(size_t)i * 3 + 1] = g;
        image[(size_t)i * 3 + 2] = b;
        saved_voxel[i] = gv;
        saved_weight[i] = wgt;
    }
}

__device__ __forceinline__ float wave_masked_sum(float x, bool take) {
    float s = take ? x : 0.0f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Scatter-add with wave aggregation: lanes that hit the same voxel are summed
// in registers (ballot on the key of the first pending lane, shuffle
// reduction) and one lane issues the atomics; after kRounds unique keys the
// remaining lanes add directly.
__global__ __launch_bounds__(256) void render_bwd_kernel(uint32_t n, const float* __restrict__ grad_image,
                                                         const uint32_t* __restrict__ saved_voxel,
                                                         const float* __restrict__ saved_weight,
                                                         float* __restrict__ grad_albedo) {
    const int kRounds = 4;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gv = 0xFFFFFFFFu;
    float gr = 0.0f, gg = 0.0f, gb = 0.0f;
    if (i < n) {
        gv = saved_voxel[i];
        float wgt = saved_weight[i];
        if (gv != 0xFFFFFFFFu && wgt != 0.0f) {
            gr = grad_image[(size_t)i * 3 + 0] * wgt;
            gg = grad_image[(size_t)i * 3 + 1] * wgt;
            gb = grad_image[(size_t)i * 3 + 2] * wgt;
        } else {
            gv = 0xFFFFFFFFu;
        }
    }
    bool pending = gv != 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    for (int round = 0; round < kRounds; round++) {
        uint64_t todo = __ballot(pending);
        if (todo == 0) break;
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
        uint32_t key = __shfl(gv, (int)leader, 64);
        bool mine = pending && gv == key;
        float sr = wave_masked_sum(gr, mine), sg = wave_masked_sum(gg, mine), sb = wave_masked_sum(gb, mine);
        if (lane == leader) {
            atomicAdd(&grad_albedo[(size_t)key * 3 + 0], sr);
            atomicAdd(&grad_albedo[(size_t)key * 3 + 1], sg);
            atomicAdd(&grad_albedo[(size_t)key * 3 + 2], sb);
        }
        if (mine) pending = false;
    }
    if (pending) {
        atomicAdd(&grad_albedo[(size_t)gv * 3 + 0], gr);
        atomicAdd(&grad_albedo[(size_t)gv * 3 + 1], gg);
        atomicAdd(&grad_albedo[(size_t)gv * 3 + 2], gb);
    }
}

// ---------------------------------------------------------------- brick-bucketed backward

// The whole gradient written once, with no global float atomics: hit pixels are bucketed by
// brick slot (a counting sort: per-slot counts with wave-aggregated integer atomics, a
// two-level exclusive scan, a placement pass), then one wave per slot sums its pixels'
// contributions in LDS and writes the slot's 512 x 3 gradient entries, zeros for a slot no
// pixel hit.  On the 16 views of the 512^3 terrain, 2.2 M hit pixels land on 287 k voxels,
// 7.8 pixels per voxel across views, so per-pixel global atomics contend, and a dense
// gradient zeroed first is written twice.
constexpr uint32_t kScanBlock = 1024;  // counts per block of the two-level scan

// Pass 1: the rank of each contributing pixel among the pixels of its slot; counts[slot] ends
// as the slot's pixel count.  Lanes of a wave with the same slot share one atomic (each round
// matches the first pending lane's slot by ballot; a lane's rank is its position among the
// matches).  Eight rounds, unrolled with one result variable each, so a wave's atomics are all
// in flight together and the wave waits for their returns once, after the last round.
__global__ __launch_bounds__(256) void bwd_count_kernel(uint32_t n, const uint32_t* __restrict__ saved_voxel,
                                                        const float* __restrict__ saved_weight,
                                                        uint32_t* __restrict__ counts, uint32_t* __restrict__ rank) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t slot = 0xFFFFFFFFu;
    if (i < n) {
        uint32_t gv = saved_voxel[i];
        if (gv != 0xFFFFFFFFu && saved_weight[i] != 0.0f) slot = gv >> 9;
    }
    bool pending = slot != 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    uint32_t my_leader = 0;
    uint64_t my_same = 0;
    bool matched = false;
    // one round: the leader's atomic result lands in its own variable b, read only after the
    // last round
    auto round = [&](uint32_t& b) {
        uint64_t todo = __ballot(pending);
        if (todo == 0) return;  // wave-uniform
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
        uint32_t key = __shfl(slot, (int)leader, 64);
        bool mine = pending && slot == key;
        uint64_t same = __ballot(mine);
        if (lane == leader) b = atomicAdd(&counts[key], (uint32_t)__popcll(same));
        if (mine) {
            my_leader = leader;
            my_same = same;
            matched = true;
            pending = false;
        }
    };
    uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0, b5 = 0, b6 = 0, b7 = 0;
    round(b0);
    round(b1);
    round(b2);
    round(b3);
    round(b4);
    round(b5);
    round(b6);
    round(b7);
    // a lane leads at most one round and its other variables stay 0, so their OR is the base of
    // the round it led
    uint32_t led = b0 | b1 | b2 | b3 | b4 | b5 | b6 | b7;
    uint32_t r = 0xFFFFFFFFu;
    uint32_t b = __shfl(led, (int)my_leader, 64);
    if (matched)
        r = b + __builtin_amdgcn_mbcnt_hi((uint32_t)(my_same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)my_same, 0u));
    if (pending) r = atomicAdd(&counts[slot], 1u);  // slots left after the unrolled rounds
    if (i < n) rank[i] = r;
}

// Pass 2a: exclusive scan of counts inside blocks of kScanBlock (256 threads x 4), into offsets;
// block_sum[b] = the block's total.
__global__ __launch_bounds__(256) void bwd_scan_blocks_kernel(uint32_t nb, const uint32_t* __restrict__ counts,
                                                              uint32_t* __restrict__ offsets,
                                                              uint32_t* __restrict__ block_sum) {
    __shared__ uint32_t wave_total[4];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t base = blockIdx.x * kScanBlock + t * 4;
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k] = base + k < nb ? counts[base + k] : 0u;
        sum += c[k];
    }
    uint32_t incl = sum;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (uint32_t w = 0; w < wave; w++) run += wave_total[w];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < nb) offsets[base + k] = run;
        run += c[k];
    }
    if (t == 255) block_sum[blockIdx.x] = run;
}

// Pass 2b: exclusive scan of the block totals in one workgroup (runs of consecutive totals per
// thread, combined across the 16 waves).
__global__ __launch_bounds__(1024) void bwd_scan_totals_kernel(uint32_t nblk, uint32_t* __restrict__ block_sum) {
    __shared__ uint32_t wave_total[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t per = (nblk + 1023) / 1024;
    uint32_t begin = min(t * per, nblk), end = min(begin + per, nblk);
    uint32_t sum = 0;
    for (uint32_t k = begin; k < end; k++) sum += block_sum[k];
    uint32_t incl = sum;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (uint32_t w = 0; w < wave; w++) run += wave_total[w];
    for (uint32_t k = begin; k < end; k++) {
        uint32_t c = block_sum[k];
        block_sum[k] = run;
        run += c;
    }
}

// Pass 3: each contributing pixel's index at its slot's offset plus its rank.
__global__ __launch_bounds__(256) void bwd_place_kernel(uint32_t n, const uint32_t* __restrict__ saved_voxel,
                                                        const uint32_t* __restrict__ rank,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ block_sum,
                                                        uint32_t* __restrict__ list) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t r = rank[i];
    if (r == 0xFFFFFFFFu) return;
    uint32_t slot = saved_voxel[i] >> 9;
    list[offsets[slot] + block_sum[slot / kScanBlock] + r] = i;
}

// Pass 4: one wave per slot.  The slot's pixels add grad_image * weight into 512 x 3 LDS
// accumulators (LDS float atomics), then the wave writes the slot's rows [slot * 512,
// min(nvox, slot * 512 + 512)) of the gradient: every entry is written, zeros where no pixel hit.
// On the 512^3 terrain's 97 k slots this moves 598 MB of gradient and ~145 MB of gathered pixel
// data in 0.133 ms (5.6 TB/s); four slots per workgroup (0.145 ms) and a grid stride over 8192
// long-lived waves (0.156 ms) measured slower.  The gradient rows are written with non-temporal
// stores (nothing in the pass reads them back): 2% off the backward in alternating runs.
__device__ __forceinline__ void store_row4_nt(float4* dst, float4 x) {
    float* q = (float*)dst;
    __builtin_nontemporal_store(x.x, q + 0);
    __builtin_nontemporal_store(x.y, q + 1);
    __builtin_nontemporal_store(x.z, q + 2);
    __builtin_nontemporal_store(x.w, q + 3);
}

__global__ __launch_bounds__(64) void bwd_reduce_kernel(uint64_t nvox, const float* __restrict__ grad_image,
                                                        const uint32_t* __restrict__ saved_voxel,
                                                        const float* __restrict__ saved_weight,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ block_sum,
                                                        const uint32_t* __restrict__ list,
                                                        float* __restrict__ grad_albedo) {
    __shared__ float acc[512 * 3];
    const uint32_t slot = blockIdx.x, lane = threadIdx.x;
    const uint64_t row0 = (uint64_t)slot * 512;
    const uint32_t rows = (uint32_t)min<uint64_t>(512, nvox - row0);
    float* out = grad_albedo + row0 * 3;
    const uint32_t cnt = counts[slot];
    if (cnt == 0) {
        if (rows == 512) {
            float4* o4 = (float4*)out;
            for (uint32_t k = lane; k < 384; k += 64) store_row4_nt(&o4[k], make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        } else {
            for (uint32_t k = lane; k < rows * 3; k += 64) out[k] = 0.0f;
        }
        return;
    }
    for (uint32_t k = lane; k < 512 * 3; k += 64) acc[k] = 0.0f;
    __syncthreads();
    const uint32_t begin = offsets[slot] + block_sum[slot / kScanBlock];
    for (uint32_t j = lane; j < cnt; j += 64) {
        uint32_t p = list[begin + j];
        uint32_t v = saved_voxel[p] & 511u;
        float w = saved_weight[p];
        atomicAdd(&acc[v * 3 + 0], grad_image[(size_t)p * 3 + 0] * w);
        atomicAdd(&acc[v * 3 + 1], grad_image[(size_t)p * 3 + 1] * w);
        atomicAdd(&acc[v * 3 + 2], grad_image[(size_t)p * 3 + 2] * w);
    }
    __syncthreads();
    if (rows == 512) {
        float4* o4 = (float4*)out;
        const float4* a4 = (const float4*)acc;
        for (uint32_t k = lane; k < 384; k += 64) store_row4_nt(&o4[k], a4[k]);
    } else {
        for (uint32_t k = lane; k < rows * 3; k += 64) out[k] = acc[k];
    }
}

__global__ void bwd_fill_kernel(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Sparse gradient: entry p is (voxel id, grad_image[p] * weight[p]), or (0, 0) for a pixel that
// contributes nothing (a miss or zero weight); duplicates are summed by whoever coalesces.
__global__ __launch_bounds__(256) void bwd_sparse_kernel(uint32_t n, const float* __restrict__ grad_image,
                                                         const uint32_t* __restrict__ saved_voxel,
                                                         const float* __restrict__ saved_weight,
                                                         int64_t* __restrict__ index, float* __restrict__ values) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t gv = saved_voxel[i];
    float w = saved_weight[i];
    bool take = gv != 0xFFFFFFFFu && w != 0.0f;
    index[i] = take ? (int64_t)gv : 0;
    values[(size_t)i * 3 + 0] = take ? grad_image[(size_t)i * 3 + 0] * w : 0.0f;
    values[(size_t)i * 3 + 1] = take ? grad_image[(size_t)i * 3 + 1] * w : 0.0f;
    values[(size_t)i * 3 + 2] = take ? grad_image[(size_t)i * 3 + 2] * w : 0.0f;
}

// Workspace of the brick-bucketed backward: counts, offsets (one u32 per slot each), block
// totals (one per kScanBlock slots), ranks and the pixel list (one u32 per pixel each), each
// part 256-byte aligned.
struct BwdWorkspace {
    uint32_t *counts, *offsets, *block_sum, *rank, *list;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t bwd_workspace_bytes(uint32_t npixels, uint64_t nslots) {
    uint64_t nblk = (nslots + kScanBlock - 1) / kScanBlock;
    return 2 * align256(nslots * 4) + align256(nblk * 4) + 2 * align256((size_t)npixels * 4);
}

static BwdWorkspace carve_bwd(void* base, uint32_t npixels, uint64_t nslots) {
    uint64_t nblk = (nslots + kScanBlock - 1) / kScanBlock;
    char* p = (char*)base;
    BwdWorkspace ws;
    ws.counts = (uint32_t*)p; p += align256(nslots * 4);
    ws.offsets = (uint32_t*)p; p += align256(nslots * 4);
    ws.block_sum = (uint32_t*)p; p += align256(nblk * 4);
    ws.rank = (uint32_t*)p; p += align256((size_t)npixels * 4);
    ws.list = (uint32_t*)p;
    return ws;
}

}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_render_fwd(const vrt_world* world, const vrt_camera* cameras, uint32_t nviews, uint32_t max_iters,
                   const float light_dir[3], float intensity, const float* albedo, float* image,
                   uint32_t* saved_voxel, float* saved_weight, void* stream) {
    return guarded([&]() -> int {
        if (!world || (nviews && !cameras) || !light_dir)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: null argument");
        if (int rcw = check_world(world, "vrt_render_fwd"); rcw != VRT_OK) return rcw;
        if (max_iters == 0 || max_iters > 65535) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: max_iters must be 1..65535");
        if ((uint64_t)world->slot_cap * 512 >= 0xFFFFFFFFull)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: world too large for 32-bit voxel ids");
        hipStream_t s = (hipStream_t)stream;
        WorldView w = world->view();
        uint64_t total = 0;
        for (uint32_t k = 0; k < nviews; k++) {
            const vrt_camera& c = cameras[k];
            if (!c.width || !c.height || !c.tile_w || !c.tile_h || c.width % c.tile_w || c.height % c.tile_h)
                return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: camera size must be a multiple of its tile");
            total += (uint64_t)c.width * c.height;
        }
        // each launch's 64-lane blocks (at most one partial block per view) and every
        // output pixel index must fit in 32 bits
        if (total > 0xFFFFF000ull) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: more than 2^32 - 4096 pixels");
        if (total && (!albedo || !image || !saved_voxel || !saved_weight))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_fwd: null argument");
        uint32_t base = 0;
        for (uint32_t k0 = 0; k0 < nviews; k0 += kFwdViews) {
            FwdViews v{};
            v.nv = std::min(kFwdViews, nviews - k0);
            uint64_t blocks = 0;
            for (uint32_t k = 0; k < v.nv; k++) {
                const vrt_camera& c = cameras[k0 + k];
                v.cam[k] = c;
                v.block_begin[k] = (uint32_t)blocks;
                v.pixel_base[k] = base;
                uint32_t n = c.width * c.height;
                blocks += ((uint64_t)n + 63) / 64;
                base += n;
            }
            v.block_begin[v.nv] = (uint32_t)blocks;
            render_fwd_kernel<<<(uint32_t)blocks, 64, 0, s>>>(w, v, max_iters, light_dir[0], light_dir[1], light_dir[2],
                                                               intensity, albedo, image, saved_voxel, saved_weight);
            VRT_HIP(hipGetLastError());
        }
        return VRT_OK;
    });
}

int vrt_render_bwd(const vrt_world* world, uint32_t npixels, const float* grad_image, const uint32_t* saved_voxel,
                   const float* saved_weight, float* grad_albedo, void* stream) {
    return guarded([&]() -> int {
        if (!world || (npixels && (!grad_image || !saved_voxel || !saved_weight || !grad_albedo)))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd: null argument");
        if (int rcw = check_world(world, "vrt_render_bwd"); rcw != VRT_OK) return rcw;
        if (npixels == 0) return VRT_OK;
        if (npixels > 0xFFFFFF00u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd: more than 2^32 - 256 pixels");
        render_bwd_kernel<<<(npixels + 255) / 256, 256, 0, (hipStream_t)stream>>>(npixels, grad_image, saved_voxel,
                                                                                  saved_weight, grad_albedo);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}

size_t vrt_render_bwd_workspace_bytes(uint32_t npixels, uint64_t nvox) {
    return bwd_workspace_bytes(npixels, (nvox + 511) / 512);
}

int vrt_render_bwd_bricks(const vrt_world* world, uint32_t npixels, const float* grad_image,
                          const uint32_t* saved_voxel, const float* saved_weight, uint64_t nvox, float* grad_albedo,
                          void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&]() -> int {
        if (!world) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: null world");
        if (int rcw = check_world(world, "vrt_render_bwd_bricks"); rcw != VRT_OK) return rcw;
        if (npixels > 0xFFFFFF00u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: more than 2^32 - 256 pixels");
        // every saved id is below slot_high_water * 512 (the world is unchanged since the forward pass)
        const uint64_t need = (uint64_t)world->arena.high_water() * 512;
        if (nvox < need)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: nvox " + std::to_string(nvox) +
                                                      " < slot_high_water * 512 = " + std::to_string(need));
        const uint64_t nslots = (nvox + 511) / 512;
        if (nslots > kMaxSlots) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: nvox above 2^37");
        if (nvox == 0) return VRT_OK;
        if (!grad_albedo || !workspace || (npixels && (!grad_image || !saved_voxel || !saved_weight)))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: null argument");
        if (workspace_bytes < bwd_workspace_bytes(npixels, nslots))
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: workspace smaller than "
                                                  "vrt_render_bwd_workspace_bytes(npixels, nvox)");
        if (((uintptr_t)grad_albedo & 15) != 0 || ((uintptr_t)workspace & 255) != 0)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_bricks: grad_albedo must be 16-byte and the "
                                                  "workspace 256-byte aligned");
        hipStream_t s = (hipStream_t)stream;
        BwdWorkspace ws = carve_bwd(workspace, npixels, nslots);
        const uint32_t ns = (uint32_t)nslots, nblk = (uint32_t)((nslots + kScanBlock - 1) / kScanBlock);
        bwd_fill_kernel<<<(ns + 255) / 256, 256, 0, s>>>(ws.counts, ns);  // a kernel, not a memset (graph-safe)
        if (npixels) bwd_count_kernel<<<(npixels + 255) / 256, 256, 0, s>>>(npixels, saved_voxel, saved_weight,
                                                                            ws.counts, ws.rank);
        bwd_scan_blocks_kernel<<<nblk, 256, 0, s>>>(ns, ws.counts, ws.offsets, ws.block_sum);
        bwd_scan_totals_kernel<<<1, 1024, 0, s>>>(nblk, ws.block_sum);
        if (npixels) bwd_place_kernel<<<(npixels + 255) / 256, 256, 0, s>>>(npixels, saved_voxel, ws.rank, ws.offsets,
                                                                            ws.block_sum, ws.list);
        bwd_reduce_kernel<<<ns, 64, 0, s>>>(nvox, grad_image, saved_voxel, saved_weight, ws.counts, ws.offsets,
                                            ws.block_sum, ws.list, grad_albedo);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}

int vrt_render_bwd_sparse(const vrt_world* world, uint32_t npixels, const float* grad_image,
                          const uint32_t* saved_voxel, const float* saved_weight, int64_t* index, float* values,
                          void* stream) {
    return guarded([&]() -> int {
        if (!world) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_sparse: null world");
        if (int rcw = check_world(world, "vrt_render_bwd_sparse"); rcw != VRT_OK) return rcw;
        if (npixels == 0) return VRT_OK;
        if (!grad_image || !saved_voxel || !saved_weight || !index || !values)
            return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_sparse: null argument");
        if (npixels > 0xFFFFFF00u) return fail(VRT_ERR_INVALID_ARGUMENT, "vrt_render_bwd_sparse: more than 2^32 - 256 pixels");
        bwd_sparse_kernel<<<(npixels + 255) / 256, 256, 0, (hipStream_t)stream>>>(npixels, grad_image, saved_voxel,
                                                                                  saved_weight, index, values);
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    });
}

}  // extern "C\"
