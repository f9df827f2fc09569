// Two-phase occlusion culling of the main view (LH_VIEW_HIZ), kernels.
//
//   k_hiz_build  one wave per 8x8 pixel tile: the tile's minimum depth
//                (reverse-Z: the farthest) from the phase-1 visibility buffer,
//                0 where a pixel is still empty.  Stored in the view's
//                radiance buffer, which the resolve overwrites later.  Also
//                records the phase-1 draw count and empties the view's
//                clip / big-triangle queues for phase 2.
//   k_hiz_cull   one lane per candidate meshlet (staged in draws[1]): its
//                conservative screen rectangle and nearest depth against the
//                tiles it touches; survivors are appended to draws[0] after
//                the phase-1 draws.
// The raster then runs again over draws[0][phase-1 count, total).
#include "hiz.h"
#include "kernels.h"

using namespace lhd;

namespace {

__global__ void __launch_bounds__(64) k_hiz_build(OneView vs) {
    const lh_view& v = vs.v;
    const uint32_t tiles_x = (v.width + hiz::kTile - 1) / hiz::kTile;
    const uint32_t tx = blockIdx.x, ty = blockIdx.y;
    const uint32_t x = tx * hiz::kTile + (threadIdx.x & 7), y = ty * hiz::kTile + (threadIdx.x >> 3);
    uint32_t d = 0xFFFFFFFFu;  // pixels outside the image do not count
    if (x < v.width && y < v.height) d = ~uint32_t(v.visbuffer[size_t(y) * v.width + x] >> 32);  // empty -> 0.0
    d = wave_min_u32(d);  // non-negative floats order as their bits
    if (threadIdx.x == 0) v.radiance[size_t(ty) * tiles_x + tx] = __uint_as_float(d);
    if (tx == 0 && ty == 0 && threadIdx.x == 0) {
        v.counters[LH_CNT_HIZ_PHASE1] = min(v.counters[LH_CNT_DRAWS], v.draw_capacity);
        v.counters[LH_CNT_CLIP] = 0;
        v.counters[LH_CNT_BIG] = 0;
    }
}

__global__ void __launch_bounds__(256) k_hiz_cull(lh_scene s, OneView vs) {
    const lh_view& v = vs.v;
    const uint32_t n = min(v.counters[LH_CNT_HIZ_CANDIDATES], v.draw_capacity);
    const MeshletReference* cand = v.draws + size_t(v.draw_capacity);
    MeshletReference* out = v.draws;
    const uint32_t tiles_x = (v.width + hiz::kTile - 1) / hiz::kTile;
    const float* tiles = v.radiance;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool draw = false;
        MeshletReference ref{0, 0};
        if (i < n) {
            ref = cand[i];
            const Instance& in = s.instances[ref.instance_index];
            const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
            const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
            f3 c;
            float r;
            world_sphere(in, ml.bounding_sphere, c, r);
            const hiz::ScreenBound b =
                hiz::sphere_bound(v.uniforms->combined_perspective_view, c, r, v.width, v.height);
            draw = !b.valid;
            if (b.valid && b.x0 <= b.x1 && b.y0 <= b.y1) {
                const uint32_t tx0 = uint32_t(b.x0) / hiz::kTile, tx1 = uint32_t(b.x1) / hiz::kTile;
                const uint32_t ty0 = uint32_t(b.y0) / hiz::kTile, ty1 = uint32_t(b.y1) / hiz::kTile;
                float farthest = 3.0e38f;
                for (uint32_t ty = ty0; ty <= ty1; ty++)
                    for (uint32_t tx = tx0; tx <= tx1; tx++) farthest = fminf(farthest, tiles[size_t(ty) * tiles_x + tx]);
                draw = !(b.zmax * (1.0f + hiz::kDepthMargin) < farthest);
            }
        }
        const uint32_t slot = wave_append(&v.counters[LH_CNT_DRAWS], draw);
        if (draw) {
            if (slot < v.draw_capacity)
                out[slot] = ref;
            else
                atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_DRAWS);
        }
    }
}

// Visibility history of the frame just rasterized: clear, then one bit per
// meshlet that owns a pixel.  Lanes of a wave that hit the same 32-bit word
// merge their bits first, so each word takes one atomic per wave.
__global__ void __launch_bounds__(256) k_history_clear(OneView vs, uint32_t words) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x)
        vs.v.history[i] = 0u;
}

__global__ void __launch_bounds__(256) k_history_mark(OneView vs, uint32_t total_meshlets) {
    const lh_view& v = vs.v;
    const uint64_t n = uint64_t(v.width) * v.height;
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < n; base += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        uint32_t word = ~0u, bit = 0u;
        if (i < n) {
            const uint64_t key = v.visbuffer[i];
            const uint32_t mid = uint32_t(key >> LH_KEY_TRI_BITS) & (LH_MAX_SCENE_MESHLETS - 1);
            if (key != LH_VISBUFFER_CLEAR && mid < total_meshlets) {
                word = mid >> 5;
                bit = 1u << (mid & 31);
            }
        }
        uint64_t pending = ballot(word != ~0u);
        while (pending) {
            const uint32_t w = __builtin_amdgcn_readlane(word, int(first_lane(pending)));
            const bool mine = word == w;
            uint32_t bits = mine ? bit : 0u;
            for (int d = 32; d >= 1; d >>= 1) bits |= __shfl_xor(bits, d, 64);
            if (lane_id() == first_lane(pending)) atomicOr(v.history + w, bits);
            pending &= ~ballot(mine);
        }
    }
}

}  // namespace

namespace lhk {

hipError_t update_history(const lh_scene& s, const OneView& vs, hipStream_t st) {
    const lh_view& v = vs.v;
    if (!v.history) return hipSuccess;
    const uint32_t words = (s.total_meshlets + 31) / 32;
    const uint64_t px = uint64_t(v.width) * v.height;
    mark("history_update", st);
    hipLaunchKernelGGL(k_history_clear, dim3(std::max(1u, std::min((words + 255) / 256, 256u))), dim3(256), 0, st,
                       vs, words);
    hipLaunchKernelGGL(k_history_mark, dim3(uint32_t(std::min<uint64_t>((px + 255) / 256, 2048))), dim3(256), 0, st,
                       vs, s.total_meshlets);
    return hipGetLastError();
}


hipError_t hiz_cull(const lh_scene& s, const OneView& vs, uint32_t blocks, hipStream_t st) {
    const lh_view& v = vs.v;
    const dim3 tiles((v.width + hiz::kTile - 1) / hiz::kTile, (v.height + hiz::kTile - 1) / hiz::kTile);
    mark("hiz_build", st);
    hipLaunchKernelGGL(k_hiz_build, tiles, dim3(64), 0, st, vs);
    mark("hiz_cull", st);
    hipLaunchKernelGGL(k_hiz_cull, dim3(std::max(blocks, 1u)), dim3(256), 0, st, s, vs);
    return hipGetLastError();
}

}  // namespace lhk
