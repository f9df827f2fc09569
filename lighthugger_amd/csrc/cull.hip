// Culling passes of the GPU-driven frame (gfx950).
//
//   k_view_reset              reset_buffers_a    src/shaders/compute/reset_buffers.glsl:9-29
//   k_clear_targets           attachment clears  src/rendering.cpp:206-211, :384
//   k_cull_instances          cull_instances     src/shaders/cull_instances.comp:7-27
//                             cull_instances_shadows               :31-56
//   k_fit_cull_instances_shadow  generate_shadow_matrices + cull_instances_shadows
//   k_cull_meshlets           write_draw_calls(_shadows)  src/shaders/write_draw_calls.comp:7-117
//
// The prefix-sum append keeps the reference's single 64-bit atomic per append
// (count in the high word, running sum in the low word,
// src/shaders/common/prefix_sum.glsl:17-30) but aggregates a whole wave into
// one atomic: lanes take consecutive slots in lane order and their inclusive
// sums from a wave scan, so the sector has exactly the invariants of
// one-by-one appends (sums strictly increasing with the slot).
//
// Every kernel is templated on its view source (one view, or a table of views
// indexed by blockIdx.z) so a batch of views runs the same arithmetic.
#include "cascade.h"
#include "hiz.h"
#include "kernels.h"
#include "lh_device.h"

using namespace lhd;

namespace {

// cull_bounding_sphere (src/shaders/common/culling.glsl:1-46).
LH_DEV bool cull_sphere(const Instance& in, const LhVec4& s, const Uniforms& u) {
    f3 c;
    float r;
    world_sphere(in, s, c, r);
    f3 v = mat4_point(u.initial_view, c);
    v.z = -v.z;
    bool visible = v.z + r > near_plane(u);
    const LhMat4& p = u.perspective;
    f3 row3{p.c[0].w, p.c[1].w, p.c[2].w}, row0{p.c[0].x, p.c[1].x, p.c[2].x}, row1{p.c[0].y, p.c[1].y, p.c[2].y};
    f3 fx = normalize3(add(row3, row0));
    f3 fy = normalize3(add(row3, row1));
    visible = visible && (v.z * fx.z + fabsf(v.x) * fx.x < r);
    visible = visible && (v.z * fy.z - fabsf(v.y) * fy.y < r);
    return !visible;
}

LH_DEV f2 shadow_view_xy(const LhMat4& view, f3 w) {
    f3 v = mat4_point(view, w);
    return {v.x, v.y};
}

// cull_bounding_sphere_shadows (culling.glsl:82-135) with
// fits_partially_inside / fits_entirely_inside(_cascade) (:48-80).
LH_DEV bool cull_sphere_shadow(const Instance& in, const LhVec4& s, const LhMat4* views, const float* radii,
                               uint32_t c) {
    f3 w;
    float r;
    world_sphere(in, s, w, r);
    f2 p = shadow_view_xy(views[c], w);
    float R = radii[c];
    if (!((fabsf(p.x) - r < R) && (fabsf(p.y) - r < R))) return true;
    for (uint32_t k = 0; k < c; k++) {
        f2 q = shadow_view_xy(views[k], w);
        float Rk = radii[k];
        if ((fabsf(q.x) + r < Rk) && (fabsf(q.y) + r < Rk)) return true;
    }
    return false;
}

// cull_cone_perspective / cull_cone_orthographic (culling.glsl:137-151).
LH_DEV bool cull_cone(const Instance& in, const Meshlet& ml, const Uniforms& u, bool ortho) {
    f3 axis = normalize3(mat3_mul_vec3(in.normal_transform, mk3(ml.cone_axis)));
    if (ortho) return dot3(mk3(u.sun_dir), axis) >= ml.cone_cutoff;
    f3 apex = mat4_point(in.transform, mk3(ml.cone_apex));
    return dot3(normalize3(sub(apex, mk3(u.camera_pos))), normalize3(axis)) >= ml.cone_cutoff;
}

// Wave-aggregated prefix_sum_inclusive_append.  Must be reached by every lane
// of the wave (lanes without an element pass visible = false).
LH_DEV void wave_prefix_append(const lh_view& v, uint32_t pass, bool visible, uint32_t index, uint32_t value) {
    uint64_t m = ballot(visible);
    if (m == 0) return;
    uint32_t incl = wave_inclusive_scan(visible ? value : 0u);
    uint32_t total = shfl(incl, 63);
    uint32_t leader = first_lane(m);
    uint64_t* sec = sector(v, pass);
    unsigned long long old = 0;
    if (lane_id() == leader)
        old = atomicAdd(reinterpret_cast<unsigned long long*>(sec),
                        (unsigned long long)total | ((unsigned long long)__popcll(m) << 32));
    old = shfl(old, int(leader));
    if (visible) {
        uint32_t slot = uint32_t(old >> 32) + lanes_below(m);
        if (slot < v.instance_capacity) {
            PrefixSumValue* vals = reinterpret_cast<PrefixSumValue*>(sec + 1);
            vals[slot] = PrefixSumValue{index, uint32_t(old) + incl};
        } else {
            atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_INSTANCES);
        }
    }
}

// prefix_sum_binary_search (prefix_sum.glsl:33-46): upper_bound on sum.
LH_DEV bool prefix_search(const uint64_t* sec, uint32_t n, uint32_t target, PrefixSumValue& out) {
    const PrefixSumValue* vals = reinterpret_cast<const PrefixSumValue*>(sec + 1);
    uint32_t first = 0, count = n;
    while (count > 0) {
        uint32_t step = count / 2, cur = first + step;
        if (target >= vals[cur].sum) {
            first = cur + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    if (first >= n) return false;
    out = vals[first];
    return true;
}

// Block-aggregated append for 256-thread blocks (every thread calls it, the
// same number of times): one device atomic per block instead of one per
// wave, since the appends of a pass all go to one counter and same-address
// atomics serialize.  Returns this thread's slot when pred holds.
constexpr uint32_t kCullThreads = 256;
LH_DEV uint32_t block_append(uint32_t* counter, bool pred, uint32_t* counts /* LDS [kCullThreads / 64 + 1] */) {
    const uint64_t m = ballot(pred);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) counts[wave] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kCullThreads / 64; w++) total += counts[w];
        counts[kCullThreads / 64] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    uint32_t slot = counts[kCullThreads / 64];
    for (uint32_t w = 0; w < wave; w++) slot += counts[w];
    slot += lanes_below(m);
    __syncthreads();  // counts is reused by the next call
    return slot;
}

}  // namespace

template <class VS>
__global__ void __launch_bounds__(64) k_view_reset(VS vs) {
    const lh_view& v = vs.get(blockIdx.z);
    uint32_t t = threadIdx.x;
    v.counters[t] = 0;  // LH_NUM_COUNTERS == 64
    if (t < LH_NUM_PASSES) sector(v, t)[0] = 0;
    if (t == 0) {
        v.misc->min_depth = 0xFFFFFFFFu;
        v.misc->max_depth = 0u;
    }
}

// Clears the visibility buffer (two keys per 16-byte store, when VIS) and the
// shadow maps (four texels per store, when SHADOW) of view v with plain
// vector stores: thread t0 of `stride` threads takes stores t0, t0 + stride, ...
LH_DEV void clear_part(const lh_view& v, bool vis, bool shadow, uint64_t t0, uint64_t stride) {
    if (vis) {
        const uint64_t nv = uint64_t(v.width) * v.height;
        const ulonglong2 kv = make_ulonglong2(LH_VISBUFFER_CLEAR, LH_VISBUFFER_CLEAR);
        for (uint64_t i = t0; 2 * i + 1 < nv; i += stride) reinterpret_cast<ulonglong2*>(v.visbuffer)[i] = kv;
        if (t0 == 0 && (nv & 1)) v.visbuffer[nv - 1] = LH_VISBUFFER_CLEAR;
    }
    if (shadow && v.shadow_maps) {
        const uint64_t ns = uint64_t(LH_NUM_CASCADES) * v.shadow_size * v.shadow_size;
        const uint4 ks = make_uint4(LH_SHADOW_CLEAR, LH_SHADOW_CLEAR, LH_SHADOW_CLEAR, LH_SHADOW_CLEAR);
        for (uint64_t i = t0; 4 * i + 3 < ns; i += stride) reinterpret_cast<uint4*>(v.shadow_maps)[i] = ks;
        if (t0 == 0)
            for (uint64_t i = ns & ~uint64_t(3); i < ns; i++) v.shadow_maps[i] = LH_SHADOW_CLEAR;
    }
}

template <class VS>
__global__ void __launch_bounds__(256) k_clear_targets(VS vs) {
    clear_part(vs.get(blockIdx.z), true, true, uint64_t(blockIdx.x) * blockDim.x + threadIdx.x,
               uint64_t(gridDim.x) * blockDim.x);
}

// Main view (shadow = false, sector 0) or cascade blockIdx.y (sector 1 + c).
template <bool SHADOW, class VS>
__global__ void __launch_bounds__(256) k_cull_instances(lh_scene s, VS vs) {
    const lh_view& v = vs.get(blockIdx.z);
    const Uniforms& u = *v.uniforms;
    uint32_t n = min(u.num_instances, s.num_instances);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cascade = SHADOW ? blockIdx.y : 0u;
    bool visible = false;
    uint32_t value = 0;
    if (i < n) {
        const Instance& in = s.instances[i];
        const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
        value = mi.num_meshlets;
        visible = SHADOW ? !cull_sphere_shadow(in, mi.bounding_sphere, v.misc->shadow_view_matrices,
                                               v.misc->shadow_sphere_radii, cascade)
                         : !cull_sphere(in, mi.bounding_sphere, u);
    }
    wave_prefix_append(v, SHADOW ? 1 + cascade : 0, visible, i, value);
}

// generate_shadow_matrices fused into cull_instances_shadows: every block
// fits the four cascades itself (lanes 0..3, identical arithmetic, so every
// block sees the same bits) into LDS and block (0, 0) publishes them to
// MiscStorage; this saves one dependent launch per frame.
template <class VS>
__global__ void __launch_bounds__(256) k_fit_cull_instances_shadow(lh_scene s, VS vs) {
    const lh_view& v = vs.get(blockIdx.z);
    __shared__ LhMat4 views[LH_NUM_CASCADES];
    __shared__ float radii[LH_NUM_CASCADES];
    const Uniforms& u = *v.uniforms;
    const uint32_t min_bits = v.misc->min_depth, max_bits = v.misc->max_depth;
    if (threadIdx.x < LH_NUM_CASCADES) {
        const cascade::Fit f = cascade::fit(u, min_bits, max_bits, threadIdx.x);
        views[threadIdx.x] = f.view;
        radii[threadIdx.x] = f.radius;
        if (blockIdx.x == 0 && blockIdx.y == 0) cascade::store(*v.misc, threadIdx.x, f);
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && min_bits == 0xFFFFFFFFu)
        atomicOr(&v.counters[LH_CNT_FLAGS], LH_FLAG_EMPTY_VIEW);
    __syncthreads();
    const uint32_t n = min(u.num_instances, s.num_instances);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = blockIdx.y;
    bool visible = false;
    uint32_t value = 0;
    if (i < n) {
        const Instance& in = s.instances[i];
        const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
        value = mi.num_meshlets;
        visible = !cull_sphere_shadow(in, mi.bounding_sphere, views, radii, c);
    }
    wave_prefix_append(v, 1 + c, visible, i, value);
}

// One pass per blockIdx.y (pass = pass0 + blockIdx.y), one view per
// blockIdx.z; grid-stride over the meshlets of the pass's visible instances.
// With LH_VIEW_HIZ (single view) the main-view survivors are split: large or
// unbounded screen footprints go to draws[0] (phase-1 occluders), small ones
// to the candidate list staged in draws[1].
// `clear` (LH_CLEAR_*): render targets of the view that the grid clears after
// its culling, as a clear kernel launched before the cull would; the stores
// follow each block's last load, so they do not hold up the cull's loads.
template <class VS>
__global__ void __launch_bounds__(kCullThreads) k_cull_meshlets(lh_scene s, VS vs, uint32_t pass0, uint32_t clear) {
    __shared__ uint32_t counts[kCullThreads / 64 + 1];
    const lh_view& v = vs.get(blockIdx.z);
    const uint32_t pass = pass0 + blockIdx.y;
    const Uniforms& u = *v.uniforms;
    const bool hiz = std::is_same<VS, OneView>::value && pass == 0 && (v.flags & LH_VIEW_HIZ) != 0;
    const uint64_t* sec = sector(v, pass);
    const uint64_t head = sec[0];
    const uint32_t total = uint32_t(head);
    const uint32_t n = min(uint32_t(head >> 32), v.instance_capacity);
    MeshletReference* out = v.draws + size_t(pass) * v.draw_capacity;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        uint32_t id = base + threadIdx.x;
        bool survive = false, candidate = false, alpha = false;
        PrefixSumValue e{0, 0};
        uint32_t local = 0;
        if (id < total && prefix_search(sec, n, id, e)) {
            const Instance& in = s.instances[e.index];
            const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
            local = id - (e.sum - mi.num_meshlets);
            const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[local];
            alpha = (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) != 0;
            if (pass == 0) {
                survive = !cull_sphere(in, ml.bounding_sphere, u) && (alpha || !cull_cone(in, ml, u, false));
                if (survive && hiz) {
                    // phase 1: visible in this view's last frame, or a large footprint
                    const uint32_t mid = s.instance_meshlet_base[e.index] + local;
                    const bool seen = v.history && ((v.history[mid >> 5] >> (mid & 31)) & 1u);
                    if (!seen) {
                        f3 c;
                        float r;
                        world_sphere(in, ml.bounding_sphere, c, r);
                        const hiz::ScreenBound b =
                            hiz::sphere_bound(u.combined_perspective_view, c, r, v.width, v.height);
                        candidate = b.valid && max(b.x1 - b.x0, b.y1 - b.y0) <= hiz::kMaxRectPx;
                        survive = !candidate;
                    }
                }
            } else {
                survive = !cull_sphere_shadow(in, ml.bounding_sphere, v.misc->shadow_view_matrices,
                                              v.misc->shadow_sphere_radii, pass - 1) &&
                          (alpha || !cull_cone(in, ml, u, true));
            }
        }
        wave_append(&v.counters[LH_CNT_ALPHA_DRAWS + pass], (survive || candidate) && alpha);
        uint32_t slot = block_append(&v.counters[LH_CNT_DRAWS + pass], survive, counts);
        if (survive) {
            if (slot < v.draw_capacity) {
                MeshletReference r;
                r.instance_index = e.index;
                r.meshlet_index = uint16_t(local);
                out[slot] = r;
            } else {
                atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_DRAWS);
            }
        }
        if (hiz) {
            const uint32_t cslot = wave_append(&v.counters[LH_CNT_HIZ_CANDIDATES], candidate);
            if (candidate) {
                if (cslot < v.draw_capacity) {
                    MeshletReference r;
                    r.instance_index = e.index;
                    r.meshlet_index = uint16_t(local);
                    v.draws[size_t(v.draw_capacity) + cslot] = r;
                } else {
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_DRAWS);
                }
            }
        }
    }
    if (clear != 0) {
        const uint64_t b = uint64_t(blockIdx.y) * gridDim.x + blockIdx.x;
        clear_part(v, (clear & lhk::kClearVisbuffer) != 0, (clear & lhk::kClearShadowMaps) != 0,
                   b * blockDim.x + threadIdx.x, uint64_t(gridDim.x) * gridDim.y * blockDim.x);
    }
}

// ---------------------------------------------------------------- launchers

namespace lhk {

template <class VS>
hipError_t view_reset(const VS& vs, hipStream_t st) {
    mark("view_reset", st);
    hipLaunchKernelGGL(k_view_reset<VS>, dim3(1, 1, vs.count()), dim3(64), 0, st, vs);
    return hipGetLastError();
}

template <class VS>
hipError_t clear_targets(const VS& vs, uint64_t max_pixels, hipStream_t st) {
    // enough 256-thread blocks for ~8 stores per thread on the largest view
    uint32_t blocks = uint32_t(std::min<uint64_t>(std::max<uint64_t>((max_pixels * 5 / 4 + 4095) / 4096, 1), 2048));
    if (vs.count() > 1) blocks = std::max(blocks / 4, 16u);
    mark("clear_targets", st);
    hipLaunchKernelGGL(k_clear_targets<VS>, dim3(blocks, 1, vs.count()), dim3(256), 0, st, vs);
    return hipGetLastError();
}

template <class VS>
hipError_t cull_instances(const lh_scene& s, const VS& vs, bool shadow, hipStream_t st) {
    uint32_t n = std::max(s.num_instances, 1u);
    dim3 grid((n + 255) / 256, shadow ? LH_NUM_CASCADES : 1, vs.count());
    mark(shadow ? "cull_instances_shadow" : "cull_instances", st);
    if (shadow)
        hipLaunchKernelGGL((k_cull_instances<true, VS>), grid, dim3(256), 0, st, s, vs);
    else
        hipLaunchKernelGGL((k_cull_instances<false, VS>), grid, dim3(256), 0, st, s, vs);
    return hipGetLastError();
}

template <class VS>
hipError_t fit_cull_instances_shadow(const lh_scene& s, const VS& vs, hipStream_t st) {
    uint32_t n = std::max(s.num_instances, 1u);
    mark("fit_cull_instances_shadow", st);
    hipLaunchKernelGGL(k_fit_cull_instances_shadow<VS>, dim3((n + 255) / 256, LH_NUM_CASCADES, vs.count()),
                       dim3(256), 0, st, s, vs);
    return hipGetLastError();
}

template <class VS>
hipError_t cull_meshlets(const lh_scene& s, const VS& vs, uint32_t pass0, uint32_t npasses, uint32_t blocks,
                         hipStream_t st, uint32_t clear) {
    if (vs.count() > 1) blocks = std::max(blocks / 8, 32u);
    mark(pass0 == 0 ? "cull_meshlets_view" : "cull_meshlets_shadow", st);
    hipLaunchKernelGGL(k_cull_meshlets<VS>, dim3(blocks, npasses, vs.count()), dim3(256), 0, st, s, vs, pass0, clear);
    return hipGetLastError();
}

#define LH_INSTANTIATE(VS)                                                                                         \
    template hipError_t view_reset<VS>(const VS&, hipStream_t);                                                    \
    template hipError_t clear_targets<VS>(const VS&, uint64_t, hipStream_t);                                       \
    template hipError_t cull_instances<VS>(const lh_scene&, const VS&, bool, hipStream_t);                        \
    template hipError_t fit_cull_instances_shadow<VS>(const lh_scene&, const VS&, hipStream_t);                   \
    template hipError_t cull_meshlets<VS>(const lh_scene&, const VS&, uint32_t, uint32_t, uint32_t, hipStream_t, uint32_t);
LH_INSTANTIATE(lhd::OneView)
LH_INSTANTIATE(lhd::ViewTable)
#undef LH_INSTANTIATE

}  // namespace lhk
