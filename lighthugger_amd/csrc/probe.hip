// Device ceilings the raster is priced against (liblh_probe.so; not part of
// the frame's C ABI).
//
//   lhp_atomic_min_u64 / _u32  throughput of device-scope atomicMin on a
//                              framebuffer-sized buffer, with the raster's
//                              access shapes: lanes of a wave on 8x8 pixel
//                              tiles of a W-wide image (the medium and big
//                              triangle paths), or on scattered pixels (one
//                              small triangle per lane).
//   lhp_copy                   a plain 16-byte-per-lane copy, the reference
//                              kernel of known traffic that scales the HBM
//                              counters (FETCH_SIZE, WRITE_SIZE).
//   lhp_rcp_check              counts the floats of a range where rcp_exact
//                              differs from the IEEE quotient 1.0f / a.
#include <hip/hip_runtime.h>

#include "lh_device.h"

#include <cstdint>

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Word index of op k of this lane.  pattern 0: the wave covers an 8x8 tile at
// a hashed tile position; pattern 1: every lane a hashed pixel.  Both stay in
// [0, W * H).
__device__ __forceinline__ uint32_t target(uint32_t wave, uint32_t lane, uint32_t k, uint32_t W, uint32_t H,
                                           uint32_t pattern) {
    const uint32_t h = mix32(wave * 0x9E3779B9u + k * 0x85EBCA6Bu + (pattern ? lane * 0xC2B2AE35u : 0u));
    if (pattern == 0) {
        const uint32_t tw = W / 8, th = H / 8;
        const uint32_t t = h % (tw * th);
        const uint32_t x = (t % tw) * 8 + (lane & 7), y = (t / tw) * 8 + (lane >> 3);
        return y * W + x;
    }
    return h % (W * H);
}

template <class T>
__global__ void __launch_bounds__(256) k_atomic_min(T* buf, uint32_t W, uint32_t H, uint32_t ops, uint32_t pattern,
                                                    uint32_t seed) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    for (uint32_t k = 0; k < ops; k++) {
        const uint32_t o = target(wave, lane, k, W, H, pattern);
        const T v = T(mix32(o ^ seed ^ (k << 8)));
        atomicMin(buf + o, v);
    }
}

__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Floats with bit patterns [lo, lo + n): one per lane of a grid-stride loop.
__global__ void __launch_bounds__(256) k_rcp_check(uint32_t lo, uint32_t n, unsigned long long* mismatches) {
    uint32_t bad = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float a = __uint_as_float(lo + i);
        bad += __float_as_uint(lhd::rcp_exact(a)) != __float_as_uint(1.0f / a) ? 1u : 0u;
    }
    if (bad) atomicAdd(mismatches, (unsigned long long)bad);
}

}  // namespace

extern "C" {

// `waves` waves of `ops` atomics each on the W x H buffer (W, H multiples of 8).
int lhp_atomic_min_u64(uint64_t* buf, uint32_t W, uint32_t H, uint32_t waves, uint32_t ops, uint32_t pattern,
                       void* stream) {
    if (!buf || W < 8 || H < 8 || W % 8 || H % 8 || waves % 4 || pattern > 1) return 1;
    hipLaunchKernelGGL((k_atomic_min<unsigned long long>), dim3(waves / 4), dim3(256), 0,
                       static_cast<hipStream_t>(stream), reinterpret_cast<unsigned long long*>(buf), W, H, ops,
                       pattern, 0x1234u);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int lhp_atomic_min_u32(uint32_t* buf, uint32_t W, uint32_t H, uint32_t waves, uint32_t ops, uint32_t pattern,
                       void* stream) {
    if (!buf || W < 8 || H < 8 || W % 8 || H % 8 || waves % 4 || pattern > 1) return 1;
    hipLaunchKernelGGL((k_atomic_min<uint32_t>), dim3(waves / 4), dim3(256), 0, static_cast<hipStream_t>(stream), buf,
                       W, H, ops, pattern, 0x1234u);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// Copies `bytes` (a multiple of 16) from src to dst.
int lhp_copy(const void* src, void* dst, uint64_t bytes, void* stream) {
    if (!src || !dst || bytes % 16) return 1;
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// Adds to *mismatches (device memory) the number of floats with bit patterns
// in [lo_bits, hi_bits] where rcp_exact(a) != 1.0f / a.
int lhp_rcp_check(uint32_t lo_bits, uint32_t hi_bits, unsigned long long* mismatches, void* stream) {
    if (!mismatches || hi_bits < lo_bits || hi_bits - lo_bits == 0xFFFFFFFFu) return 1;
    hipLaunchKernelGGL(k_rcp_check, dim3(8192), dim3(256), 0, static_cast<hipStream_t>(stream), lo_bits,
                       hi_bits - lo_bits + 1, mismatches);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // extern "C"
