// device_common.hpp — device-side building blocks shared by every gfx950 kernel of the
// SPH step: smoothing kernels, cell hashing, wave reductions.
//
// Arithmetic contract (matches the FP32 oracle): every value is float, evaluated in the
// reference's expression order, with FMA contraction disabled (-ffp-contract=off) and
// correctly rounded division and square root. Cell coordinates are computed in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mps {
namespace dev {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kBlock = 256;  // 4 wave64s per workgroup for particle-parallel kernels

// --------------------------------------------------------------------- float3 helpers
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// glm order: (x*x + y*y) + z*z
__device__ __forceinline__ float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return __builtin_sqrtf(dot3(a, a)); }

// --------------------------------------------------------------------- smoothing kernels
// Constants of MPSSPHKernel.cuh:9-15 rounded to float.
constexpr float kW = static_cast<float>(1.5666814710608447114749495456982);
constexpr float kG = static_cast<float>(14.323944878270580219199538703526);
constexpr float kCubicW = static_cast<float>(2.5464790894703253723021402139602);
constexpr float kCubicG = static_cast<float>(15.278874536821952233812841283761);
constexpr float kSTW = static_cast<float>(7.2756545413437867780061148970292);
constexpr float kSTG = static_cast<float>(21.826963624031360334018344691088);
constexpr float kEps = static_cast<float>(1.0e-20);
constexpr float kFltEps = 1.19209290e-7f;  // FLT_EPSILON

// W(d, 1/h): poly6 (CUBIC=false) or cubic spline (MPSSPHKernel.cuh:28-45)
template <bool CUBIC>
__device__ __forceinline__ float Wk(float dist, float invh) {
    const float ratio = dist * invh;
    if (ratio >= 1.0f) return 0.0f;
    if (CUBIC) {
        if (ratio >= 0.5f) {
            const float temp = (1.0f - ratio) * invh;
            return kCubicW * 2.0f * temp * temp * temp;
        }
        return kCubicW * invh * invh * invh * ((6.0f * ratio - 6.0f) * ratio * ratio + 1.0f);
    }
    const float tmp = (1.0f - ratio * ratio) * invh;
    return kW * tmp * tmp * tmp;
}
// G(d, 1/h) with grad W_ij = G * x_ij: spiky or cubic (MPSSPHKernel.cuh:46-65)
template <bool CUBIC>
__device__ __forceinline__ float Gk(float dist, float invh) {
    const float ratio = dist * invh;
    if (ratio < kEps || ratio >= 1.0f) return 0.0f;
    const float invh2 = invh * invh;
    if (CUBIC) {
        if (ratio > 0.5f) {
            const float temp = (1.0f - ratio) * invh2;
            return -kCubicG / dist * temp * temp;
        }
        return kCubicG / dist * invh2 * invh2 * (3.0f * ratio - 2.0f) * ratio;
    }
    const float tmp = (1.0f - ratio) * invh2;
    return -kG / dist * tmp * tmp;
}
// Pair factor of the viscosity operator, G(d) / ((d^2 + h2) (d + FLT_EPSILON))
// (MPSSPHNonPressureUtil.cuh:617-618), regrouped to one division: the 1/d inside G joins the
// denominators, Phi = (G d) / ((d (d^2 + h2)) (d + FLT_EPSILON)). The oracle evaluates the
// same expression in the same order (its Phik).
template <bool CUBIC>
__device__ __forceinline__ float Phik(float dist, float invh, float h2) {
    const float ratio = dist * invh;
    if (ratio < kEps || ratio >= 1.0f) return 0.0f;
    const float invh2 = invh * invh;
    float num;
    if (CUBIC) {
        if (ratio > 0.5f) {
            const float temp = (1.0f - ratio) * invh2;
            num = -kCubicG * temp * temp;
        } else {
            num = kCubicG * invh2 * invh2 * (3.0f * ratio - 2.0f) * ratio;
        }
    } else {
        const float tmp = (1.0f - ratio) * invh2;
        num = -kG * tmp * tmp;
    }
    return num / ((dist * (dist * dist + h2)) * (dist + kFltEps));
}
// surface-tension kernels (MPSSPHKernel.cuh:67-84)
__device__ __forceinline__ float STWk(float dist, float invh) {
    float ratio = dist * invh;
    ratio = ratio < 0.5f ? 0.5f : ratio;
    if (ratio >= 1.0f) return 0.0f;
    const float temp = (1.0f - ratio) * invh;
    return kSTW * temp * temp * temp;
}
__device__ __forceinline__ float STGk(float dist, float invh) {
    const float ratio = dist * invh;
    if (ratio <= 0.5f || ratio >= 1.0f) return 0.0f;
    const float temp = (1.0f - ratio) * invh * invh;
    return -kSTG / dist * temp * temp;
}

// --------------------------------------------------------------------- hashing
struct HashGeom {
    uint32_t hx, hy, hz;  // powers of two
    double inv_cell;      // 1.0 / double(cell)
};
struct i3 {
    int x, y, z;
};
// trunc(x * (1/cell)) + H/2, product in double (MPSSpatialHash.cuh:9-19)
__device__ __forceinline__ i3 grid_pos(float x, float y, float z, const HashGeom& g) {
    return {static_cast<int>(static_cast<double>(x) * g.inv_cell) + static_cast<int>(g.hx >> 1),
            static_cast<int>(static_cast<double>(y) * g.inv_cell) + static_cast<int>(g.hy >> 1),
            static_cast<int>(static_cast<double>(z) * g.inv_cell) + static_cast<int>(g.hz >> 1)};
}
// linear cell index with per-axis wrap (MPSSpatialHash.cuh:21-27)
__device__ __forceinline__ uint32_t grid_index(int x, int y, int z, const HashGeom& g) {
    return ((static_cast<uint32_t>(z) & (g.hz - 1u)) * g.hy + (static_cast<uint32_t>(y) & (g.hy - 1u))) * g.hx +
           (static_cast<uint32_t>(x) & (g.hx - 1u));
}
__device__ __forceinline__ uint32_t split_by3(uint32_t x) {
    if (x == 1024u) x--;
    x = (x | x << 16u) & 0x030000FFu;
    x = (x | x << 8u) & 0x0300F00Fu;
    x = (x | x << 4u) & 0x030C30C3u;
    x = (x | x << 2u) & 0x09249249u;
    return x;
}
// Workgroups of a launch go to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and
// each XCD has its own L2. Renumbering so that XCD x runs one contiguous range of logical
// workgroups keeps the particles a workgroup gathers, which are near it in slot order, in
// the L2 that already holds them. A bijection on [0, n) for every n.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    constexpr uint32_t X = 8u;
    const uint32_t per = n / X, rem = n % X, x = b % X;
    return x * per + min(x, rem) + b / X;
}

// Morton key; adds H/2 on top of grid_pos's H/2 (MPSSpatialHash.cuh:29-45)
__device__ __forceinline__ uint32_t grid_zindex(i3 p, const HashGeom& g) {
    const uint32_t x = static_cast<uint32_t>(p.x + static_cast<int>(g.hx >> 1u)) & (g.hx - 1u);
    const uint32_t y = static_cast<uint32_t>(p.y + static_cast<int>(g.hy >> 1u)) & (g.hy - 1u);
    const uint32_t z = static_cast<uint32_t>(p.z + static_cast<int>(g.hz >> 1u)) & (g.hz - 1u);
    return split_by3(x) | (split_by3(y) << 1u) | (split_by3(z) << 2u);
}

// --------------------------------------------------------------------- wave/block reductions
// Sum of a double over the 256-thread block in a fixed tree order; result valid in thread 0.
__device__ __forceinline__ double block_sum_d(double v, double* lds /*[kBlock/64]*/) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    }
    __syncthreads();
    return s;
}

}  // namespace dev
}  // namespace mps
