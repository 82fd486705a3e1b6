// Shared helpers for the gfx950 PF-FLIP kernels: error plumbing, index maps,
// BF16/FP16 storage conversions and deterministic wave/block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "pfflip.h"

#define PF_API extern "C" __attribute__((visibility("default")))

namespace pf {

int set_error(int code, const char* fmt, ...);

inline hipStream_t S(pf_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned cdiv(int64_t n, int64_t b) { return (unsigned)((n + b - 1) / b); }

}  // namespace pf

#define PF_CHECK_HIP(expr)                                                               \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return pf::set_error((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define PF_CHECK_LAUNCH() PF_CHECK_HIP(hipGetLastError())

#define PF_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) return pf::set_error(PF_ERR_INVALID, __VA_ARGS__); \
    } while (0)

namespace pf {

// ------------------------------------------------------------------ geometry
struct Dims {
    int nx, ny, nz;
};

__host__ __device__ inline int cell_index(const Dims& d, int i, int j, int k) {
    return (i * d.ny + j) * d.nz + k;
}

// face arrays of axis A have one extra entry along A
template <int A>
__host__ __device__ inline int face_index(const Dims& d, int i, int j, int k) {
    if (A == 0) return (i * d.ny + j) * d.nz + k;
    if (A == 1) return (i * (d.ny + 1) + j) * d.nz + k;
    return (i * d.ny + j) * (d.nz + 1) + k;
}

__host__ __device__ inline int64_t face_count(const Dims& d, int a) {
    return (int64_t)(d.nx + (a == 0)) * (d.ny + (a == 1)) * (d.nz + (a == 2));
}

// ------------------------------------------------------------------ storage types
// BF16 / FP16 are stored as raw 16-bit words; arithmetic is always FP32.
struct bf16s {
    uint16_t u;
};
struct fp16s {
    uint16_t u;
};

__device__ inline float to_f32(float v) { return v; }
__device__ inline float to_f32(bf16s v) { return __uint_as_float((uint32_t)v.u << 16); }
__device__ inline float to_f32(fp16s v) {
    _Float16 h;
    __builtin_memcpy(&h, &v.u, 2);
    return (float)h;
}

template <typename T>
__device__ inline T from_f32(float v);
template <>
__device__ inline float from_f32<float>(float v) { return v; }
template <>
__device__ inline bf16s from_f32<bf16s>(float v) {
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return bf16s{(uint16_t)(u >> 16)};
}
template <>
__device__ inline fp16s from_f32<fp16s>(float v) {
    _Float16 h = (_Float16)v;
    fp16s r;
    __builtin_memcpy(&r.u, &h, 2);
    return r;
}

// ------------------------------------------------------------------ reductions
// Wave64 butterfly reductions: every lane ends with the full result, and the
// combination order is fixed, so results are bitwise reproducible.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Block sum of up to NV doubles per thread; result valid in thread 0.
// `sh` needs NV * (blockDim/64) doubles.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[q * 16 + wid] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sh[q * 16 + w];
            v[q] = s;
        }
    }
    __syncthreads();
}

__device__ inline float block_max(float v, float* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_max(v);
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = sh[0];
        for (int w = 1; w < nw; ++w) m = fmaxf(m, sh[w]);
        v = m;
    }
    __syncthreads();
    return v;
}

// Grid-wide "last block" ticket.  Every block publishes its partial result,
// then takes a ticket; the block that takes the last ticket combines the
// partials in block order and resets the ticket for the next launch.
__device__ inline bool take_last_ticket(unsigned* ticket) {
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
        if (last) {
            __threadfence();
            *ticket = 0u;
        }
    }
    __syncthreads();
    return last;
}

// Non-negative float max through the integer ordering of IEEE bits.
__device__ inline void atomic_max_nonneg(float* addr, float v) {
    atomicMax(reinterpret_cast<unsigned*>(addr), __float_as_uint(v));
}

// Blocks of a launch are dealt round-robin to the 8 XCDs.  Remapping the
// block id so that each XCD receives a contiguous range of work items keeps
// neighbouring tiles (which share halo faces) in one XCD's L2.
__device__ inline int xcd_block(int b, int nb) {
    const int nxcd = 8;
    if (nb < nxcd * 2) return b;
    const int per = nb / nxcd, rem = nb % nxcd;
    const int x = b % nxcd, q = b / nxcd;
    // XCD x owns items [x*per + min(x,rem), ...) of length per + (x < rem)
    const int base = x * per + (x < rem ? x : rem);
    const int len = per + (x < rem ? 1 : 0);
    if (q < len) return base + q;
    return b;  // unreachable for nb = nxcd*q_max layouts; keeps the map total
}

// ------------------------------------------------------------------ random stream
// PCG hash stream (Jarzynski & Olano, "Hash Functions for GPU Rendering",
// JCGT 2020): one 32-bit state per (seed, index) pair.
__host__ __device__ inline uint32_t pcg_hash(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

__device__ inline uint32_t rng_init(int32_t seed, int32_t offset) {
    return pcg_hash((uint32_t)seed + pcg_hash((uint32_t)offset));
}

// uniform in [0, 1): 24 bits
__device__ inline float rng_uniform(uint32_t& s) {
    s = pcg_hash(s);
    return (float)(s >> 8) * (1.0f / 16777216.0f);
}

// standard normal (Box-Muller); u1 = 0 is moved to 2^-25 so log stays finite
__device__ inline float rng_normal(uint32_t& s) {
    const float u1 = rng_uniform(s);
    const float u2 = rng_uniform(s);
    const float r = sqrtf(-2.0f * logf(fmaxf(u1, 2.98023224e-8f)));
    return r * cosf(6.28318530718f * u2);
}

}  // namespace pf
