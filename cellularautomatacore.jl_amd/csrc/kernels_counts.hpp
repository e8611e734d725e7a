// Cell counts of an Int8 lattice by value, and the number of cells equal to a
// second lattice of the same shape: state_counts (src/measures.jl:14-20) and the
// agreement count of cohens_kappa (src/evaluation.jl:43-47) on the device, so
// that an ensemble member's summary leaves the GPU as a few counters instead of
// the whole lattice.
//
// k_counts<K>: a wave covers 1024 bytes of one row, 16 per lane.  Every 32-bit
// word is compared with the values 0..K-1 at once: a byte of x = w ^ v*0x01010101
// is zero exactly where the cell equals v, and
//   ~(((x & 0x7F7F7F7F) + 0x7F7F7F7F) | x | 0x7F7F7F7F)
// sets the top bit of every zero byte and of nothing else (the addition cannot
// carry between bytes, since each byte sum is at most 0xFE).  The popcount of
// that word, masked to the row's cells, adds to the lane's counter for v.  The
// counters stay in registers; each wave adds its sums to out[] with one atomic
// per value.  Values outside 0..K-1 are not counted here: the host compares the
// total with the cell count and, when cells are left over, runs k_hist256.
//
// k_hist256: a histogram of all 256 byte values with LDS atomics, for the rare
// lattice holding values the rule does not produce.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

constexpr int CNT_SEG = 1024;      // bytes of a row per wave-iteration
constexpr int CNT_SLOTS = 258;     // per lattice: counts[256], equal, spare

__device__ inline uint32_t zero_byte_tops(uint32_t x) {
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}

__device__ inline unsigned long long wave_sum64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// s, o: pitched rows (pitch bytes apart, n1 cells each; o may be null); the
// pitch is a multiple of 16, so a 16-byte load that starts inside a row stays
// inside its pitch.  out[k] += cells equal to k (k < K); out[256] += cells where
// s and o agree.
template <int K>
__global__ __launch_bounds__(256) void k_counts(const uint8_t* __restrict__ s, const uint8_t* __restrict__ o,
                                                int64_t pitch, int64_t n1, int64_t n2,
                                                unsigned long long* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t segs = (n1 + CNT_SEG - 1) / CNT_SEG;
    const int64_t total = segs * n2;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    uint32_t c[K];
#pragma unroll
    for (int k = 0; k < K; ++k) c[k] = 0;
    uint32_t eq = 0;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < total; w += nw) {
        const int64_t y = w / segs;
        const int64_t x = (w - y * segs) * CNT_SEG + lane * 16;
        if (x >= n1) continue;
        const uint4 v = *reinterpret_cast<const uint4*>(s + y * pitch + x);
        const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
        uint32_t m[4];
        const int64_t nv = n1 - x;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t r = nv - 4 * j;
            m[j] = r >= 4 ? 0x80808080u : r <= 0 ? 0u : 0x80808080u & ((1u << (8 * (int)r)) - 1u);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t pat = 0x01010101u * (uint32_t)k;
#pragma unroll
            for (int j = 0; j < 4; ++j) c[k] += __popc(zero_byte_tops(wd[j] ^ pat) & m[j]);
        }
        if (o) {
            const uint4 u = *reinterpret_cast<const uint4*>(o + y * pitch + x);
            const uint32_t ud[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) eq += __popc(zero_byte_tops(wd[j] ^ ud[j]) & m[j]);
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const unsigned long long t = wave_sum64(c[k]);
        if (lane == 0 && t) atomicAdd(out + k, t);
    }
    if (o) {
        const unsigned long long t = wave_sum64(eq);
        if (lane == 0 && t) atomicAdd(out + 256, t);
    }
}

// out[b] += cells holding byte value b, over all 256 values.
__global__ __launch_bounds__(256) void k_hist256(const uint8_t* __restrict__ s, int64_t pitch, int64_t n1,
                                                 int64_t n2, unsigned long long* __restrict__ out) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n = n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = i / n1;
        atomicAdd(&h[s[y * pitch + (i - y * n1)]], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(out + threadIdx.x, (unsigned long long)h[threadIdx.x]);
}
