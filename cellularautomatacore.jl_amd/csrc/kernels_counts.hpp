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
// counters stay in registers; a block's four waves add theirs in LDS and the
// block stores its sums in a slot of its own, and k_counts_reduce adds the slots
// (one atomic per wave on a handful of shared counters serialises in L2: 0.3 ms
// on a 16384^2 lattice).  Values outside 0..K-1 are not counted here: the host
// compares the total with the cell count and, when cells are left over, runs
// k_hist256.
//
// k_hist256: a histogram of all 256 byte values with LDS atomics, for the rare
// lattice holding values the rule does not produce.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

constexpr int CNT_SEG = 1024;      // bytes of a row per wave-iteration
constexpr int CNT_SLOTS = 258;     // per lattice: counts[256], equal, spare
constexpr int CNT_BLOCKS = 2048;   // most blocks of one k_counts launch
constexpr int CNT_PART = 17;       // slots of partial sums: counts of 0..15, equal

__device__ inline uint32_t zero_byte_tops(uint32_t x) {
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}

__device__ inline unsigned long long wave_sum64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// s, o: pitched rows (pitch bytes apart, n1 cells each; o may be null); the
// pitch is a multiple of 16, so a 16-byte load that starts inside a row stays
// inside its pitch.  part[k * CNT_BLOCKS + b] = block b's cells equal to k
// (k < K), part[16 * CNT_BLOCKS + b] = its cells where s and o agree (0
// without o): slot-major, so the reduction reads each slot coalesced.
template <int K>
__global__ __launch_bounds__(256) void k_counts(const uint8_t* __restrict__ s, const uint8_t* __restrict__ o,
                                                int64_t pitch, int64_t n1, int64_t n2,
                                                unsigned long long* __restrict__ part) {
    static_assert(K <= 16, "counted in registers: values 0..15 at most");
    __shared__ unsigned long long sums[4][K + 1];
    const int lane = threadIdx.x & 63;
    const int64_t segs = (n1 + CNT_SEG - 1) / CNT_SEG;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    uint32_t c[K];
#pragma unroll
    for (int k = 0; k < K; ++k) c[k] = 0;
    uint32_t eq = 0;
    // the counts of one lane's 16 bytes at column x (x < n1) of a row
    auto count16 = [&](uint4 v, uint4 u, int64_t x) {
        const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
        uint32_t m[4] = {0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u};
        const int32_t nv = (int32_t)(n1 - x < 16 ? n1 - x : 16);
        if (nv < 16) {  // the lane that reaches the row's end
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t r = nv - 4 * j;
                m[j] = r >= 4 ? 0x80808080u : r <= 0 ? 0u : 0x80808080u & ((1u << (8 * r)) - 1u);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t pat = 0x01010101u * (uint32_t)k;
#pragma unroll
            for (int j = 0; j < 4; ++j) c[k] += __popc(zero_byte_tops(wd[j] ^ pat) & m[j]);
        }
        if (o) {
            const uint32_t ud[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) eq += __popc(zero_byte_tops(wd[j] ^ ud[j]) & m[j]);
        }
    };
    // segment w = y * segs + g; the grid steps w by nw, so y and g advance by
    // fixed amounts with one carry (no division in the loop)
    // (a wave past the last segment starts at y >= n2 and skips the loop, but
    // still takes part in the block's sums below)
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t dy = nw / segs, dg = nw - dy * segs;
    int64_t y = w0 / segs, g = w0 - y * segs;
    const uint4 zero = {0u, 0u, 0u, 0u};
    for (; y < n2;) {
        const int64_t x = g * CNT_SEG + lane * 16;
        const bool ok = x < n1;
        uint4 v = zero, u = zero;
        if (ok) {
            v = *reinterpret_cast<const uint4*>(s + y * pitch + x);
            if (o) u = *reinterpret_cast<const uint4*>(o + y * pitch + x);
        }
        if (ok) count16(v, u, x);
        y += dy;
        g += dg;
        if (g >= segs) {
            g -= segs;
            ++y;
        }
    }
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const unsigned long long t = wave_sum64(c[k]);
        if (lane == 0) sums[wv][k] = t;
    }
    const unsigned long long te = wave_sum64(eq);
    if (lane == 0) sums[wv][K] = te;
    __syncthreads();
    if (threadIdx.x <= K) {
        const int k = threadIdx.x;
        const unsigned long long t = sums[0][k] + sums[1][k] + sums[2][k] + sums[3][k];
        part[(int64_t)(k == K ? 16 : k) * CNT_BLOCKS + blockIdx.x] = t;
    }
}

// out[0 .. CNT_SLOTS): the sums of k_counts' nblocks slots (counts of 0..K-1,
// the agreement in out[256]); every other slot is set to 0.  One block of
// CNT_RED threads, two partial sums each at most.
constexpr int CNT_RED = 1024;
template <int K>
__global__ __launch_bounds__(CNT_RED) void k_counts_reduce(const unsigned long long* __restrict__ part,
                                                           int nblocks, unsigned long long* __restrict__ out) {
    __shared__ unsigned long long w[CNT_RED / 64][K + 1];
    unsigned long long t[K + 1];
#pragma unroll
    for (int k = 0; k <= K; ++k) t[k] = 0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < K; ++k) t[k] += part[(int64_t)k * CNT_BLOCKS + b];
        t[K] += part[(int64_t)16 * CNT_BLOCKS + b];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= K; ++k) {
        const unsigned long long v = wave_sum64(t[k]);
        if (lane == 0) w[wv][k] = v;
    }
    for (int i = threadIdx.x; i < CNT_SLOTS; i += blockDim.x) out[i] = 0;
    __syncthreads();
    if (threadIdx.x <= K) {
        const int k = threadIdx.x;
        unsigned long long v = 0;
        for (int i = 0; i < CNT_RED / 64; ++i) v += w[i][k];
        out[k == K ? 256 : k] = v;
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

// The counters of n lattices (CNT_SLOTS words each: counts of the 256 byte
// values, then the agreement) in the layout of cah_batch_counts' outputs:
// counts[i * 256 + v] and equal[i].
__global__ __launch_bounds__(256) void k_counts_pack(const unsigned long long* __restrict__ slots, int32_t n,
                                                     unsigned long long* __restrict__ counts,
                                                     unsigned long long* __restrict__ equal) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < (int64_t)n * 257;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = j / 257, v = j - i * 257;
        if (v < 256) counts[i * 256 + v] = slots[i * CNT_SLOTS + v];
        else equal[i] = slots[i * CNT_SLOTS + 256];
    }
}

// n words set to 0 (a kernel rather than a memset, so that the calls that use
// it can be captured in a graph).
__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* __restrict__ p, int64_t n) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        p[j] = 0;
}
