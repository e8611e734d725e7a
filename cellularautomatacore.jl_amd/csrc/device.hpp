// device.hpp -- per-cell building blocks shared by every sweep kernel (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cahip.h"

#define CAH_DEV __device__ __forceinline__
// the index arithmetic that the host-side launch checks share with the kernels
#define CAH_HD __host__ __device__ __forceinline__

// Philox4x32-10 (src/random.jl:6-37).  The key schedule is the same for every
// lane of a wave, so the compiler keeps it in scalar registers; counter words
// that are wave-uniform (the time, the high half of the cell id) are multiplied
// on the scalar unit in the first two rounds.
// Each three-way XOR is one v_bitop3_b32 (truth table 0x96) on gfx950.
CAH_DEV uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;

// One Philox4x32 round given the round's first product p0 = M0 * c0.
CAH_DEV void philox_round(uint64_t p0, uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                          uint32_t& k0, uint32_t& k1) {
    const uint64_t p1 = (uint64_t)PHILOX_M1 * (uint64_t)c2;
    const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
}

// Rounds R0 .. 9 of Philox4x32-10.
template <int R0 = 0>
CAH_DEV void philox_rounds(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                           uint32_t k1) {
#pragma unroll
    for (int r = R0; r < 10; ++r) philox_round((uint64_t)PHILOX_M0 * (uint64_t)c0, c0, c1, c2, c3, k0, k1);
}

CAH_DEV void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                      uint32_t k1) {
    philox_rounds<0>(c0, c1, c2, c3, k0, k1);
}

// _philox_draw + _uniform (src/random.jl:102-120): FP64 in [0, 1).
CAH_DEV double uniform64(uint64_t seed, uint64_t cellid, int64_t t, uint32_t k) {
    uint32_t c0 = (uint32_t)cellid, c1 = (uint32_t)(cellid >> 32), c2 = (uint32_t)(uint64_t)t,
             c3 = k;
    philox10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t bits = ((uint64_t)c0 << 32) | (uint64_t)c1;
    return (double)(bits >> 11) * 0x1.0p-53;
}

// _uniform32 (src/random.jl:109-114): 24 random bits as a Float32.
CAH_DEV float uniform32(uint64_t seed, uint64_t cellid, int64_t t, uint32_t k) {
    uint32_t c0 = (uint32_t)cellid, c1 = (uint32_t)(cellid >> 32), c2 = (uint32_t)(uint64_t)t,
             c3 = k;
    philox10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    return (float)(c0 >> 8) * (1.0f / 16777216.0f);
}

// The Philox4x32-10 draw of cell rowid + J (1-based) at step t with counter word
// c3 (the draw's k): r0, r1 are the first two output words.  With fewer than
// 2^32 cells (SMALL) the cell id is a 32-bit word without carries, so the first
// round's product M0 * (rowid + J) is one multiply-add, M0 * rowid + M0 * J.
template <bool SMALL, int J>
CAH_DEV void cell_philox(uint32_t c3, uint64_t rowid, int64_t t, uint64_t seed, uint32_t& r0, uint32_t& r1) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t c0, c1, c2 = (uint32_t)(uint64_t)t;
    if constexpr (SMALL) {
        c0 = (uint32_t)rowid;
        c1 = 0;
        const uint64_t p0 = (uint64_t)PHILOX_M0 * c0 + (uint64_t)PHILOX_M0 * (uint32_t)J;
        philox_round(p0, c0, c1, c2, c3, k0, k1);
        philox_rounds<1>(c0, c1, c2, c3, k0, k1);
    } else {
        const uint64_t cellid = rowid + (uint64_t)J;
        c0 = (uint32_t)cellid;
        c1 = (uint32_t)(cellid >> 32);
        philox10(c0, c1, c2, c3, k0, k1);
    }
    r0 = c0;
    r1 = c1;
}


// The 53-bit draw of _uniform, x = (r0 << 32 | r1) >> 11, so u = x * 2^-53.
CAH_DEV uint64_t draw53(uint32_t r0, uint32_t r1) { return (((uint64_t)r0 << 32) | r1) >> 11; }

// u < p for u = x * 2^-53 as an integer test x < thr53(p): p * 2^53 is exact
// (a power-of-two scaling), and an integer x is below a real P exactly when it is
// below ceil(P).  NaN and p <= 0 give 0 (never), p >= 1 gives 2^53 (always).
CAH_DEV uint64_t thr53(double p) {
    const double P = p * 0x1.0p53;
    if (!(P > 0.0)) return 0;
    if (P >= 0x1.0p53) return 1ull << 53;
    return (uint64_t)ceil(P);
}

// x < T for the 53-bit draw x = (r0 << 32 | r1) >> 11 is r0:r1 < T * 2^11, so the
// high word alone decides it unless r0 equals hi53(T) = floor(T * 2^11 / 2^32):
// r0 < hi53(T) is a hit, r0 > hi53(T) a miss, and a tie (probability 2^-32) needs
// r1.  T = 2^53 (always) maps to 0xFFFFFFFF, so that every r0 but 0xFFFFFFFF is
// a hit and that one is a tie.
CAH_DEV uint32_t hi53(uint64_t T) { return T >= (1ull << 53) ? 0xFFFFFFFFu : (uint32_t)(T >> 21); }

// The same for the Float32 draw u = x24 * 2^-24 (_uniform32) against a Float32 p.
CAH_DEV uint32_t thr24(float p) {
    const float P = p * 16777216.0f;
    if (!(P > 0.0f)) return 0;
    if (P >= 16777216.0f) return 1u << 24;
    return (uint32_t)ceilf(P);
}

// bernoulli(ctx, p, k) (src/random.jl:129-130): the comparison stays in FP64 for a
// Float64 p, and uses the 24-bit draw for a Float32 p.
CAH_DEV bool bernoulli(uint64_t seed, uint64_t cellid, int64_t t, uint32_t k, double p,
                       int p_f32) {
    if (p_f32) return uniform32(seed, cellid, t, k) < (float)p;
    return uniform64(seed, cellid, t, k) < p;
}

// _map_axis (src/spaces.jl:46-51) on 0-based indices.  `out` is set only for a
// Constant axis and an index off the lattice.
CAH_HD int32_t map_axis0(int32_t kind, int32_t g, int32_t n, bool& out) {
    out = false;
    if (g >= 0 && g < n) return g;
    if (kind == CAH_BOUNDARY_PERIODIC) {
        if (g >= -n && g < 2 * n) return g < 0 ? g + n : g - n;  // within one period: no division
        int32_t m = g % n;
        return m < 0 ? m + n : m;
    }
    if (kind == CAH_BOUNDARY_REFLECTING) {
        if (g >= -n && g < 2 * n) return g < 0 ? -1 - g : 2 * n - 1 - g;
        const int32_t n2 = 2 * n;
        int32_t m = g % n2;
        if (m < 0) m += n2;
        return m < n ? m : n2 - 1 - m;
    }
    out = true;
    return g;
}

// Lattice geometry and edges as the kernels see it.  The state is stored dense
// per column-major row (first axis fastest) with a row pitch in elements.
struct Lat {
    int64_t pitch;
    int32_t n1, n2;
    int32_t b1, b2;
    int32_t two_d;
};

// The block's index in an XCD-aware order.  Workgroups are dealt to the eight
// XCDs round robin (block b to XCD b % 8), and each XCD has its own L2.
// Numbering the blocks so that XCD x runs one contiguous range of them puts
// neighbouring tiles, which read each other's edges, on one L2.  A bijection of
// 0 .. gridDim.x - 1 for any grid size.
CAH_DEV int block_id_xcd() {
    const int n = gridDim.x, b = blockIdx.x;
    const int q = n >> 3, r = n & 7, xcd = b & 7;
    return xcd * q + min(xcd, r) + (b >> 3);
}

// _boundary_read (src/backends.jl:177-200): the value at 0-based (g1, g2), which may
// lie off the lattice.  When both axes are off a Constant edge, axis 2 wins.
template <typename T>
CAH_DEV T read_mapped(const T* u, const Lat& L, T cv1, T cv2, int32_t g1, int32_t g2) {
    bool o1, o2;
    const int32_t m2 = map_axis0(L.b2, g2, L.n2, o2);
    if (o2) return cv2;
    const int32_t m1 = map_axis0(L.b1, g1, L.n1, o1);
    if (o1) return cv1;
    return u[(int64_t)m2 * L.pitch + m1];
}

// _gate_open (src/schemes.jl:84-89).  `gate` is the Checkerboard half (0: even
// index sum first, 1: odd); cellid is 1-based on the unpadded lattice.
CAH_DEV bool gate_open(int32_t scheme, int32_t gate, double sp, int32_t sp_f32, uint64_t seed,
                       uint64_t cellid, int64_t t, int32_t x, int32_t y, int32_t two_d) {
    if (scheme == CAH_SCHEME_STOCHASTIC) {
        if (sp_f32) return uniform32(seed, cellid, t, 0) < (float)sp;
        return uniform64(seed, cellid, t, 0) < sp;
    }
    if (scheme == CAH_SCHEME_CHECKERBOARD) {
        const int32_t s = (x + 1) + (two_d ? (y + 1) : 0);
        return (s & 1) == gate;
    }
    return true;
}

// ---------------------------------------------------------------- SWAR on bytes

// Bytes equal to `v` (v in 0..255 replicated) -> 0x01 in those bytes, 0 elsewhere.
// Exact for every byte value (no borrow crosses a byte).
CAH_HD uint32_t bytes_eq(uint32_t x, uint32_t vrep) {
    const uint32_t t = x ^ vrep;                                   // 0 where equal
    const uint32_t nz = (((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
    return (nz >> 7) ^ 0x01010101u;
}

CAH_DEV uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t sh) {
    return __builtin_amdgcn_alignbit(hi, lo, sh);
}
