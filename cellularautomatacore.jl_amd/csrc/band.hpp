// band.hpp -- building blocks of the band-walk kernels: compile-time unrolling,
// row loads with the boundary map, plain vector loads and stores, and lane
// shifts within a wave.
#pragma once
#include <type_traits>

#include "device.hpp"


template <int I, int N>
struct Unroll {
    template <class F>
    static CAH_DEV void run(F&& f) {
        f(std::integral_constant<int, I>{});
        Unroll<I + 1, N>::run(f);
    }
};
template <int N>
struct Unroll<N, N> {
    template <class F>
    static CAH_DEV void run(F&&) {}
};

constexpr int fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// One row of a band: VB dwords per lane starting at byte x.  Off-lattice bytes
// take the boundary value; a row beyond a Constant axis-2 edge is its value.
template <int VB>
CAH_DEV void load_row(const int8_t* __restrict__ u, const Lat& L, int8_t cv1, int8_t cv2,
                      int32_t x, int32_t y, uint32_t (&X)[VB]) {
    bool o2;
    const int32_t my = map_axis0(L.b2, y, L.n2, o2);
    if (o2) {
        const uint32_t v = 0x01010101u * (uint8_t)cv2;
#pragma unroll
        for (int i = 0; i < VB; ++i) X[i] = v;
        return;
    }
    const int8_t* row = u + (int64_t)my * L.pitch;
    if (x >= 0 && x + 4 * VB <= L.n1) {
        if constexpr (VB == 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(row + x);
            X[0] = v.x; X[1] = v.y; X[2] = v.z; X[3] = v.w;
        } else if constexpr (VB == 2) {
            const uint2 v = *reinterpret_cast<const uint2*>(row + x);
            X[0] = v.x; X[1] = v.y;
        } else {
#pragma unroll
            for (int i = 0; i < VB; ++i) X[i] = reinterpret_cast<const uint32_t*>(row + x)[i];
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t w = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            bool o1;
            const int32_t mx = map_axis0(L.b1, x + 4 * i + b, L.n1, o1);
            const uint8_t v = o1 ? (uint8_t)cv1 : (uint8_t)row[mx];
            w |= (uint32_t)v << (8 * b);
        }
        X[i] = w;
    }
}

// A lane's VB output dwords of row y.  Ordinary stores: the next step reads
// them, and a lattice that fits the 256 MB Infinity Cache is then read from it
// (streaming stores, which skip it, measured 30% slower on Larger-than-Life
// 8192^2).
template <int VB>
CAH_DEV void store_row(int8_t* __restrict__ u, const Lat& L, int32_t x, int32_t y,
                       const uint32_t (&X)[VB]) {
    // x >= 0 for every writer lane: an unsigned offset from the wave-uniform row
    int8_t* row = u + (int64_t)y * L.pitch + (uint32_t)x;
    if constexpr (VB == 4) {
        *reinterpret_cast<uint4*>(row) = make_uint4(X[0], X[1], X[2], X[3]);
    } else if constexpr (VB == 2) {
        *reinterpret_cast<uint2*>(row) = make_uint2(X[0], X[1]);
    } else {
#pragma unroll
        for (int i = 0; i < VB; ++i) reinterpret_cast<uint32_t*>(row)[i] = X[i];
    }
}

// Neighbour lanes' values within a wave (DPP wave shifts: lane i receives lane
// i-1, or lane i+1).  Lanes at the wave's ends receive 0; they are guard lanes.
// bound_ctrl: the lane without a source reads 0, so the move can fold into the
// instruction that uses it as a DPP operand.
CAH_DEV uint32_t from_prev_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);
}
CAH_DEV uint32_t from_next_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);
}

template <int VB>
CAH_DEV void load_vec(const int8_t* p, uint32_t (&X)[VB]) {
    if constexpr (VB == 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        X[0] = v.x; X[1] = v.y; X[2] = v.z; X[3] = v.w;
    } else if constexpr (VB == 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        X[0] = v.x; X[1] = v.y;
    } else {
        X[0] = *reinterpret_cast<const uint32_t*>(p);
    }
}

// The wave's index in the grid, the same in every lane.
CAH_DEV int wave_id() {
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
}

// The wave index in the XCD-aware block order of block_id_xcd: neighbouring row
// chunks, which read each other's edge rows, share one L2.
CAH_DEV int wave_id_xcd() {
    return __builtin_amdgcn_readfirstlane(block_id_xcd() * (blockDim.x / 64) + threadIdx.x / 64);
}


// ---------------------------------------------------------------- mapped rows
//
// Boundary conditions without divergent branches, for lattices whose first
// side is a multiple of the lane width 4*VB and at least 64 lanes wide: every
// lane's 4*VB bytes are then either inside the lattice or a whole aligned group
// beyond one edge, which maps onto an aligned group inside (shifted for a
// Periodic axis, mirrored byte for byte for a Reflecting one) or onto the
// Constant value.  Rows map the same way with wave-uniform arithmetic.

struct LaneMap {
    int32_t src;    // first byte of the group to load (always inside the row)
    bool mirror;    // Reflecting: the group is read back to front
    bool fill;      // Constant: the lane holds the axis-1 value
};

template <int VB>
CAH_HD LaneMap lane_map(const Lat& L, int32_t x) {
    constexpr int W = 4 * VB;
    LaneMap m{x, false, false};
    if (x >= 0 && x + W <= L.n1) return m;
    if (L.b1 == CAH_BOUNDARY_PERIODIC) {
        m.src = x < 0 ? x + L.n1 : x - L.n1;
    } else if (L.b1 == CAH_BOUNDARY_REFLECTING) {
        m.src = x < 0 ? -x - W : 2 * L.n1 - x - W;
        m.mirror = true;
    } else {
        m.fill = true;
    }
    // lanes far past an edge only need a valid address
    m.src = m.src < 0 ? 0 : (m.src > L.n1 - W ? L.n1 - W : m.src);
    return m;
}

// The row a (wave-uniform) index y reads, and whether it lies beyond a
// Constant axis-2 edge.  Valid for -n2 < y < 2 n2.
CAH_HD int32_t row_map(const Lat& L, int32_t y, bool& fill) {
    fill = false;
    if (y >= 0 && y < L.n2) return y;
    if (L.b2 == CAH_BOUNDARY_PERIODIC) return y < 0 ? y + L.n2 : y - L.n2;
    if (L.b2 == CAH_BOUNDARY_REFLECTING) return y < 0 ? -1 - y : 2 * L.n2 - 1 - y;
    fill = true;
    return 0;
}

// The part of a band walk one wave owns: wave `wid` walks band wid % nbands over
// the rows [ys, ye) of chunk wid / nbands, and its lane reads the 4*VB bytes
// from column x; lanes 1..62 inside the lattice write them (the end lanes are
// guards).  Bands are 62 * 4 * VB bytes wide.
struct BandPart {
    int32_t x, ys, ye;
    bool writer, up;
};

template <int VB>
CAH_HD BandPart band_part(int32_t nbands, int32_t rows, const Lat& L, int wid, int lane) {
    constexpr int OW = 62 * 4 * VB;
    const int band = wid % nbands, chunk = wid / nbands;
    BandPart b;
    b.ys = chunk * rows;
    b.ye = b.ys + rows < L.n2 ? b.ys + rows : L.n2;
    b.x = band * OW - 4 * VB + lane * 4 * VB;
    b.writer = lane >= 1 && lane <= 62 && b.x < L.n1;
    b.up = chunk & 1;
    return b;
}

CAH_DEV uint32_t bswap32(uint32_t v) { return __builtin_amdgcn_perm(v, v, 0x00010203u); }

// Buffer addressing for the mapped rows: one resource per buffer (its base in
// scalar registers, gfx9 descriptor word 3 = 0x00020000, no range clamp: every
// address is inside the lattice by construction), the wave-uniform row offset
// in a scalar register and the lane's byte offset in a vector register, so a
// row load or store costs no vector instructions for its address.  Offsets fit
// 32 bits: mapped_ok requires pitch * n2 < 2^31.
CAH_DEV __amdgpu_buffer_rsrc_t row_rsrc(const int8_t* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(p), 0, 0xFFFFFFFF, 0x00020000);
}

template <int VB>
CAH_DEV void buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, uint32_t (&X)[VB]) {
    if constexpr (VB == 4) {
        typedef uint32_t v4 __attribute__((ext_vector_type(4)));
        const v4 v = __builtin_bit_cast(v4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
        X[0] = v.x; X[1] = v.y; X[2] = v.z; X[3] = v.w;
    } else if constexpr (VB == 2) {
        typedef uint32_t v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_bit_cast(v2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
        X[0] = v.x; X[1] = v.y;
    } else {
        X[0] = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    }
}

template <int VB>
CAH_DEV void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, const uint32_t (&X)[VB]) {
    if constexpr (VB == 4) {
        typedef uint32_t v4 __attribute__((ext_vector_type(4)));
        __builtin_amdgcn_raw_buffer_store_b128(v4{X[0], X[1], X[2], X[3]}, r, voff, soff, 0);
    } else if constexpr (VB == 2) {
        typedef uint32_t v2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(v2{X[0], X[1]}, r, voff, soff, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(X[0], r, voff, soff, 0);
    }
}

template <int VB>
CAH_DEV void load_mapped(__amdgpu_buffer_rsrc_t u, const Lat& L, const LaneMap& m,
                         uint32_t fill1, uint32_t fill2, int32_t y, uint32_t (&X)[VB]) {
    bool frow;
    const int32_t my = row_map(L, y, frow);
    buf_load<VB>(u, (uint32_t)m.src, (uint32_t)my * (uint32_t)L.pitch, X);
    uint32_t R[VB];
#pragma unroll
    for (int i = 0; i < VB; ++i) R[i] = bswap32(X[VB - 1 - i]);
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t v = m.mirror ? R[i] : X[i];
        v = m.fill ? fill1 : v;
        X[i] = frow ? fill2 : v;
    }
}

// A writer lane's row store in the mapped walk (x >= 0).
template <int VB>
CAH_DEV void store_mapped(__amdgpu_buffer_rsrc_t u, const Lat& L, int32_t x, int32_t y,
                          const uint32_t (&X)[VB]) {
    buf_store<VB>(u, (uint32_t)x, (uint32_t)y * (uint32_t)L.pitch, X);
}

// Whether a lattice's rows and lanes can all use load_mapped.
template <int VB>
CAH_HD bool mapped_ok(const Lat& L, int32_t halo_rows) {
    return L.n1 % (4 * VB) == 0 && L.n1 >= 64 * 4 * VB && L.n2 > halo_rows &&
           L.pitch * (int64_t)L.n2 < (int64_t(1) << 31);
}
