// gate.hpp -- which cells of a band walk update: the mask and the scheme's gate,
// four cells a dword (_update_cell!, src/backends.jl:444-455).  A cell that is
// masked out or whose gate is closed keeps its value; the others take the rule's.
#pragma once
#include "band.hpp"

struct BandGate {
    const uint8_t* mask;  // n1 x n2 bytes without pitch (0: the cell never updates); nullptr: none
    int32_t scheme;       // CAH_SCHEME_*
    int32_t gate;         // Checkerboard: 0 the cells of even index sum (1-based), 1 the odd ones
    int32_t sp_f32;       // Stochastic: the probability is a Float32 (the 24-bit draw)
    double sp;            // Stochastic: the update probability
};

// Whether any cell can be kept: a mask, or a gate other than Synchronous's.
CAH_HD bool band_gated(const BandGate& g) {
    return g.mask != nullptr || g.scheme != CAH_SCHEME_SYNCHRONOUS;
}

// The mask bytes of row y for a lane's 4*VB cells from column x.  FAST: the
// lattice suits load_mapped (n1 is a multiple of 4*VB, n1 * n2 < 2^31), and the
// lane reads its group at `src` (x itself for every writer lane) with one buffer
// load; otherwise byte by byte, bytes off the lattice reading as 0.
template <int VB, bool FAST>
CAH_DEV void mask_row(const uint8_t* mask, __amdgpu_buffer_rsrc_t rm, const Lat& L, int32_t src, int32_t x,
                      int32_t y, uint32_t (&M)[VB]) {
    if constexpr (FAST) {
        buf_load<VB>(rm, (uint32_t)src, (uint32_t)y * (uint32_t)L.n1, M);
    } else {
        const uint8_t* row = mask + (int64_t)y * L.n1;
#pragma unroll
        for (int i = 0; i < VB; ++i) {
            uint32_t w = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int32_t xx = x + 4 * i + b;
                w |= (uint32_t)(xx >= 0 && xx < L.n1 ? row[xx] : 0) << (8 * b);
            }
            M[i] = w;
        }
    }
}

// Checkerboard's open cells (0x01 bytes) of a dword of row y whose first column
// is a multiple of 4: the cells whose 1-based index sum (x + b + 1) + (y + 1)
// has the gate's parity (0 even, 1 odd), i.e. the bytes b with b + y of that
// parity.
CAH_HD uint32_t checker_open(int32_t y, int32_t gate) { return ((y + gate) & 1) ? 0x01000100u : 0x00010001u; }

// The cells of a lane's VB dwords of row y that update, as 0x01 bytes: the mask
// bytes M (when the sweep has a mask) and the gate (_gate_open,
// src/schemes.jl:84-89).  x, the lane's first column, is a multiple of 4, and
// rowid the 1-based cell id of that cell.  Stochastic draws _uniform(seed,
// cellid, t, 0) per cell (cell_philox with k = 0; SMALL: fewer than 2^32 cells)
// and opens it when the draw is below p: for a Float32 p as x24 < thr24(p), for
// a Float64 p as x53 < thr53(p) (device.hpp), both exact.  The Float64 test
// looks at the draw's high word against hi53 and forms a lane's draws again in
// full only on a tie (one draw in 2^32).
template <int VB, bool SMALL>
CAH_DEV void band_open(const BandGate& g, const uint32_t (&M)[VB], uint64_t rowid, int32_t y, uint64_t seed,
                       int64_t t, uint64_t sp53, uint32_t sp24, uint32_t (&open)[VB]) {
#pragma unroll
    for (int i = 0; i < VB; ++i) open[i] = g.mask ? bytes_eq(M[i], 0u) ^ 0x01010101u : 0x01010101u;
    if (g.scheme == CAH_SCHEME_CHECKERBOARD) {
        const uint32_t pat = checker_open(y, g.gate);
#pragma unroll
        for (int i = 0; i < VB; ++i) open[i] &= pat;
    } else if (g.scheme == CAH_SCHEME_STOCHASTIC) {
        // One 32-bit limit on the draw's first word for both widths: the
        // Float32 test (r0 >> 8) < sp24 is r0 < 256 sp24 exactly, and the
        // Float64 one is decided by r0 against hi53 but on a tie.  A word equal
        // to the limit (a Float64 tie, or 2^32 - 1 under p >= 1 in Float32)
        // takes the lane's draws again with the exact tests.
        const bool f32 = g.sp_f32 != 0;
        const uint32_t lim = f32 ? (sp24 >= (1u << 24) ? 0xFFFFFFFFu : sp24 << 8) : hi53(sp53);
        uint32_t bits[VB];
        bool tie = false;
        Unroll<0, VB>::run([&](auto I) {
            constexpr int i = decltype(I)::value;
            bits[i] = 0;
            Unroll<0, 4>::run([&](auto B) {
                constexpr int b = decltype(B)::value;
                uint32_t r0, r1;
                cell_philox<SMALL, 4 * i + b>(0u, rowid, t, seed, r0, r1);
                tie |= r0 == lim;
                bits[i] |= (r0 < lim ? 1u : 0u) << (8 * b);
            });
        });
        if (tie) {
            Unroll<0, VB>::run([&](auto I) {
                constexpr int i = decltype(I)::value;
                bits[i] = 0;
                Unroll<0, 4>::run([&](auto B) {
                    constexpr int b = decltype(B)::value;
                    uint32_t r0, r1;
                    cell_philox<SMALL, 4 * i + b>(0u, rowid, t, seed, r0, r1);
                    const bool hit = f32 ? (r0 >> 8) < sp24 : draw53(r0, r1) < sp53;
                    bits[i] |= (hit ? 1u : 0u) << (8 * b);
                });
            });
        }
#pragma unroll
        for (int i = 0; i < VB; ++i) open[i] &= bits[i];
    }
}

// Each byte of `out` where `open` holds 0x01, else the byte of `keep`.
CAH_HD uint32_t open_select(uint32_t open, uint32_t out, uint32_t keep) {
    const uint32_t m = (open << 8) - open;  // 0xFF bytes where open
    return (out & m) | (keep & ~m);
}
