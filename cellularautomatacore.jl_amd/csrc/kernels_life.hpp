// kernels_life.hpp -- Life and Generations with Moore(R <= 6).
//
// Each lane owns 4*VB consecutive cells of a row and walks down a column band,
// keeping the vertical (2R+1)-row running count in registers and forming the
// horizontal (2R+1)-wide sum with byte-parallel (SWAR) adds and funnel shifts;
// the rule is a union of count intervals tested four cells at a time.  A wave
// covers 64 lanes and writes the middle 62 (the end lanes are guards).
//
//   k_life_swar   Life, Synchronous, no mask
//   k_life_gated  Life or Generations under any gate (Synchronous, Stochastic,
//                 Checkerboard), with or without a mask
#pragma once
#include "band.hpp"
#include "gate.hpp"

constexpr int LIFE_MAX_IV = 2;

struct LifeArgs {
    const int8_t* cur;
    int8_t* next;
    Lat L;
    int32_t cv1, cv2;
    int32_t nbands, rows, nwaves;
    int32_t nb, ns;  // birth / survive interval counts
    // interval [lo, hi] of counts: bytes n with n + (128 - lo) >= 128 and
    // n + (127 - hi) < 128, replicated over the four bytes
    uint32_t b_lo[LIFE_MAX_IV], b_hi[LIFE_MAX_IV], s_lo[LIFE_MAX_IV], s_hi[LIFE_MAX_IV];
    // k_life_gated only: the rule (gen_C = 0 Life, else Generations with gen_C
    // states: gen_k = max(0, 129 - C) and gen_dying = (C > 2 ? 2 : 0), each
    // replicated over the four bytes), the gate and mask, the seed and the step
    int32_t gen_C;
    uint32_t gen_k, gen_dying;
    BandGate g;
    uint64_t seed;
    int64_t t;
};

// Runs of set bits of a count mask -> [lo, hi] intervals, as the byte-parallel
// constants of life_rule and gen_rule (host side).  Slots left over hold 0, which
// matches no count (n & ~n = 0).  -1 when the mask has more than LIFE_MAX_IV runs.
inline int life_intervals(uint64_t m, uint32_t* lo, uint32_t* hi) {
    for (int k = 0; k < LIFE_MAX_IV; ++k) {
        lo[k] = 0;
        hi[k] = 0;
    }
    int k = 0;
    for (int c = 0; c < 64;) {
        if (!((m >> c) & 1u)) {
            ++c;
            continue;
        }
        int e = c;
        while (e + 1 < 64 && ((m >> (e + 1)) & 1u)) ++e;
        if (k == LIFE_MAX_IV) return -1;
        lo[k] = (uint32_t)(128 - c) * 0x01010101u;
        hi[k] = (uint32_t)(127 - e) * 0x01010101u;
        ++k;
        c = e + 1;
    }
    return k;
}

// Horizontal (2R+1)-wide byte sums of the dwords E (G guard dwords on each side
// of the lane's VB dwords): W[i] = sum over d in [-R, R] of the bytes at offset d.
// Offsets d = 4q + rho with the same rho share one funnel shift of a range sum
// (no byte sum reaches 256, so shifting sums equals summing shifts).
template <int R, int VB, int G>
CAH_DEV void hsum(const uint32_t (&E)[VB + 2 * G], uint32_t (&W)[VB]) {
    // prefix sums over E (bytes stay below 256: at most (VB+2G)(2R+1) <= 88)
    uint32_t P[VB + 2 * G];
    P[0] = E[0];
#pragma unroll
    for (int j = 1; j < VB + 2 * G; ++j) P[j] = P[j - 1] + E[j];
    auto range = [&](int a, int b) -> uint32_t { return a > 0 ? P[b] - P[a - 1] : P[b]; };
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t w = 0;
#pragma unroll
        for (int rho = 0; rho < 4; ++rho) {
            const int qlo = -fdiv(R + rho, 4);  // ceil((-R - rho) / 4)
            const int qhi = fdiv(R - rho, 4);
            if (qlo > qhi) continue;
            const uint32_t a0 = range(G + i + qlo, G + i + qhi);
            if (rho == 0) {
                w += a0;
            } else {
                const uint32_t a1 = range(G + i + 1 + qlo, G + i + 1 + qhi);
                w += funnel(a1, a0, 8 * rho);
            }
        }
        W[i] = w;
    }
}

// The new states of VB dwords (four cells each) from the live counts n (centre
// left out) and the centre cells c (bytes 0 or 1): bit n of survive for a live
// cell, of birth otherwise, each a union of count intervals.
// NIV intervals per mask are tested; unused ones are neutral (lo = 128 never
// matches a count below 128).
template <int VB, int NIV>
CAH_HD void life_rule(const LifeArgs& a, const uint32_t (&n)[VB], const uint32_t (&c)[VB],
                       uint32_t (&out)[VB]) {
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t inb = 0, ins = 0;
#pragma unroll
        for (int k = 0; k < NIV; ++k) {
            inb |= (n[i] + a.b_lo[k]) & ~(n[i] + a.b_hi[k]);
            ins |= (n[i] + a.s_lo[k]) & ~(n[i] + a.s_hi[k]);
        }
        const uint32_t sel = c[i] << 7;  // 0x80 in the bytes of live cells
        out[i] = (((sel & ins) | (~sel & inb)) & 0x80808080u) >> 7;
    }
}

// local_rule(::Generations) (src/rules.jl:446-456) on four cells a dword: c the
// cells as stored, n their live counts (neighbours equal to 1).  A dead cell (0)
// is born by the birth intervals; a live one (1) survives by the survive
// intervals or starts dying (2; 0 with two states); any other value s moves to
// s + 1 while s + 1 < C and to 0 after.  s + 1 < C is s <= C - 2 as a signed
// byte: every negative byte, and a byte s7 < 128 with s7 + (129 - C) < 128 (C is
// at most 129 here: a larger C passes every byte).  Exact for every byte value.
template <int VB, int NIV>
CAH_HD void gen_rule(const LifeArgs& a, const uint32_t (&n)[VB], const uint32_t (&c)[VB], uint32_t (&out)[VB]) {
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t inb = 0, ins = 0;
#pragma unroll
        for (int k = 0; k < NIV; ++k) {
            inb |= (n[i] + a.b_lo[k]) & ~(n[i] + a.b_hi[k]);
            ins |= (n[i] + a.s_lo[k]) & ~(n[i] + a.s_hi[k]);
        }
        const uint32_t born = (inb & 0x80808080u) >> 7, keep = (ins & 0x80808080u) >> 7;
        const uint32_t z = bytes_eq(c[i], 0u), o = bytes_eq(c[i], 0x01010101u);
        const uint32_t zm = (z << 8) - z, om = (o << 8) - o, km = (keep << 8) - keep;
        const uint32_t c7 = c[i] & 0x7F7F7F7Fu, sign = c[i] & 0x80808080u;
        const uint32_t inc = (c7 + 0x01010101u) ^ sign;  // s + 1 in each byte, no carry between bytes
        const uint32_t le = ((sign | ~(c7 + a.gen_k)) & 0x80808080u) >> 7;
        const uint32_t older = inc & ((le << 8) - le);
        const uint32_t live = keep | (~km & a.gen_dying);
        out[i] = (zm & born) | (om & live) | (~(zm | om) & older);
    }
}

// gen_rule for C <= 4 on cells that are all 0, 1, 2 or 3 (bits b1 b0 of each
// byte): a dead cell (0) is born by the birth intervals, a live one (1) survives
// by the survive intervals or starts dying (2 when C > 2, else 0), a dying cell
// s (2 or 3) moves to s + 1 while s + 1 < C (only 2 -> 3 with C = 4) and to 0
// after.
template <int VB, int NIV>
CAH_HD void gen_rule_small(const LifeArgs& a, const uint32_t (&n)[VB], const uint32_t (&c)[VB],
                           uint32_t (&out)[VB]) {
    const uint32_t c4 = a.gen_C == 4 ? 0x01010101u : 0u;  // 2 -> 3
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        uint32_t inb = 0, ins = 0;
#pragma unroll
        for (int k = 0; k < NIV; ++k) {
            inb |= (n[i] + a.b_lo[k]) & ~(n[i] + a.b_hi[k]);
            ins |= (n[i] + a.s_lo[k]) & ~(n[i] + a.s_hi[k]);
        }
        const uint32_t born = (inb & 0x80808080u) >> 7, keep = (ins & 0x80808080u) >> 7;
        const uint32_t b0 = c[i] & 0x01010101u, b1 = (c[i] >> 1) & 0x01010101u;
        const uint32_t one = b0 & ~b1, zero = ~(b0 | b1) & 0x01010101u, to3 = ~b0 & b1 & c4;
        const uint32_t dying = one & ~keep & (a.gen_dying >> 1);
        out[i] = (zero & born) | (one & keep) | (dying << 1) | to3 | (to3 << 1);
    }
}

// The walk of one wave down rows [ys, ye) of its band (ys, ye wave-uniform), or
// up them when `up`: the walk runs over rows y' = ys + ye - 1 - y (the Moore sums
// are symmetric in the row offset).  FAST: the lattice suits load_mapped, so every
// row is one vector load per lane with the boundary applied branch-free;
// otherwise rows and bytes go through the general boundary map.  Rows live in a
// ring of 2P = 4R+2 register slots, row r in slot (r - ys + R) mod 2P, and each
// load writes its final slot directly: the slot of the row that just left the
// window.  P = 2R+1 rows are in flight per wave.
//
// GATED (k_life_gated): the ring keeps the cells as stored and the counts take
// the bytes equal to 1 (for R <= 2 converted once, into a ring of their own
// beside it; wider walks convert a row when it enters and when it leaves the
// count, to stay within the registers; BIN: the cells are 0 or 1 and count as
// they are, and a kept cell stays 0 or 1); the rule is Life or Generations; a cell
// the mask or the gate keeps is written back unchanged.  The mask rows go
// through a further ring (row r in slot (r - ys) mod 2P): the mask of row y is
// loaded right after row y + R, the row whose wait precedes row y's output, so
// it adds no wait of its own (loads complete in order).  SMALL (gated walk only):
// every value read is 0, 1, 2 or 3 and the rule is Life or Generations with
// C <= 4, so a cell == 1 is bit 0 without bit 1 and the rule takes
// gen_rule_small.  MASK
// (gated walk): the walk may have a mask; without, it keeps no mask ring (2P
// registers a lane's dword fewer: the radius-5 walk then holds three waves a
// SIMD instead of two).
template <int R, int VB, bool BIN, int NIV, bool FAST, bool GATED, bool SMALL = false, bool MASK = true>
CAH_DEV void life_walk(const LifeArgs& a, int32_t x, int32_t ys, int32_t ye, bool writer, bool up) {
    constexpr int P = 2 * R + 1, S = 2 * P;
    constexpr int G = (R + 3) / 4;  // guard dwords needed on each side
    static_assert(G <= VB, "a lane must hold the guard dwords of its neighbours");
    const int8_t cv1 = (int8_t)a.cv1, cv2 = (int8_t)a.cv2;
    const uint32_t fill1 = 0x01010101u * (uint8_t)cv1, fill2 = 0x01010101u * (uint8_t)cv2;
    const LaneMap lm = lane_map<VB>(a.L, x);
    const __amdgpu_buffer_rsrc_t rcur = row_rsrc(a.cur), rnext = row_rsrc(a.next);
    const int32_t last = ye + R - 1;  // the last row any output of the chunk reads
    const int32_t mirror = ys + ye - 1;
    auto load = [&](int32_t y, uint32_t (&X)[VB]) {
        const int32_t ya = up ? mirror - y : y;
        if constexpr (FAST) load_mapped<VB>(rcur, a.L, lm, fill1, fill2, ya, X);
        else load_row<VB>(a.cur, a.L, cv1, cv2, x, ya, X);
    };
    auto ones = [&](uint32_t (&X)[VB]) {
        if constexpr (!BIN && !GATED) {
#pragma unroll
            for (int i = 0; i < VB; ++i) X[i] = bytes_eq(X[i], 0x01010101u);
        }
    };
    // a cell's contribution to the counts (the ring holds it converted unless
    // GATED); ONES: the gated walk's converted rows, in slots of their own
    constexpr bool ONES = GATED && !BIN && !SMALL && R <= 2;
    uint32_t ring[S][VB];
    uint32_t one[ONES ? S : 1][VB];
    auto live = [&](uint32_t v) -> uint32_t {
        if constexpr (GATED && SMALL) return v & ~(v >> 1) & 0x01010101u;
        else if constexpr (GATED && !BIN) return bytes_eq(v, 0x01010101u);
        else return v;
    };
    // the counts of ring slot k: converted now (into its ONES slot), or again
    auto enter = [&](int k, int i) -> uint32_t {
        if constexpr (ONES) return one[k][i] = live(ring[k][i]);
        else return live(ring[k][i]);
    };
    auto leave = [&](int k, int i) -> uint32_t {
        if constexpr (ONES) return one[k][i];
        else return live(ring[k][i]);
    };
    // the mask (GATED): its rows [ys, ye) only, the clamp repeats the chunk's last
    constexpr bool MR = GATED && MASK;  // a mask ring
    const bool masked = MR && a.g.mask != nullptr;
    const __amdgpu_buffer_rsrc_t rmask = row_rsrc(reinterpret_cast<const int8_t*>(a.g.mask));
    uint32_t mring[MR ? S : 1][VB];
    auto mload = [&](int32_t r, uint32_t (&X)[VB]) {
        if constexpr (MR) {
            if (masked) mask_row<VB, FAST>(a.g.mask, rmask, a.L, lm.src, x, up ? mirror - r : r, X);
        }
    };
    uint64_t sp53 = 0;
    uint32_t sp24 = 0;
    if constexpr (GATED) {
        sp53 = thr53(a.g.sp);
        sp24 = thr24((float)a.g.sp);
    }
    const bool small = (uint64_t)a.L.n1 * (uint64_t)a.L.n2 < 0xFFFFFFFFull;

    // rows ys-R .. ys+R-1 counted now; rows ys+R .. ys+R+P in flight, each with the
    // mask of the row R above it
    Unroll<0, 2 * R>::run([&](auto K) { load(ys - R + decltype(K)::value, ring[decltype(K)::value]); });
    Unroll<2 * R, S>::run([&](auto K) {
        constexpr int k = decltype(K)::value;
        load(min(ys + k - R, last), ring[k]);
        mload(min(ys + k - 2 * R, ye - 1), mring[MR ? k - 2 * R : 0]);
    });
    uint32_t V[VB];
#pragma unroll
    for (int i = 0; i < VB; ++i) V[i] = 0;
    Unroll<0, 2 * R>::run([&](auto K) {
        constexpr int k = decltype(K)::value;
        ones(ring[k]);
#pragma unroll
        for (int i = 0; i < VB; ++i) V[i] += enter(k, i);
    });

    int32_t y = ys;
    auto body = [&](auto PH, bool more) {
        constexpr int ph = decltype(PH)::value;
        constexpr int s_new = (ph + 2 * R) % S;  // row y + R
        constexpr int s_c = (ph + R) % S;        // row y
        constexpr int s_old = ph % S;            // row y - R
        ones(ring[s_new]);
#pragma unroll
        for (int i = 0; i < VB; ++i) V[i] += enter(s_new, i);
        uint32_t E[VB + 2 * G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            E[j] = from_prev_lane(V[VB - G + j]);
            E[G + VB + j] = from_next_lane(V[j]);
        }
#pragma unroll
        for (int i = 0; i < VB; ++i) E[G + i] = V[i];
        uint32_t Wd[VB], n[VB], c1[VB], out[VB];
        hsum<R, VB, G>(E, Wd);
#pragma unroll
        for (int i = 0; i < VB; ++i) {
            c1[i] = leave(s_c, i);
            n[i] = Wd[i] - c1[i];  // Moore leaves the centre out
            // R = 6: counts up to (2R+1)^2 - 1 = 168 fit a byte, but the interval
            // tests take bytes below 128; a count of 128 or more is set to 127,
            // which no mask holds (counts 0..63)
            if constexpr (R >= 6) n[i] = (n[i] | (((n[i] >> 7) & 0x01010101u) * 0x7Fu)) & 0x7F7F7F7Fu;
        }
        const int32_t yr = up ? mirror - y : y;  // the row written
        if constexpr (GATED) {
            if (a.gen_C) {
                if constexpr (SMALL) gen_rule_small<VB, NIV>(a, n, ring[s_c], out);
                else gen_rule<VB, NIV>(a, n, ring[s_c], out);
            } else {
                life_rule<VB, NIV>(a, n, c1, out);
            }
            if (band_gated(a.g)) {
                uint32_t open[VB];
                const uint64_t rowid = (uint64_t)yr * (uint64_t)a.L.n1 + (uint64_t)(int64_t)x + 1;
                if (small) band_open<VB, true>(a.g, mring[MR ? ph % S : 0], rowid, yr, a.seed, a.t, sp53, sp24, open);
                else band_open<VB, false>(a.g, mring[MR ? ph % S : 0], rowid, yr, a.seed, a.t, sp53, sp24, open);
#pragma unroll
                for (int i = 0; i < VB; ++i) out[i] = open_select(open[i], out[i], ring[s_c][i]);
            }
        } else {
            life_rule<VB, NIV>(a, n, c1, out);
        }
        if (writer) {
            if constexpr (FAST) store_mapped<VB>(rnext, a.L, x, yr, out);
            else store_row<VB>(a.next, a.L, x, yr, out);
        }
#pragma unroll
        for (int i = 0; i < VB; ++i) V[i] -= leave(s_old, i);
        // the slot of row y - R now takes row y - R + 2P = y + R + P + 1
        // (clamped: rows past the chunk are never read, the clamp repeats the
        // last row), and the mask ring the mask of row y + P + 1
        if (more) {
            load(min(y + R + P + 1, last), ring[s_old]);
            mload(min(y + P + 1, ye - 1), mring[MR ? (ph + P + 1) % S : 0]);
        }
        ++y;
    };
    // whole blocks of 2P rows, then the remaining (< 2P) rows
    while (y + S <= ye) Unroll<0, S>::run([&](auto PH) { body(PH, true); });
    Unroll<0, S>::run([&](auto PH) {
        if (y < ye) body(PH, y + P + 1 < ye);
    });
}

// One wave's work: band `wid % nbands`, rows chunk `wid / nbands` (wid is
// wave-uniform, so the row bookkeeping stays in scalar registers).  Odd chunks
// walk up: the two chunks beside a chunk boundary then read the rows around it
// at the same time, so the second read finds them in L2.
template <int R, int VB, bool BIN, int NIV, bool GATED = false, bool SMALL = false, bool MASK = true>
CAH_DEV void life_band(const LifeArgs& a, int wid) {
    const BandPart b = band_part<VB>(a.nbands, a.rows, a.L, wid, threadIdx.x & 63);
    if (mapped_ok<VB>(a.L, 2 * R + 1))
        life_walk<R, VB, BIN, NIV, true, GATED, SMALL, MASK>(a, b.x, b.ys, b.ye, b.writer, b.up);
    else
        life_walk<R, VB, BIN, NIV, false, GATED, SMALL, MASK>(a, b.x, b.ys, b.ye, b.writer, b.up);
}

template <int R, int VB, bool BIN, int NIV>
__global__ __launch_bounds__(256) void k_life_swar(LifeArgs a) {
    const int wid = wave_id_xcd();
    if (wid >= a.nwaves) return;  // whole waves only
    life_band<R, VB, BIN, NIV>(a, wid);
}

// Life or Generations (a.gen_C) under the gate and mask of a.g, one sweep;
// NIV intervals per count mask; BIN: every cell read is 0 or 1; SMALL: every
// value read is 0, 1, 2 or 3 and a Generations rule has at most 4 states; MASK:
// a.g.mask may be set (false: it is null).
template <int R, int VB, int NIV, bool BIN, bool SMALL = false, bool MASK = true>
__global__ __launch_bounds__(256) void k_life_gated(LifeArgs a) {
    const int wid = wave_id_xcd();
    if (wid >= a.nwaves) return;
    life_band<R, VB, BIN, NIV, true, SMALL, MASK>(a, wid);
}

// ---------------------------------------------------------------- batched ensembles
//
// One launch advances every trajectory of a batch by one step.  The waves of all
// trajectories form one flat list: wave w belongs to item wave_item[w] and is
// that item's wave (w - wave_base[item]).  Each item is the trajectory's own
// argument block (buffers for parity 0, lattice, band layout); the buffer
// parity comes from the launch.

template <int R, int VB, bool BIN, int NIV>
__global__ __launch_bounds__(256) void k_life_batch(const LifeArgs* __restrict__ items,
                                                    const int32_t* __restrict__ wave_item,
                                                    const int32_t* __restrict__ wave_base,
                                                    int32_t nwaves, int32_t parity) {
    const int wid = wave_id();
    if (wid >= nwaves) return;
    const int k = __builtin_amdgcn_readfirstlane(wave_item[wid]);
    LifeArgs a = items[k];
    if (parity) {
        const int8_t* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<int8_t*>(c);
    }
    life_band<R, VB, BIN, NIV>(a, wid - wave_base[k]);
}
