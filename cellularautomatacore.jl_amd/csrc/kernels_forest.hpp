// kernels_forest.hpp -- the forest fire with VonNeumann(1): the band walk of
// kernels_life.hpp with one Philox4x32-10 draw per cell.
//
//   k_forest_vn1    Synchronous, no mask
//   k_forest_gated  any gate (Synchronous, Stochastic, Checkerboard), with or
//                   without a mask
#pragma once
#include "band.hpp"
#include "gate.hpp"

struct ForestArgs {
    const int8_t* cur;
    int8_t* next;
    Lat L;
    int32_t cv1, cv2;
    int32_t nbands, rows, nwaves;
    uint64_t seed;
    int64_t t;
    double p, f;
    int32_t p_f32, f_f32;
    BandGate g;  // k_forest_gated only
};

// The forest fire (README.md:56-60) on four cells at once.  Each cell forms
// exactly one Philox draw, k = 1 (growth) for an empty cell and k = 2 (lightning)
// otherwise; it decides the cell only where the reference rule draws it.  The
// FP64 test u < p with u = x * 2^-53 (x the 53-bit draw) is made exactly as the
// integer test x < thr53(p), the Float32 one as x24 < thr24(p) (device.hpp).
// PREC says which of p (bit 0) and f (bit 1) are Float32 and use the 24-bit draw.
// The cell is rowid + J, and k (1 or 2) comes in as c3 of cell_philox
// (device.hpp).  With both p and f in FP64 the walk tests the draw's high word
// against hi53 (device.hpp) and repeats a lane's draws through forest_draw only
// on a tie, so the last round forms one word and the test is one 32-bit compare.
template <bool SMALL, int PREC, int J>
CAH_DEV bool forest_draw(uint32_t c3, uint64_t rowid, int64_t t, uint64_t seed, uint64_t tp53,
                         uint64_t tf53, uint32_t tp24, uint32_t tf24) {
    const bool empty = c3 == 1u;
    uint32_t c0, c1;
    cell_philox<SMALL, J>(c3, rowid, t, seed, c0, c1);
    if constexpr (PREC == 0) {
        return draw53(c0, c1) < (empty ? tp53 : tf53);
    } else if constexpr (PREC == 3) {
        return (c0 >> 8) < (empty ? tp24 : tf24);
    } else {
        const bool hit64 = draw53(c0, c1) < (empty ? tp53 : tf53);
        const bool hit32 = (c0 >> 8) < (empty ? tp24 : tf24);
        return (empty ? (PREC & 1) : (PREC & 2)) ? hit32 : hit64;
    }
}

// The rule's result for four cells (one byte each), from 0/1 bytes: e4 empty,
// f4 burning, nf4 a VonNeumann neighbour burning, hit4 the cell's draw hit.  An
// empty cell grows a tree on a hit; a burning cell burns out; any other cell is a
// tree, which catches fire (2) beside a fire or on a lightning hit, else stays 1.
CAH_DEV uint32_t forest_bytes(uint32_t e4, uint32_t f4, uint32_t nf4, uint32_t hit4) {
    const uint32_t em = (e4 << 8) - e4, fm = (f4 << 8) - f4;  // 0xFF bytes where set
    const uint32_t tree = 0x01010101u + (nf4 | hit4);
    return ((em & hit4) | (~em & tree)) & ~fm;
}

// Burning (== 2) and empty (== 0) cells of four bytes, as 0x01 per byte.  TRI: the
// walk reads only 0, 1 and 2 (the lattice has held nothing else since its last
// step or scan, and no Constant edge holds another value), so bit 1 alone marks a
// fire and a zero pair of low bits an empty cell; otherwise the exact byte test.
template <bool TRI>
CAH_DEV uint32_t forest_fire4(uint32_t x) {
    if constexpr (TRI) return (x >> 1) & 0x01010101u;
    else return bytes_eq(x, 0x02020202u);
}
template <bool TRI>
CAH_DEV uint32_t forest_empty4(uint32_t x) {
    if constexpr (TRI) return ~(x | (x >> 1)) & 0x01010101u;
    else return bytes_eq(x, 0u);
}

// The walk of one wave down rows [ys, ye) of its band, as life_walk: rows in a
// ring of 6 register slots (row r in slot (r - ys + 1) mod 6), each load writing
// its final slot, 3 rows in flight.  GATED (k_forest_gated): a cell the mask or
// the gate keeps is written back unchanged; the mask of row y is loaded right
// after row y + 1, into a second ring (row r in slot (r - ys) mod 6).
template <int VB, bool SMALL, int PREC, bool FAST, bool TRI, bool GATED>
CAH_DEV void forest_walk(const ForestArgs& a, int32_t x, int32_t ys, int32_t ye, bool writer) {
    constexpr int S = 6;
    const int8_t cv1 = (int8_t)a.cv1, cv2 = (int8_t)a.cv2;
    const uint32_t fill1 = 0x01010101u * (uint8_t)cv1, fill2 = 0x01010101u * (uint8_t)cv2;
    const LaneMap lm = lane_map<VB>(a.L, x);
    const __amdgpu_buffer_rsrc_t rcur = row_rsrc(a.cur), rnext = row_rsrc(a.next);
    const int32_t last = ye;  // the last row any output of the chunk reads
    auto load = [&](int32_t y, uint32_t (&X)[VB]) {
        if constexpr (FAST) load_mapped<VB>(rcur, a.L, lm, fill1, fill2, y, X);
        else load_row<VB>(a.cur, a.L, cv1, cv2, x, y, X);
    };
    const uint64_t tp53 = thr53(a.p), tf53 = thr53(a.f);
    const uint32_t tp24 = thr24((float)a.p), tf24 = thr24((float)a.f);
    const uint32_t hp53 = hi53(tp53), hf53 = hi53(tf53);
    const bool masked = GATED && a.g.mask != nullptr;
    const __amdgpu_buffer_rsrc_t rmask = row_rsrc(reinterpret_cast<const int8_t*>(a.g.mask));
    uint32_t mring[GATED ? S : 1][VB];
    auto mload = [&](int32_t r, uint32_t (&X)[VB]) {
        if constexpr (GATED) {
            if (masked) mask_row<VB, FAST>(a.g.mask, rmask, a.L, lm.src, x, r, X);
        }
    };
    uint64_t sp53 = 0;
    uint32_t sp24 = 0;
    if constexpr (GATED) {
        sp53 = thr53(a.g.sp);
        sp24 = thr24((float)a.g.sp);
    }

    uint32_t raw[S][VB], fire[S][VB];
    load(ys - 1, raw[0]);
    load(ys, raw[1]);
    Unroll<2, S>::run([&](auto K) {
        constexpr int k = decltype(K)::value;
        load(min(ys + k - 1, last), raw[k]);
        mload(min(ys + k - 2, ye - 1), mring[GATED ? k - 2 : 0]);
    });
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        fire[0][i] = forest_fire4<TRI>(raw[0][i]);
        fire[1][i] = forest_fire4<TRI>(raw[1][i]);
    }

    int32_t y = ys;
    auto body = [&](auto PH, bool more) {
        constexpr int ph = decltype(PH)::value;
        constexpr int s_dn = (ph + 2) % S, s_c = (ph + 1) % S, s_up = ph % S;
#pragma unroll
        for (int i = 0; i < VB; ++i) fire[s_dn][i] = forest_fire4<TRI>(raw[s_dn][i]);
        uint32_t E[VB + 2];
        E[0] = from_prev_lane(fire[s_c][VB - 1]);
        E[VB + 1] = from_next_lane(fire[s_c][0]);
#pragma unroll
        for (int i = 0; i < VB; ++i) E[1 + i] = fire[s_c][i];
        uint32_t out[VB], e4[VB], hit4[VB];
        const uint64_t rowid = (uint64_t)y * (uint64_t)a.L.n1 + (uint64_t)x + 1;
        bool tie = false;
        Unroll<0, VB>::run([&](auto I) {
            constexpr int i = decltype(I)::value;
            e4[i] = forest_empty4<TRI>(raw[s_c][i]);
            const uint32_t k4 = 0x02020202u - e4[i];  // the draw's k: 1 empty, 2 otherwise
            hit4[i] = 0;
            Unroll<0, 4>::run([&](auto B) {
                constexpr int b = decltype(B)::value;
                const uint32_t k = (k4 >> (8 * b)) & 0xFFu;
                bool hit;
                if constexpr (PREC == 0) {  // the high word decides all but a tie
                    uint32_t r0, r1;
                    cell_philox<SMALL, 4 * i + b>(k, rowid, a.t, a.seed, r0, r1);
                    const uint32_t lim = k == 1u ? hp53 : hf53;
                    hit = r0 < lim;
                    tie |= r0 == lim;
                } else {
                    hit = forest_draw<SMALL, PREC, 4 * i + b>(k, rowid, a.t, a.seed, tp53, tf53, tp24, tf24);
                }
                hit4[i] |= (hit ? 1u : 0u) << (8 * b);
            });
        });
        if constexpr (PREC == 0) {
            if (tie) {  // about one lane in 2^29 rows: every draw of the lane again, in full
                Unroll<0, VB>::run([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    const uint32_t k4 = 0x02020202u - e4[i];
                    hit4[i] = 0;
                    Unroll<0, 4>::run([&](auto B) {
                        constexpr int b = decltype(B)::value;
                        const bool hit = forest_draw<SMALL, 0, 4 * i + b>((k4 >> (8 * b)) & 0xFFu, rowid, a.t,
                                                                          a.seed, tp53, tf53, tp24, tf24);
                        hit4[i] |= (hit ? 1u : 0u) << (8 * b);
                    });
                });
            }
        }
        Unroll<0, VB>::run([&](auto I) {
            constexpr int i = decltype(I)::value;
            // VonNeumann(1): left, right, up, down
            const uint32_t nf = funnel(E[1 + i], E[i], 24) | funnel(E[2 + i], E[1 + i], 8) |
                                fire[s_up][i] | fire[s_dn][i];
            out[i] = forest_bytes(e4[i], fire[s_c][i], nf, hit4[i]);
        });
        if constexpr (GATED) {
            if (band_gated(a.g)) {
                uint32_t open[VB];
                band_open<VB, SMALL>(a.g, mring[ph], rowid, y, a.seed, a.t, sp53, sp24, open);
#pragma unroll
                for (int i = 0; i < VB; ++i) out[i] = open_select(open[i], out[i], raw[s_c][i]);
            }
        }
        if (writer) {
            if constexpr (FAST) store_mapped<VB>(rnext, a.L, x, y, out);
            else store_row<VB>(a.next, a.L, x, y, out);
        }
        // the slot of row y - 1 now takes row y + 5 (clamped past the chunk), and
        // the mask ring the mask of row y + 4
        if (more) {
            load(min(y + 5, last), raw[s_up]);
            mload(min(y + 4, ye - 1), mring[GATED ? (ph + 4) % S : 0]);
        }
        ++y;
    };
    while (y + S <= ye) Unroll<0, S>::run([&](auto PH) { body(PH, true); });
    Unroll<0, S>::run([&](auto PH) {
        if (y < ye) body(PH, y + 4 < ye);
    });
}

template <int VB, bool SMALL, int PREC, bool TRI, bool GATED = false>
CAH_DEV void forest_band(const ForestArgs& a, int wid) {
    const BandPart b = band_part<VB>(a.nbands, a.rows, a.L, wid, threadIdx.x & 63);
    if (mapped_ok<VB>(a.L, 3)) forest_walk<VB, SMALL, PREC, true, TRI, GATED>(a, b.x, b.ys, b.ye, b.writer);
    else forest_walk<VB, SMALL, PREC, false, TRI, GATED>(a, b.x, b.ys, b.ye, b.writer);
}

// Waves in the XCD-aware block order (wave_id_xcd): a wave's neighbouring bands,
// which share the cache lines at the band edges, run on the same XCD.
template <int VB, bool SMALL, int PREC, bool TRI>
__global__ __launch_bounds__(256) void k_forest_vn1(ForestArgs a) {
    const int wid = wave_id_xcd();
    if (wid >= a.nwaves) return;
    forest_band<VB, SMALL, PREC, TRI>(a, wid);
}

// The forest fire under the gate and mask of a.g, one sweep (4 cells a lane).
template <bool SMALL, int PREC, bool TRI>
__global__ __launch_bounds__(256) void k_forest_gated(ForestArgs a) {
    const int wid = wave_id_xcd();
    if (wid >= a.nwaves) return;
    forest_band<1, SMALL, PREC, TRI, true>(a, wid);
}

// ---------------------------------------------------------------- batched ensembles
//
// One launch advances every trajectory of a batch by one step.  The waves of all
// trajectories form one flat list: wave w belongs to item wave_item[w] and is
// that item's wave (w - wave_base[item]).  Each item is the trajectory's own
// argument block (buffers for parity 0, lattice, seed, band layout, p and f, and
// its time base); the launch adds the step offset dt and, when tbase is not
// null, the time base *tbase read on the device (so that a captured sequence of
// launches serves any starting time) to the item's t, and swaps the buffers when
// parity is set.  A launch over the first nwaves waves advances
// only the items those waves belong to, so items ordered by steps left can stop
// at different steps of one sequence of launches.

template <int VB, bool SMALL, int PREC, bool TRI>
__global__ __launch_bounds__(256) void k_forest_batch(const ForestArgs* __restrict__ items,
                                                      const int32_t* __restrict__ wave_item,
                                                      const int32_t* __restrict__ wave_base,
                                                      int32_t nwaves, int32_t parity, int64_t dt,
                                                      const int64_t* __restrict__ tbase) {
    const int wid = wave_id();
    if (wid >= nwaves) return;
    const int k = __builtin_amdgcn_readfirstlane(wave_item[wid]);
    ForestArgs a = items[k];
    if (parity) {
        const int8_t* c = a.cur;
        a.cur = a.next;
        a.next = const_cast<int8_t*>(c);
    }
    a.t += dt + (tbase ? *tbase : 0);
    forest_band<VB, SMALL, PREC, TRI>(a, wid - wave_base[k]);
}
