// The launch checks' models (csrc/launch_check.hpp) on planted mistakes: each
// case builds a launch whose geometry or buffers are wrong in one way and
// requires the model to report it, and the same launch made right to pass.
// Prints "ok" and exits 0 when every case behaves.
#include <cstdio>
#include <string>

#include "kernels_box.hpp"
#include "kernels_generic.hpp"
#include "kernels_life_tb.hpp"
#include "launch.hpp"
#include "lenia_launch.hpp"
#include "launch_check.hpp"

static int bad = 0;

static void expect(const char* name, bool should_pass, const lchk::Chk& c) {
    if (c.ok() != should_pass) {
        std::printf("%s: expected %s, got %s\n", name, should_pass ? "pass" : "failure",
                    c.ok() ? "pass" : c.err.c_str());
        ++bad;
    }
}

static Lat lat(int n1, int n2, int64_t pitch, int b1, int b2) {
    Lat L{};
    L.pitch = pitch;
    L.n1 = n1;
    L.n2 = n2;
    L.b1 = b1;
    L.b2 = b2;
    L.two_d = 1;
    return L;
}

int main() {
    const int P = CAH_BOUNDARY_PERIODIC, R = CAH_BOUNDARY_REFLECTING;
    // a 1000 x 300 Int8 lattice, pitch 1024: two buffers of the right size and
    // one a row short
    const Lat L = lat(1000, 300, 1024, R, R);
    auto* cur = (const int8_t*)lchk::dry_alloc(1024 * 300, "cur");
    auto* next = (int8_t*)lchk::dry_alloc(1024 * 300, "next");
    auto* short_next = (int8_t*)lchk::dry_alloc(1024 * 299, "short next");
    const int32_t nbands = (1000 + 247) / 248, rows = 16, chunks = (300 + 15) / 16;
    {
        lchk::Chk c("band walk, right");
        lchk::band_walk<1>(c, L, cur, next, nbands, rows, nbands * chunks, 4 * ((nbands * chunks + 3) / 4), 1, 3);
        expect("band walk, right", true, c);
    }
    {
        lchk::Chk c("band walk, one wave too few");
        lchk::band_walk<1>(c, L, cur, next, nbands, rows, nbands * chunks - 1, nbands * chunks, 1, 3);
        expect(c.kernel.c_str(), false, c);
    }
    {
        lchk::Chk c("band walk, a band too few");
        lchk::band_walk<1>(c, L, cur, next, nbands - 1, rows, (nbands - 1) * chunks, 4096, 1, 3);
        expect(c.kernel.c_str(), false, c);
    }
    {
        lchk::Chk c("band walk, grid too small");
        lchk::band_walk<1>(c, L, cur, next, nbands, rows, nbands * chunks, nbands * chunks - 4, 1, 3);
        expect(c.kernel.c_str(), false, c);
    }
    {
        lchk::Chk c("band walk, output buffer a row short");
        lchk::band_walk<1>(c, L, cur, short_next, nbands, rows, nbands * chunks, 4096, 1, 3);
        expect(c.kernel.c_str(), false, c);
    }
    {
        // a pitch below the row: the writer lanes pass it
        lchk::Chk c("band walk, pitch below the row");
        const Lat bad_pitch = lat(1000, 300, 992, R, R);
        lchk::band_walk<1>(c, bad_pitch, cur, next, nbands, rows, nbands * chunks, 4096, 1, 3);
        expect(c.kernel.c_str(), false, c);
    }
    {
        lchk::Chk c("pointer in no allocation");
        c.range((const void*)0x10, 0, 16, "stray");
        expect(c.kernel.c_str(), false, c);
    }
    {
        // radius-5 walk on a lattice whose rows the mapped walk may read beyond
        lchk::Chk c("LtL walk, right");
        const Lat T = lat(1000, 300, 1024, P, P);
        const int32_t nb2 = (1000 + 495) / 496;
        lchk::band_walk<2>(c, T, cur, next, nb2, 35, nb2 * ((300 + 34) / 35), 4096, 5, 11);
        expect("LtL walk, right", true, c);
    }
    // k_life_tb: the right grid, then one tile row short
    {
        lchk::Chk c("k_life_tb, right");
        lchk::life_tb(c, L, cur, next, 4, 32, TB_K, 16, (1000 + 127) / 128, (300 + 31) / 32);
        expect("k_life_tb, right", true, c);
        lchk::Chk d("k_life_tb, a tile row short");
        lchk::life_tb(d, L, cur, next, 4, 32, TB_K, 16, (1000 + 127) / 128, (300 + 31) / 32 - 1);
        expect(d.kernel.c_str(), false, d);
        lchk::Chk e("k_life_tb, too many steps");
        lchk::life_tb(e, L, cur, next, 4, 32, TB_K, TB_K + 1, (1000 + 127) / 128, (300 + 31) / 32);
        expect(e.kernel.c_str(), false, e);
    }
    // an LDS tile over 64 KiB, and a grid that does not cover the lattice
    {
        lchk::Chk c("tile kernel, LDS over 64 KiB");
        lchk::tile_kernel(c, L, cur, 1, 64, 16, 8, 16, 19, 70000);
        expect(c.kernel.c_str(), false, c);
        lchk::Chk d("tile kernel, grid short");
        lchk::tile_kernel(d, L, cur, 1, 64, 16, 8, 15, 19, 4096);
        expect(d.kernel.c_str(), false, d);
        lchk::Chk e("tile kernel, right");
        lchk::tile_kernel(e, L, cur, 1, 64, 16, 8, 16, 19, 4096);
        expect("tile kernel, right", true, e);
        // the general kernel's tile: halo 5 (columns from 64 bx - 8, 88 a row, 26 rows)
        const size_t lds = (size_t)gen_w(5) * gen_rows(5);
        lchk::Chk f("general tile, right");
        lchk::generic_tile(f, L, cur, 5, 16, 19, lds);
        expect("general tile, right", true, f);
        lchk::Chk g("general tile, LDS for a smaller halo");
        lchk::generic_tile(g, L, cur, 5, 16, 19, (size_t)gen_w(4) * gen_rows(4));
        expect(g.kernel.c_str(), false, g);
        lchk::Chk h("general tile, grid short");
        lchk::generic_tile(h, L, cur, 5, 16, 18, lds);
        expect(h.kernel.c_str(), false, h);
        lchk::Chk i("general tile, halo past 32");
        lchk::generic_tile(i, L, cur, 33, 16, 19, (size_t)gen_w(33) * gen_rows(33));
        expect(i.kernel.c_str(), false, i);
        // the window-sum kernel: 64 x 64 blocks, 16 x 5 of them over 1000 x 300
        lchk::Chk j("box tile, Moore(8), right");
        lchk::box_tile(j, L, cur, next, false, 8, 16, 5, box_lds(false, 8));
        expect("box tile, Moore(8), right", true, j);
        lchk::Chk k("box tile, Moore(32), right");
        lchk::box_tile(k, L, cur, next, false, 32, 16, 5, box_lds(false, 32));
        expect("box tile, Moore(32), right", true, k);
        lchk::Chk l("box tile, VonNeumann(8), right");
        lchk::box_tile(l, L, cur, next, true, 8, 16, 5, box_lds(true, 8));
        expect("box tile, VonNeumann(8), right", true, l);
        lchk::Chk m("box tile, VonNeumann(1), right");
        lchk::box_tile(m, L, cur, next, true, 1, 16, 5, box_lds(true, 1));
        expect("box tile, VonNeumann(1), right", true, m);
        lchk::Chk n("box tile, Moore radius past 32");
        lchk::box_tile(n, L, cur, next, false, 33, 16, 5, box_lds(false, 33));
        expect(n.kernel.c_str(), false, n);
        lchk::Chk o("box tile, VonNeumann radius past 8");
        lchk::box_tile(o, L, cur, next, true, 9, 16, 5, box_lds(true, 9));
        expect(o.kernel.c_str(), false, o);
        lchk::Chk q("box tile, LDS sized for von Neumann's smaller planes");
        lchk::box_tile(q, L, cur, next, true, 8, 16, 5, box_lds(true, 7));
        expect(q.kernel.c_str(), false, q);
        lchk::Chk u("box tile, grid a row of blocks short");
        lchk::box_tile(u, L, cur, next, false, 8, 16, 4, box_lds(false, 8));
        expect(u.kernel.c_str(), false, u);
        lchk::Chk v("box tile, output a row short");
        lchk::box_tile(v, L, cur, short_next, false, 8, 16, 5, box_lds(false, 8));
        expect(v.kernel.c_str(), false, v);
    }
    // k_lenia_mfma: 2048^2 FP32, H = 12, rows -11 .. 11 (12 pairs): fragments of
    // the right size, then planted mistakes
    {
        const Lat F = lat(2048, 2048, 2048, P, P);
        const float* fc = (const float*)lchk::dry_alloc((size_t)2048 * 2048 * 4, "lenia cur");
        float* fn = (float*)lchk::dry_alloc((size_t)2048 * 2048 * 4, "lenia next");
        const size_t fb = lenia_mfma_fragment_bytes(12);
        LeniaMfmaArgs a{};
        a.cur = fc;
        a.next = fn;
        a.L = F;
        a.H = 12;
        a.dy0 = -11;
        a.npairs = 12;
        a.nbx = 32;
        a.afrag = lchk::dry_alloc(fb, "fragments");
        lchk::Chk c("k_lenia_mfma, right");
        lchk::lenia_mfma(c, a, LM_F16, 32, 32, fb);
        expect("k_lenia_mfma, right", true, c);
        lchk::Chk d("k_lenia_mfma, split fragments missing their low half");
        lchk::lenia_mfma(d, a, LM_F16X2, 32, 32, fb);
        expect(d.kernel.c_str(), false, d);
        lchk::Chk f("k_lenia_mfma, a tile count a row that is not the lattice's");
        a.nbx = 31;
        lchk::lenia_mfma(f, a, LM_F16, 32, 32, fb);
        expect(f.kernel.c_str(), false, f);
        a.nbx = 32;
        lchk::Chk h("k_lenia_mfma, fragments for fewer pairs");
        lchk::lenia_mfma(h, a, LM_F16, 32, 32, lenia_mfma_fragment_bytes(11));
        expect(h.kernel.c_str(), false, h);
        lchk::Chk p("k_lenia_mfma, one row-offset pair too many (rows past the tile)");
        a.npairs = 13;
        lchk::lenia_mfma(p, a, LM_F16, 32, 32, lenia_mfma_fragment_bytes(13));
        expect(p.kernel.c_str(), false, p);
        a.npairs = 12;
        lchk::Chk g("k_lenia_mfma, pairs past the rows of a smaller H (H = 8)");
        a.H = 8;
        lchk::lenia_mfma(g, a, LM_F16, 32, 32, fb);
        expect(g.kernel.c_str(), false, g);
        a.H = 17;
        lchk::Chk e("k_lenia_mfma, H past the tile");
        lchk::lenia_mfma(e, a, LM_F16, 32, 32, fb);
        expect(e.kernel.c_str(), false, e);
    }
    // a batch table whose entries skip an item
    {
        const int32_t map[6] = {0, 0, 1, 1, 2, 2}, base[3] = {0, 2, 4};
        lchk::Chk c("item table, right");
        lchk::item_table(c, map, base, 6, 3, [](int64_t) { return (int64_t)2; });
        expect("item table, right", true, c);
        const int32_t map2[6] = {0, 0, 2, 2, 2, 2};
        lchk::Chk d("item table, an item skipped");
        lchk::item_table(d, map2, base, 6, 3, [](int64_t) { return (int64_t)2; });
        expect(d.kernel.c_str(), false, d);
    }
    std::printf(bad ? "%d failures\n" : "ok\n", bad);
    return bad ? 1 : 0;
}
