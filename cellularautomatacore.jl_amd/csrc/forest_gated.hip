// forest_gated.hip -- instantiations and launcher of the gated forest-fire
// kernel (Stochastic and Checkerboard gates, masks).
#include "launch.hpp"

template <bool SMALL, bool TRI>
static void forest_gated_s(int prec, int blocks, const ForestArgs& a, hipStream_t st) {
    switch (prec) {
    case 0: k_forest_gated<SMALL, 0, TRI><<<blocks, 256, 0, st>>>(a); break;
    case 1: k_forest_gated<SMALL, 1, TRI><<<blocks, 256, 0, st>>>(a); break;
    case 2: k_forest_gated<SMALL, 2, TRI><<<blocks, 256, 0, st>>>(a); break;
    default: k_forest_gated<SMALL, 3, TRI><<<blocks, 256, 0, st>>>(a); break;
    }
}

void launch_forest_gated(bool small, bool tri, int blocks, const ForestArgs& a, hipStream_t st) {
    const int prec = (a.p_f32 ? 1 : 0) | (a.f_f32 ? 2 : 0);
    if (small) {
        if (tri) forest_gated_s<true, true>(prec, blocks, a, st);
        else forest_gated_s<true, false>(prec, blocks, a, st);
    } else {
        if (tri) forest_gated_s<false, true>(prec, blocks, a, st);
        else forest_gated_s<false, false>(prec, blocks, a, st);
    }
}
