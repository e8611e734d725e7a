// forest.hip -- instantiations and launchers of the single-lattice forest-fire
// kernels (the batched ones are in forest_batch.hip: the two units compile in
// parallel).
#include "launch.hpp"

template <int VB, bool SMALL, bool TRI>
static void forest_vs(int prec, int blocks, const ForestArgs& a, hipStream_t st) {
    switch (prec) {
    case 0: k_forest_vn1<VB, SMALL, 0, TRI><<<blocks, 256, 0, st>>>(a); break;
    case 1: k_forest_vn1<VB, SMALL, 1, TRI><<<blocks, 256, 0, st>>>(a); break;
    case 2: k_forest_vn1<VB, SMALL, 2, TRI><<<blocks, 256, 0, st>>>(a); break;
    default: k_forest_vn1<VB, SMALL, 3, TRI><<<blocks, 256, 0, st>>>(a); break;
    }
}

template <bool TRI>
static void forest_t(int vb, bool small, int prec, int blocks, const ForestArgs& a, hipStream_t st) {
    if (vb == 2) {
        if (small) forest_vs<2, true, TRI>(prec, blocks, a, st);
        else forest_vs<2, false, TRI>(prec, blocks, a, st);
    } else {
        if (small) forest_vs<1, true, TRI>(prec, blocks, a, st);
        else forest_vs<1, false, TRI>(prec, blocks, a, st);
    }
}

void launch_forest(int vb, bool small, bool tri, int blocks, const ForestArgs& a, hipStream_t st) {
    const int prec = (a.p_f32 ? 1 : 0) | (a.f_f32 ? 2 : 0);
    if (tri) forest_t<true>(vb, small, prec, blocks, a, st);
    else forest_t<false>(vb, small, prec, blocks, a, st);
}
