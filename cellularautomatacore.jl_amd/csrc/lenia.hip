// lenia.hip -- Lenia's MFMA implicit-GEMM field and FP32 FFT field.
#include <hipfft/hipfft.h>

#include <cstring>
#include <vector>

#include "lenia_launch.hpp"

void launch_lenia_mfma(int mode, int32_t ntiles, const LeniaMfmaArgs& a, hipStream_t st) {
    if (mode == LM_BF16) k_lenia_mfma<LM_BF16><<<ntiles, LM_NT, 0, st>>>(a);
    else if (mode == LM_F16X2) k_lenia_mfma<LM_F16X2><<<ntiles, LM_NT, 0, st>>>(a);
    else k_lenia_mfma<LM_F16><<<ntiles, LM_NT, 0, st>>>(a);
}

void launch_lenia_mfma_batch(int mode, int nblocks, const LeniaMfmaArgs* items,
                             const int32_t* block_item, const int32_t* block_base,
                             int32_t parity, int64_t dt, hipStream_t st) {
    if (mode == LM_BF16)
        k_lenia_mfma_batch<LM_BF16><<<nblocks, LM_NT, 0, st>>>(items, block_item, block_base, parity, dt);
    else if (mode == LM_F16X2)
        k_lenia_mfma_batch<LM_F16X2><<<nblocks, LM_NT, 0, st>>>(items, block_item, block_base, parity, dt);
    else
        k_lenia_mfma_batch<LM_F16><<<nblocks, LM_NT, 0, st>>>(items, block_item, block_base, parity, dt);
}

// The row offsets the MFMA loop visits: from the first row holding a nonzero
// weight, in pairs, up to the last such row (the last pair may take one row
// more, at most H + 1, whose weights are zero).
void lenia_mfma_pairs(int H, const double* wd, int32_t& dy0, int32_t& npairs) {
    const int K = 2 * H + 1;
    int lo = K, hi = -1;
    for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c)
            if (wd[(size_t)r * K + c] != 0.0) {
                lo = r < lo ? r : lo;
                hi = r;
            }
    if (hi < 0) {  // no weight: one pair of zero fragments
        dy0 = 0;
        npairs = 1;
        return;
    }
    dy0 = lo - H;
    npairs = (hi - lo + 2) / 2;
}

// Fragment (p, r) lane l holds A[m][k] for m = l % 16 and k = 8 g + i, g = l / 16,
// i = 0..7 (the operand order of v_mfma_f32_16x16x32_*; the B operand takes the
// same k order, so the sum is the same under any fixed k permutation):
// A[m][k] = W(16 (r - HC) + 8 (g & 1) + i - m, dy0 + 2 p + (g >> 1)), zero off
// the (2H + 1)^2 grid.
template <typename E>
static void fill_fragments(int H, const double* wd, int32_t dy0, int32_t npairs, std::vector<E>& hi,
                           std::vector<E>& lo) {
    const int K = 2 * H + 1, HC = LM_HC, NR = LM_NR;
    const size_t nfrag = (size_t)npairs * NR * 64;
    hi.assign(nfrag * 8, (E)0.0f);
    lo.assign(nfrag * 8, (E)0.0f);
    for (int p = 0; p < npairs; ++p)
        for (int r = 0; r < NR; ++r)
            for (int lane = 0; lane < 64; ++lane)
                for (int i = 0; i < 8; ++i) {
                    const int m = lane % 16, g = lane / 16;
                    const int dx = 16 * (r - HC) + 8 * (g & 1) + i - m, dy = dy0 + 2 * p + (g >> 1);
                    double w = 0.0;
                    if (dx >= -H && dx <= H && dy >= -H && dy <= H) w = wd[(size_t)(dy + H) * K + (dx + H)];
                    const size_t at = (((size_t)p * NR + r) * 64 + lane) * 8 + i;
                    const E h = (E)(float)w;
                    hi[at] = h;
                    lo[at] = (E)(float)(w - (double)(float)h);
                }
}

void lenia_mfma_fragments(int mode, int H, const double* wd, int32_t dy0, int32_t npairs, std::vector<char>& buf) {
    const size_t plane = lenia_mfma_fragment_bytes(npairs);
    buf.assign(plane * 2, 0);
    if (mode == LM_BF16) {
        std::vector<__bf16> hi, lo;
        fill_fragments(H, wd, dy0, npairs, hi, lo);
        std::memcpy(buf.data(), hi.data(), plane);
        std::memcpy(buf.data() + plane, lo.data(), plane);
    } else {
        std::vector<_Float16> hi, lo;
        fill_fragments(H, wd, dy0, npairs, hi, lo);
        std::memcpy(buf.data(), hi.data(), plane);
        std::memcpy(buf.data() + plane, lo.data(), plane);
    }
}

// ---------------------------------------------------------------- FFT field

struct LeniaFFT {
    hipfftHandle fwd = 0, inv = 0;
    float2* spec = nullptr;
    float2* kspec = nullptr;
    float* field = nullptr;
    int n1 = 0, n2 = 0;
    int64_t nspec = 0;
    bool dry = false;  // buffers at fake addresses, no plans
};

__global__ void k_spec_conj_scale(float2* k, int64_t n, float scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) k[i] = make_float2(k[i].x * scale, -k[i].y * scale);
}

__global__ void k_spec_mul(float2* s, const float2* k, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 a = s[i], b = k[i];
        s[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
}

LeniaFFT* lenia_fft_make_dry(int n1, int n2, void* spec, void* kspec, float* field) {
    LeniaFFT* f = new LeniaFFT();
    f->n1 = n1;
    f->n2 = n2;
    f->nspec = (int64_t)(n1 / 2 + 1) * n2;
    f->spec = (float2*)spec;
    f->kspec = (float2*)kspec;
    f->field = field;
    f->dry = true;
    return f;
}

void lenia_fft_destroy(LeniaFFT* f) {
    if (!f) return;
    if (f->dry) {
        delete f;
        return;
    }
    if (f->fwd) hipfftDestroy(f->fwd);
    if (f->inv) hipfftDestroy(f->inv);
    for (void* p : {(void*)f->spec, (void*)f->kspec, (void*)f->field})
        if (p) (void)hipFree(p);  // teardown: nothing to report to
    delete f;
}

int lenia_fft_create(LeniaFFT** out, int n1, int n2, int64_t pitch, const int32_t* taps,
                     const double* w, int ntaps, hipStream_t st) {
    LeniaFFT* f = new LeniaFFT();
    f->n1 = n1;
    f->n2 = n2;
    const int nc = n1 / 2 + 1;
    f->nspec = (int64_t)nc * n2;
    auto bail = [&](int rc) {
        lenia_fft_destroy(f);
        return rc;
    };
    if (hipMalloc(&f->spec, f->nspec * sizeof(float2)) != hipSuccess ||
        hipMalloc(&f->kspec, f->nspec * sizeof(float2)) != hipSuccess ||
        hipMalloc(&f->field, (size_t)n1 * n2 * sizeof(float)) != hipSuccess)
        return bail(-4);
    // column-major (n1 fastest) = row-major (n2, n1); the state rows have a pitch
    int n[2] = {n2, n1};
    int in_embed[2] = {n2, (int)pitch}, spec_embed[2] = {n2, nc}, dense_embed[2] = {n2, n1};
    if (hipfftPlanMany(&f->fwd, 2, n, in_embed, 1, (int)pitch * n2, spec_embed, 1, nc * n2,
                       HIPFFT_R2C, 1) != HIPFFT_SUCCESS)
        return bail(-7);
    if (hipfftPlanMany(&f->inv, 2, n, spec_embed, 1, nc * n2, dense_embed, 1, n1 * n2, HIPFFT_C2R,
                       1) != HIPFFT_SUCCESS)
        return bail(-7);
    hipfftSetStream(f->fwd, st);
    hipfftSetStream(f->inv, st);
    // the kernel wrapped onto the lattice: kpad[mod(o, dims)] += w over the read
    // offsets (ext/CellularAutomataCoreFFTWExt.jl:17-24), laid out with the state pitch
    std::vector<float> kpad((size_t)pitch * n2, 0.0f);
    std::vector<double> acc((size_t)n1 * n2, 0.0);
    for (int k = 0; k < ntaps; ++k) {
        const int a = ((taps[2 * k] % n1) + n1) % n1, b = ((taps[2 * k + 1] % n2) + n2) % n2;
        acc[(size_t)b * n1 + a] += w[k];
    }
    for (int b = 0; b < n2; ++b)
        for (int a = 0; a < n1; ++a) kpad[(size_t)b * pitch + a] = (float)acc[(size_t)b * n1 + a];
    float* kdev = nullptr;
    if (hipMalloc(&kdev, kpad.size() * sizeof(float)) != hipSuccess) return bail(-4);
    if (hipMemcpy(kdev, kpad.data(), kpad.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(kdev);
        return bail(-7);
    }
    hipfftResult r = hipfftExecR2C(f->fwd, kdev, (hipfftComplex*)f->kspec);
    const hipError_t sync = hipStreamSynchronize(st);
    (void)hipFree(kdev);
    if (r != HIPFFT_SUCCESS || sync != hipSuccess) return bail(-7);
    // conj(rfft(kpad)) / (n1 n2): the inverse transform is unnormalised
    const float scale = 1.0f / ((float)n1 * (float)n2);
    k_spec_conj_scale<<<(unsigned)((f->nspec + 255) / 256), 256, 0, st>>>(f->kspec, f->nspec, scale);
    if (hipStreamSynchronize(st) != hipSuccess) return bail(-7);
    *out = f;
    return 0;
}

int lenia_fft_field(LeniaFFT* f, const float* cur, hipStream_t st) {
    // the state's stream may change after the plan was made (a server's states)
    if (hipfftSetStream(f->fwd, st) != HIPFFT_SUCCESS || hipfftSetStream(f->inv, st) != HIPFFT_SUCCESS)
        return -7;
    if (hipfftExecR2C(f->fwd, const_cast<float*>(cur), (hipfftComplex*)f->spec) != HIPFFT_SUCCESS)
        return -7;
    k_spec_mul<<<(unsigned)((f->nspec + 255) / 256), 256, 0, st>>>(f->spec, f->kspec, f->nspec);
    if (hipfftExecC2R(f->inv, (hipfftComplex*)f->spec, f->field) != HIPFFT_SUCCESS) return -7;
    return 0;
}

float* lenia_fft_buffer(LeniaFFT* f) { return f->field; }
