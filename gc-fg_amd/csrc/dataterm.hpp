// Colour data term on the device (the stage behind DataTerm.h).
//
// Replaces the reference's 256^3 Gaussian-KDE colour model
// (DataTermKernels.cu:19-150, DataTerm.cpp:28-137) with:
//   per class k (1 = foreground, 2 = background seeds) and channel c a 16-bin
//   histogram n_{k,c}[v >> 4] of the seed pixels,
//   p_k(I) = prod_c (n_{k,c}[I_c >> 4] + 1) / (N_k + 16),
//   pos = llround(64 * -ln p_bg(I)), neg = llround(64 * -ln p_fg(I)),
//   beta = sum_{4-neighbour pairs} |I_p - I_q|^2 / #pairs (exact int64 sum),
//   w = llround(lambda * exp(-|I_p - I_q|^2 / (2 beta)))  (lambda if beta == 0).
// Device work: one stats pass (LDS histograms + int64 beta sum) and one links
// pass (t-link gather from a 2 x 4096 LUT, n-link step function from lambda
// integer thresholds). The LUT and thresholds are built on the host in double
// from the 96 histogram counts and beta, so the quantised capacities round
// exactly as a direct per-pixel double evaluation does.
//
// The reference's own colour model is kept as a second mode (GCFG_DT_MODEL_KDE,
// DataTermKernels.cu:19-31,61-83,116-150; DataTerm.cpp:97-137): a Gaussian
// kernel density over the seed colours,
//   g[k] = (float)(1 / (sigma * 2.507) * expf(-0.5 (k/sigma)^2)),  k = 0..100,
//   p(I) = sum_j g[floor(|I - s_j|)]  over seed colours s_j with |I - s_j| < 101,
//   pos = (int)(300000 * p),  neg = (int)(300000 * (1 - p))   (p rounded to float).
// The reference scatters every seed into a 256^3 float volume with float
// atomics (4.3 M atomics per seed pixel, order-dependent sums). Here the
// density is gathered per pixel instead: the seed colours are captured at
// selection time into a device list, and each pixel sums its kernel values in
// f64 in list order, so the result is deterministic and the same value the
// volume holds up to float rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "gcfg.h"

namespace gcfg {

constexpr int KDE_MAX_SEEDS = 65536;  // seed colours kept for the kernel density
constexpr int KDE_TAPS = 101;         // kernel support: floor(distance) 0..100 (DataTermKernels.cu:11)

struct DtStats {
    int32_t hist[2][3][16];
    int32_t counts[2];
    unsigned long long beta_sum;
};

class DtContext {
   public:
    ~DtContext();
    int init(int W, int H, hipStream_t stream, bool own_outputs);
    // Inputs (device pointers; the image may be owned or borrowed).
    int upload_image(const uint8_t *h_rgb);
    void use_image(const uint8_t *d_rgb) { img_ = d_rgb; }
    int upload_seeds(const uint8_t *h_seeds);
    int add_seed_rect(int x0, int y0, int w, int h, int cls);
    int clear_seeds();
    // Full compute; outputs go to the owned planes or to the given ones.
    int compute(int32_t *pos, int32_t *neg, int32_t *up, int32_t *down, int32_t *left, int32_t *right);
    int compute_owned() { return compute(pos_, neg_, up_, down_, left_, right_); }
    // Kernel-density mode: selectOnMap's 5 x 5 square at (x..x+4, y..y+4),
    // colours read from the current image now (pixels outside it are skipped).
    int kde_select(int x, int y, int size);
    // Video frames: keep the colour model of the current seeds (histograms;
    // the kernel-density seed colours are kept anyway) for later images, as the
    // reference keeps its density across frames (ComplexApplication.cu:202-207).
    void freeze_model() { frozen = computed; frozen_stats = stats; }
    // Compute into the second set of output planes (allocated on first use)
    // and swap it with the first: the previous outputs stay readable through
    // prev_*() until the next swap.
    int compute_swap();
    const int32_t *prev_pos() const { return alt_[0]; }
    int set_model(int m);
    int set_sigma(int s);
    int unquantized(double *h_tpos, double *h_tneg, double *h_right, double *h_down);

    int W = 0, H = 0;
    int32_t lambda = 50;
    int32_t model = GCFG_DT_MODEL_HISTOGRAM;
    int32_t sigma = 5;     // DEFAULT_SIGMA, DataTerm.cpp:6
    int32_t kde_count = 0; // seed colours in the list
    bool frozen = false;
    DtStats frozen_stats{};
    float gauss[KDE_TAPS] = {};
    DtStats stats{};
    int64_t pairs = 0;
    double beta = 0.0;
    bool computed = false;
    int32_t *pos_ = nullptr, *neg_ = nullptr, *up_ = nullptr, *down_ = nullptr, *left_ = nullptr,
            *right_ = nullptr;
    size_t device_bytes() const { return bytes_; }

   private:
    int build_tables();
    int build_gauss();
    hipStream_t stream_ = nullptr;
    const uint8_t *img_ = nullptr;
    uint8_t *own_img_ = nullptr, *seeds_ = nullptr;
    DtStats *d_stats_ = nullptr, *h_stats_ = nullptr;
    int32_t *d_lut_ = nullptr, *h_lut_ = nullptr;  // [2*4096 t-link LUT][thresholds]
    uint32_t *d_kde_ = nullptr;                    // seed colours 0x00RRGGBB-packed (c0 << 16 | c1 << 8 | c2)
    int32_t *alt_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // second output set
    float *d_gauss_ = nullptr;
    std::vector<void *> allocs_;
    size_t bytes_ = 0;
};

}  // namespace gcfg
