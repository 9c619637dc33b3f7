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
// Device work, three launches per image or per batch of images and no host
// round trip:
//   * stats: one workgroup per 4096-pixel chunk; seed histograms (LDS, then
//     added to the image's accumulator only by chunks that hold seeds) and the
//     chunk's exact sum of squared colour differences over right and down
//     pairs (one partial per chunk, no same-address atomics);
//   * table: eight workgroups per image; each reduces the partials, then
//     evaluates its share of the t-link costs of the 2 x 16^3 colour bins and
//     of the lambda n-link thresholds in f64, the same expressions the oracle
//     evaluates per pixel;
//   * links: one workgroup per chunk, four pixels per lane (aligned dword
//     loads of the interleaved RGB, the rows above and below through a
//     shifted window), each horizontal pair's weight computed once (vertical
//     pairs once from each side), n-link weights from an f32 estimate
//     confirmed against the thresholds, six int32 planes written with 16-byte
//     stores.
// Chunks go to the XCDs in contiguous eighths, so the rows a chunk reads from
// its neighbours are in its L2.
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
constexpr int DT_MAX_LAMBDA = 4096;   // n-link thresholds live in LDS: (lambda + 1) * 4 bytes
constexpr int DT_CHUNK = 4096;        // pixels per workgroup of the stats and links passes
constexpr int DT_LUT = 2 * 4096;      // t-link costs: foreground class [0, 4096), background [4096, 8192)
constexpr int DT_HIST = 2 * 3 * 16 + 2;  // histogram counters and the two seed counts

struct DtStats {
    int32_t hist[2][3][16];
    int32_t counts[2];
    unsigned long long beta_sum;
};

// One image's data term: inputs, outputs and device work areas.
struct DtJob {
    const uint8_t *img;      // W*H*3 interleaved RGB, 4-byte aligned
    const uint8_t *seeds;    // W*H seed classes (histogram model); null: no seeds
    int32_t *out[6];         // pos, neg, up, down, left, right (W*H each)
    int32_t W, H;
    int32_t chunk0;          // first chunk of this image in a batched launch
    int32_t lambda;
    int32_t model;           // GCFG_DT_MODEL_*
    int32_t kde_n;           // seed colours in kde
    int32_t aligned;         // 1: every output plane is 16-byte aligned (vector stores)
    int32_t pad_;
    DtStats *freeze_to;      // dt_freeze copies the statistics here (a session's kept colour model), or null
    const uint32_t *kde;     // kernel-density seed colours (0x00RRGGBB order c0, c1, c2)
    const float *gauss;      // kernel table (KDE_TAPS)
    unsigned long long *beta_part;  // one partial sum per chunk (stats pass)
    int32_t *hist_acc;       // DT_HIST counters, zero between computes (the table pass clears them)
    DtStats *stats;          // reduced statistics (table pass)
    const DtStats *frozen;   // colour model kept from an earlier image, or null
    uint16_t *tlut;          // DT_LUT t-link costs
    int32_t *thr;            // lambda + 1 n-link thresholds
};

// Kernel arguments: one job by value (njobs == 0), or a device table of
// njobs jobs and the job of every chunk.
struct DtLaunch {
    DtJob one;
    const DtJob *jobs;
    const int32_t *cmap;
    int32_t njobs;
    int32_t lds_lambda;  // largest lambda of the launch (dynamic LDS of the links pass)
    int32_t units;       // work units of the launching pass (set by dt_launch)
    // Workgroups per chunk in the stats and links passes (1, 2 or 4, each a
    // share of the chunk's rounds; single images only, whose beta partials
    // hold DT_SPLIT_MAX slots per chunk) and table workgroups per image
    // (8, 16 or 32); 0: 1 and 8.
    int32_t split;
    int32_t tparts;
};
constexpr int DT_SPLIT_MAX = 4;

__host__ __device__ inline int dt_chunks(int W, int H) { return (int)(((int64_t)W * H + DT_CHUNK - 1) / DT_CHUNK); }
// Stats, table and links passes over every chunk of the launch.
void dt_launch(const DtLaunch &L, int total_chunks, hipStream_t s);
// After a batched launch: each job's statistics copied to its freeze_to
// (one launch for the table of njobs device jobs).
void dt_freeze(const DtJob *d_jobs, int njobs, hipStream_t s);

class DtContext {
   public:
    ~DtContext();
    // Buffers for images of up to max(W * H, cap_px) pixels. With `block`
    // (device memory of at least block_bytes(cap) bytes, owned by the caller)
    // every buffer is carved from it and nothing is allocated, and the seed
    // map is not cleared (the caller uploads one before computing).
    int init(int W, int H, hipStream_t stream, bool own_outputs, size_t cap_px = 0, void *block = nullptr);
    static size_t block_bytes(size_t cap_px);
    // Reuse the buffers for a W x H image (W * H <= capacity()): the
    // kernel-density colours and the frozen colour model are dropped; the
    // seed map is left as it is (upload or clear one before computing).
    int reshape(int W, int H);
    size_t capacity() const { return cap_px_; }
    // Image and seed map copied from the host on another stream (the copies
    // are ordered before this context's later work by the caller).
    int upload_on(hipStream_t s, const uint8_t *h_rgb, const uint8_t *h_seeds);
    // Inputs (device pointers; the image may be owned or borrowed).
    int upload_image(const uint8_t *h_rgb);
    int use_image(const uint8_t *d_rgb);
    int upload_seeds(const uint8_t *h_seeds);
    int add_seed_rect(int x0, int y0, int w, int h, int cls);
    int clear_seeds();
    // Full compute, queued on the stream (no host synchronisation); outputs
    // go to the owned planes or to the given ones.
    int compute(int32_t *pos, int32_t *neg, int32_t *up, int32_t *down, int32_t *left, int32_t *right);
    int compute_owned() { return compute(pos_, neg_, up_, down_, left_, right_); }
    // Kernel-density mode: selectOnMap's 5 x 5 square at (x..x+4, y..y+4),
    // colours read from the current image now (pixels outside it are skipped).
    int kde_select(int x, int y, int size);
    // Video frames: keep the colour model of the last compute (histograms;
    // the kernel-density seed colours are kept anyway) for later images, as the
    // reference keeps its density across frames (ComplexApplication.cu:202-207).
    int freeze_model();
    // A compute into the owned planes as one job of a batched launch over
    // several contexts (histogram model): the job, whose statistics dt_freeze
    // keeps as the colour model (freeze_to), and, once the caller has queued
    // both on this context's stream, batch_queued() to record it.
    DtJob batch_job() const;
    void batch_queued() {
        computed = true;
        frozen = true;
    }
    // Compute into the second set of output planes (allocated on first use)
    // and swap it with the first: the previous outputs stay readable through
    // prev_*() until the next swap.
    int compute_swap();
    const int32_t *prev_pos() const { return alt_[0]; }
    int set_model(int m);
    int set_sigma(int s);
    // Statistics of the last compute to the host (waits for it).
    int fetch_stats();
    int unquantized(double *h_tpos, double *h_tneg, double *h_right, double *h_down);

    int W = 0, H = 0;
    int32_t lambda = 50;
    int32_t model = GCFG_DT_MODEL_HISTOGRAM;
    int32_t sigma = 5;     // DEFAULT_SIGMA, DataTerm.cpp:6
    int32_t kde_count = 0; // seed colours in the list
    bool frozen = false;
    float gauss[KDE_TAPS] = {};
    DtStats stats{};       // host copy (fetch_stats)
    int64_t pairs = 0;
    double beta = 0.0;
    bool computed = false;
    int32_t *pos_ = nullptr, *neg_ = nullptr, *up_ = nullptr, *down_ = nullptr, *left_ = nullptr,
            *right_ = nullptr;
    size_t device_bytes() const { return bytes_; }

   private:
    int build_gauss();
    DtJob job(int32_t *const *outs) const;
    hipStream_t stream_ = nullptr;
    const uint8_t *img_ = nullptr;
    uint8_t *own_img_ = nullptr, *seeds_ = nullptr;
    // stats, frozen model, table areas and per-chunk partials
    DtStats *d_stats_ = nullptr, *d_frozen_ = nullptr, *h_stats_ = nullptr;
    unsigned long long *d_beta_part_ = nullptr;
    int32_t *d_hist_acc_ = nullptr;
    uint16_t *d_tlut_ = nullptr;
    int32_t *d_thr_ = nullptr;
    uint32_t *d_kde_ = nullptr;  // seed colours (c0 << 16 | c1 << 8 | c2)
    int32_t *alt_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // second output set
    float *d_gauss_ = nullptr;
    int gauss_sigma_ = 0;  // sigma the device kernel table was built for (0: none)
    size_t cap_px_ = 0;
    std::vector<void *> allocs_;
    size_t bytes_ = 0;
    bool carved_ = false;
};

}  // namespace gcfg
