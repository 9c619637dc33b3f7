// Colour data-term kernels and the DataTerm.h C ABI (DataTerm.cpp:12-165,
// DataTermKernels.cu:19-150 re-specified as histogram / -log-likelihood /
// contrast-weighted n-links; formulas in dataterm.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "dataterm.hpp"
#include "engine.hpp"

namespace gcfg {

constexpr int DT_NT = 256;
constexpr int MAX_LAMBDA = 4096;  // thresholds live in LDS: (MAX_LAMBDA + 1) * 4 bytes
constexpr int MAX_SQDIFF = 3 * 255 * 255;

__global__ void k_dt_zero(DtStats *s) {
    int *p = reinterpret_cast<int *>(s);
    for (int i = threadIdx.x; i < (int)(sizeof(DtStats) / sizeof(int)); i += blockDim.x) p[i] = 0;
}

__device__ __forceinline__ int sqdiff(const uint8_t *a, const uint8_t *b) {
    const int d0 = (int)a[0] - b[0], d1 = (int)a[1] - b[1], d2 = (int)a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// One pass over the image: seed histograms (LDS int atomics) and the exact sum
// of squared colour differences over right and down neighbour pairs.
__global__ __launch_bounds__(DT_NT) void k_dt_stats(const uint8_t *img, const uint8_t *seeds, int W, int H,
                                                    DtStats *out) {
    __shared__ int hist[2 * 3 * 16 + 2];
    __shared__ unsigned long long red[DT_NT / 64];
    for (int i = threadIdx.x; i < 2 * 3 * 16 + 2; i += DT_NT) hist[i] = 0;
    __syncthreads();
    const int64_t n = (int64_t)W * H;
    unsigned long long bs = 0;
    for (int64_t i = blockIdx.x * (int64_t)DT_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * DT_NT) {
        const int x = (int)(i % W), y = (int)(i / W);
        const uint8_t *a = img + 3 * i;
        if (x + 1 < W) bs += sqdiff(a, a + 3);
        if (y + 1 < H) bs += sqdiff(a, a + 3 * (int64_t)W);
        const int s = seeds[i];
        if (s == GCFG_SEED_FG || s == GCFG_SEED_BG) {
            const int k = s - 1;
            atomicAdd(&hist[k * 48 + 0 * 16 + (a[0] >> 4)], 1);
            atomicAdd(&hist[k * 48 + 1 * 16 + (a[1] >> 4)], 1);
            atomicAdd(&hist[k * 48 + 2 * 16 + (a[2] >> 4)], 1);
            atomicAdd(&hist[96 + k], 1);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)bs, o, 64), hi = __shfl_xor((unsigned)(bs >> 32), o, 64);
        bs += ((unsigned long long)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bs;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < DT_NT / 64; ++w) t += red[w];
        if (t) atomicAdd(&out->beta_sum, t);
    }
    int *oh = &out->hist[0][0][0];
    for (int i = threadIdx.x; i < 96; i += DT_NT)
        if (hist[i]) atomicAdd(&oh[i], hist[i]);
    if (threadIdx.x < 2 && hist[96 + threadIdx.x]) atomicAdd(&out->counts[threadIdx.x], hist[96 + threadIdx.x]);
}

// n-link weight from the integer thresholds: w(s) = #{k in 1..lambda : s <= thr[k]}
// (thr is non-increasing in k, so this is a binary search).
__device__ __forceinline__ int nlink_weight(int s, const int32_t *thr, int lambda) {
    int lo = 0, hi = lambda;  // largest k with thr[k] >= s (thr[0] = +inf)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (thr[mid] >= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(DT_NT) void k_dt_links(const uint8_t *img, const int32_t *lut, int lambda, int W,
                                                    int H, int32_t *pos, int32_t *neg, int32_t *up, int32_t *down,
                                                    int32_t *left, int32_t *right) {
    extern __shared__ int32_t thr[];
    const int32_t *gthr = lut + 2 * 4096;
    for (int i = threadIdx.x; i <= lambda; i += DT_NT) thr[i] = gthr[i];
    __syncthreads();
    const int64_t n = (int64_t)W * H;
    for (int64_t i = blockIdx.x * (int64_t)DT_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * DT_NT) {
        const int x = (int)(i % W), y = (int)(i / W);
        const uint8_t *a = img + 3 * i;
        const int idx = ((a[0] >> 4) << 8) | ((a[1] >> 4) << 4) | (a[2] >> 4);
        pos[i] = lut[4096 + idx];  // cost of the background label
        neg[i] = lut[idx];         // cost of the foreground label
        right[i] = x + 1 < W ? nlink_weight(sqdiff(a, a + 3), thr, lambda) : 0;
        left[i] = x > 0 ? nlink_weight(sqdiff(a, a - 3), thr, lambda) : 0;
        down[i] = y + 1 < H ? nlink_weight(sqdiff(a, a + 3 * (int64_t)W), thr, lambda) : 0;
        up[i] = y > 0 ? nlink_weight(sqdiff(a, a - 3 * (int64_t)W), thr, lambda) : 0;
    }
}

// ---------------------------------------------------------------------------
// Kernel-density colour model (the reference's, DataTermKernels.cu:19-150).

// floor(sqrt(s)) for 0 <= s < 2^24. The hardware square root is within an ulp,
// which can land just below an integer root; the two integer checks make the
// result exact (the reference truncates a correctly rounded sqrtf of the same
// integer, DataTermKernels.cu:140,147, which is also exact for these s).
__device__ __forceinline__ int isqrt24(int s) {
    int k = (int)__builtin_sqrtf((float)s);
    if ((k + 1) * (k + 1) <= s) ++k;
    if (k * k > s) --k;
    return k;
}

// p(I) = sum over the seed list of g[floor(|I - s_j|)] for |I - s_j| < 101,
// in f64 and in list order (the list index is uniform across the wavefront,
// so the seed colours are scalar loads).
__device__ __forceinline__ double kde_density(const uint8_t *a, const uint32_t *__restrict__ cols, int n,
                                              const float *g) {
    const int c0 = a[0], c1 = a[1], c2 = a[2];
    double p = 0.0;
    for (int j = 0; j < n; ++j) {
        const uint32_t c = cols[j];
        const int d0 = c0 - (int)(c >> 16), d1 = c1 - (int)((c >> 8) & 255u), d2 = c2 - (int)(c & 255u);
        const int s = d0 * d0 + d1 * d1 + d2 * d2;
        if (s < KDE_TAPS * KDE_TAPS) p += (double)g[isqrt24(s)];
    }
    return p;
}

// (int) of a double as C truncates it, saturated to the int32 range (the
// reference's cast is undefined beyond it).
__device__ __forceinline__ int32_t trunc_sat(double v) {
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int32_t)0x80000000u;
    return (int32_t)v;
}

// Pre-quantisation values in f64 on the device (verification output).
// Kernel-density model (kde_cols != null): tpos = 300000 p, tneg = 300000 (1 - p).
__global__ __launch_bounds__(DT_NT) void k_dt_unquantized(const uint8_t *img, const DtStats *st, double beta,
                                                          int lambda, int W, int H, const uint32_t *kde_cols,
                                                          int kde_n, const float *gauss, double *tpos,
                                                          double *tneg, double *rightf, double *downf) {
    const int64_t n = (int64_t)W * H;
    for (int64_t i = blockIdx.x * (int64_t)DT_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * DT_NT) {
        const int x = (int)(i % W), y = (int)(i / W);
        const uint8_t *a = img + 3 * i;
        if (kde_cols) {
            const double p = kde_density(a, kde_cols, kde_n, gauss);
            tpos[i] = 300000.0 * p;
            tneg[i] = 300000.0 * (1.0 - p);
        } else {
            double cost[2];
            for (int k = 0; k < 2; ++k) {
                double p = 1.0;
                for (int c = 0; c < 3; ++c)
                    p *= (double)(st->hist[k][c][a[c] >> 4] + 1) / (double)(st->counts[k] + 16);
                cost[k] = 64.0 * -log(p);
            }
            tpos[i] = cost[1];
            tneg[i] = cost[0];
        }
        double wr = 0.0, wd = 0.0;
        if (x + 1 < W) wr = beta > 0.0 ? (double)lambda * exp(-(double)sqdiff(a, a + 3) / (2.0 * beta)) : lambda;
        if (y + 1 < H)
            wd = beta > 0.0 ? (double)lambda * exp(-(double)sqdiff(a, a + 3 * (int64_t)W) / (2.0 * beta)) : lambda;
        rightf[i] = wr;
        downf[i] = wd;
    }
}

// Seed colours of the in-image part of a selection square, in the
// reference's launch order (x offset outer, y offset inner, DataTerm.cpp:114-123).
__global__ void k_dt_kde_gather(const uint8_t *img, int W, int x0, int y0, int ni, int nj, uint32_t *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj) return;
    const int X = x0 + t / nj, Y = y0 + t % nj;
    const uint8_t *a = img + 3 * ((int64_t)Y * W + X);
    out[t] = ((uint32_t)a[0] << 16) | ((uint32_t)a[1] << 8) | (uint32_t)a[2];
}

// t-links from the kernel density (K_CalcTerms, DataTermKernels.cu:61-83,
// with p rounded to float as the reference's volume holds it), n-links as in
// k_dt_links.
__global__ __launch_bounds__(DT_NT) void k_dt_kde_links(const uint8_t *img, const uint32_t *__restrict__ cols,
                                                        int ncols, const float *gauss, const int32_t *lut,
                                                        int lambda, int W, int H, int32_t *pos, int32_t *neg,
                                                        int32_t *up, int32_t *down, int32_t *left,
                                                        int32_t *right) {
    extern __shared__ int32_t thr[];
    __shared__ float g[KDE_TAPS];
    const int32_t *gthr = lut + 2 * 4096;
    for (int i = threadIdx.x; i <= lambda; i += DT_NT) thr[i] = gthr[i];
    for (int i = threadIdx.x; i < KDE_TAPS; i += DT_NT) g[i] = gauss[i];
    __syncthreads();
    const int64_t n = (int64_t)W * H;
    for (int64_t i = blockIdx.x * (int64_t)DT_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * DT_NT) {
        const int x = (int)(i % W), y = (int)(i / W);
        const uint8_t *a = img + 3 * i;
        const float pf = (float)kde_density(a, cols, ncols, g);
        pos[i] = trunc_sat(300000.0 * (double)pf);
        neg[i] = trunc_sat(300000.0 * (double)(1.0f - pf));
        right[i] = x + 1 < W ? nlink_weight(sqdiff(a, a + 3), thr, lambda) : 0;
        left[i] = x > 0 ? nlink_weight(sqdiff(a, a - 3), thr, lambda) : 0;
        down[i] = y + 1 < H ? nlink_weight(sqdiff(a, a + 3 * (int64_t)W), thr, lambda) : 0;
        up[i] = y > 0 ? nlink_weight(sqdiff(a, a - 3 * (int64_t)W), thr, lambda) : 0;
    }
}

__global__ void k_fill_rect_u8(uint8_t *map, int W, int H, int x0, int y0, int w, int h, uint8_t v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < w * h; i += gridDim.x * blockDim.x) {
        const int x = x0 + i % w, y = y0 + i / w;
        if (x >= 0 && y >= 0 && x < W && y < H) map[(int64_t)y * W + x] = v;
    }
}

__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int grid_for(int64_t n) {
    const int64_t g = (n + DT_NT - 1) / DT_NT;
    return (int)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

// ---------------------------------------------------------------------------
DtContext::~DtContext() {
    // Teardown: errors here have nowhere to go.
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void *p : allocs_) (void)hipFree(p);
    if (h_stats_) (void)hipHostFree(h_stats_);
    if (h_lut_) (void)hipHostFree(h_lut_);
}

int DtContext::init(int w, int h, hipStream_t stream, bool own_outputs) {
    if (w <= 0 || h <= 0) {
        set_error("data term: shape must be positive");
        return GCFG_ERR_INVALID;
    }
    W = w;
    H = h;
    stream_ = stream;
    const size_t n = (size_t)W * H;
    auto dalloc = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            set_error("data term: device allocation failed");
            return GCFG_ERR_NOMEM;
        }
        allocs_.push_back(*p);
        bytes_ += bytes;
        return GCFG_OK;
    };
    int rc = GCFG_OK;
    rc |= dalloc((void **)&own_img_, 3 * n);
    rc |= dalloc((void **)&seeds_, n);
    rc |= dalloc((void **)&d_stats_, sizeof(DtStats));
    rc |= dalloc((void **)&d_lut_, (2 * 4096 + MAX_LAMBDA + 1) * sizeof(int32_t));
    rc |= dalloc((void **)&d_kde_, KDE_MAX_SEEDS * sizeof(uint32_t));
    rc |= dalloc((void **)&d_gauss_, KDE_TAPS * sizeof(float));
    if (own_outputs) {
        rc |= dalloc((void **)&pos_, n * 4);
        rc |= dalloc((void **)&neg_, n * 4);
        rc |= dalloc((void **)&up_, n * 4);
        rc |= dalloc((void **)&down_, n * 4);
        rc |= dalloc((void **)&left_, n * 4);
        rc |= dalloc((void **)&right_, n * 4);
    }
    if (rc != GCFG_OK) return GCFG_ERR_NOMEM;
    GCFG_HIP(hipHostMalloc((void **)&h_stats_, sizeof(DtStats), hipHostMallocDefault));
    GCFG_HIP(hipHostMalloc((void **)&h_lut_, (2 * 4096 + MAX_LAMBDA + 1) * sizeof(int32_t), hipHostMallocDefault));
    img_ = own_img_;
    return clear_seeds();
}

int DtContext::upload_image(const uint8_t *h_rgb) {
    if (!h_rgb) {
        set_error("data term: null image");
        return GCFG_ERR_INVALID;
    }
    GCFG_HIP(hipMemcpyAsync(own_img_, h_rgb, (size_t)3 * W * H, hipMemcpyHostToDevice, stream_));
    img_ = own_img_;
    computed = false;
    return GCFG_OK;
}

int DtContext::upload_seeds(const uint8_t *h_seeds) {
    if (!h_seeds) {
        set_error("data term: null seed map");
        return GCFG_ERR_INVALID;
    }
    GCFG_HIP(hipMemcpyAsync(seeds_, h_seeds, (size_t)W * H, hipMemcpyHostToDevice, stream_));
    computed = false;
    return GCFG_OK;
}

int DtContext::add_seed_rect(int x0, int y0, int w, int h, int cls) {
    if (w <= 0 || h <= 0 || cls < 0 || cls > 2) {
        set_error("data term: invalid seed rectangle");
        return GCFG_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_fill_rect_u8, dim3(grid_for((int64_t)w * h)), dim3(DT_NT), 0, stream_, seeds_, W, H, x0,
                       y0, w, h, (uint8_t)cls);
    GCFG_HIP(hipGetLastError());
    computed = false;
    return GCFG_OK;
}

int DtContext::clear_seeds() {
    kde_count = 0;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for((int64_t)W * H)), dim3(DT_NT), 0, stream_, seeds_, (int64_t)W * H,
                       (uint8_t)0);
    GCFG_HIP(hipGetLastError());
    computed = false;
    return GCFG_OK;
}

int DtContext::set_model(int m) {
    if (m != GCFG_DT_MODEL_HISTOGRAM && m != GCFG_DT_MODEL_KDE) {
        set_error("data term: unknown colour model");
        return GCFG_ERR_INVALID;
    }
    model = m;
    computed = false;
    return GCFG_OK;
}

int DtContext::set_sigma(int s) {
    if (s < 1 || s > 1000) {
        set_error("data term: sigma must be in 1..1000");
        return GCFG_ERR_INVALID;
    }
    sigma = s;
    computed = false;
    return GCFG_OK;
}

int DtContext::kde_select(int x, int y, int size) {
    const int xa = std::max(x, 0), xb = std::min(x + size, W);
    const int ya = std::max(y, 0), yb = std::min(y + size, H);
    const int ni = xb > xa ? xb - xa : 0, nj = yb > ya ? yb - ya : 0;
    if (ni * nj == 0) return GCFG_OK;
    if (kde_count + ni * nj > KDE_MAX_SEEDS) {
        set_error("data term: kernel-density seed list is full (65536 colours)");
        return GCFG_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_dt_kde_gather, dim3(1), dim3(DT_NT), 0, stream_, img_, W, xa, ya, ni, nj,
                       d_kde_ + kde_count);
    GCFG_HIP(hipGetLastError());
    kde_count += ni * nj;
    computed = false;
    return GCFG_OK;
}

// Kernel values of K_InitGaussian (DataTermKernels.cu:19-31) with its mixed
// float/double arithmetic: x / sigma in float, the exponent in double rounded
// to float for expf, the factor 1 / (sigma * 2.507) in double, stored as float.
int DtContext::build_gauss() {
    for (int k = 0; k < KDE_TAPS; ++k) {
        const float t = (float)k / (float)sigma;
        const double arg = -0.5 * (double)t * (double)t;
        gauss[k] = (float)(1.0 / ((double)(float)sigma * 2.507) * (double)expf((float)arg));
    }
    GCFG_HIP(hipMemcpyAsync(d_gauss_, gauss, sizeof(gauss), hipMemcpyHostToDevice, stream_));
    return GCFG_OK;
}

// t-link LUT over the 16^3 colour bins (entries [0,4096): foreground cost,
// [4096, 8192): background cost) and the n-link thresholds thr[k] = largest
// squared difference s with llround(lambda * exp(-s / (2 beta))) >= k.
int DtContext::build_tables() {
    for (int k = 0; k < 2; ++k) {
        double r[3][16];
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < 16; ++b)
                r[c][b] = (double)(stats.hist[k][c][b] + 1) / (double)(stats.counts[k] + 16);
        for (int idx = 0; idx < 4096; ++idx) {
            double p = 1.0;
            p *= r[0][(idx >> 8) & 15];
            p *= r[1][(idx >> 4) & 15];
            p *= r[2][idx & 15];
            h_lut_[k * 4096 + idx] = (int32_t)llround(64.0 * -log(p));
        }
    }
    int32_t *thr = h_lut_ + 2 * 4096;
    thr[0] = 0x7fffffff;
    auto weight = [&](int s) -> long long {
        return beta > 0.0 ? llround((double)lambda * exp(-(double)s / (2.0 * beta))) : (long long)lambda;
    };
    for (int k = 1; k <= lambda; ++k) {
        if (weight(MAX_SQDIFF) >= k) {
            thr[k] = MAX_SQDIFF;
            continue;
        }
        if (weight(0) < k) {
            thr[k] = -1;
            continue;
        }
        int lo = 0, hi = MAX_SQDIFF;  // weight(lo) >= k > weight(hi)
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            if (weight(mid) >= k) lo = mid; else hi = mid;
        }
        thr[k] = lo;
    }
    GCFG_HIP(hipMemcpyAsync(d_lut_, h_lut_, (2 * 4096 + lambda + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                            stream_));
    return GCFG_OK;
}

int DtContext::compute(int32_t *pos, int32_t *neg, int32_t *up, int32_t *down, int32_t *left, int32_t *right) {
    if (lambda < 0 || lambda > MAX_LAMBDA) {
        set_error("data term: lambda out of range");
        return GCFG_ERR_INVALID;
    }
    const int64_t n = (int64_t)W * H;
    hipLaunchKernelGGL(k_dt_zero, dim3(1), dim3(128), 0, stream_, d_stats_);
    hipLaunchKernelGGL(k_dt_stats, dim3(grid_for(n)), dim3(DT_NT), 0, stream_, img_, seeds_, W, H, d_stats_);
    GCFG_HIP(hipGetLastError());
    GCFG_HIP(hipMemcpyAsync(h_stats_, d_stats_, sizeof(DtStats), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    stats = *h_stats_;
    if (frozen) {  // keep the model's histograms; beta follows the current image
        std::memcpy(stats.hist, frozen_stats.hist, sizeof(stats.hist));
        std::memcpy(stats.counts, frozen_stats.counts, sizeof(stats.counts));
        std::memcpy(h_stats_, &stats, sizeof(DtStats));
        GCFG_HIP(hipMemcpyAsync(d_stats_, h_stats_, sizeof(DtStats), hipMemcpyHostToDevice, stream_));
    }
    pairs = (int64_t)(W - 1) * H + (int64_t)W * (H - 1);
    beta = pairs > 0 ? (double)stats.beta_sum / (double)pairs : 0.0;
    int rc = build_tables();
    if (rc != GCFG_OK) return rc;
    if (model == GCFG_DT_MODEL_KDE) {
        rc = build_gauss();
        if (rc != GCFG_OK) return rc;
        hipLaunchKernelGGL(k_dt_kde_links, dim3(grid_for(n)), dim3(DT_NT), (lambda + 1) * sizeof(int32_t), stream_,
                           img_, d_kde_, kde_count, d_gauss_, d_lut_, lambda, W, H, pos, neg, up, down, left, right);
    } else {
        hipLaunchKernelGGL(k_dt_links, dim3(grid_for(n)), dim3(DT_NT), (lambda + 1) * sizeof(int32_t), stream_,
                           img_, d_lut_, lambda, W, H, pos, neg, up, down, left, right);
    }
    GCFG_HIP(hipGetLastError());
    computed = true;
    return GCFG_OK;
}

int DtContext::compute_swap() {
    if (!pos_) {
        set_error("data term: no owned output planes");
        return GCFG_ERR_STATE;
    }
    const size_t n = (size_t)W * H;
    for (int k = 0; k < 6; ++k)
        if (!alt_[k]) {
            if (hipMalloc((void **)&alt_[k], n * 4) != hipSuccess) {
                alt_[k] = nullptr;
                set_error("data term: device allocation failed");
                return GCFG_ERR_NOMEM;
            }
            allocs_.push_back(alt_[k]);
            bytes_ += n * 4;
        }
    int rc = compute(alt_[0], alt_[1], alt_[2], alt_[3], alt_[4], alt_[5]);
    if (rc != GCFG_OK) return rc;
    int32_t **cur[6] = {&pos_, &neg_, &up_, &down_, &left_, &right_};
    for (int k = 0; k < 6; ++k) std::swap(*cur[k], alt_[k]);
    return GCFG_OK;
}

int DtContext::unquantized(double *h_tpos, double *h_tneg, double *h_right, double *h_down) {
    if (!computed) {
        set_error("data term: compute before reading unquantised values");
        return GCFG_ERR_STATE;
    }
    const size_t n = (size_t)W * H;
    double *buf = nullptr;
    GCFG_HIP(hipMalloc((void **)&buf, 4 * n * sizeof(double)));
    const bool kde = model == GCFG_DT_MODEL_KDE;
    hipLaunchKernelGGL(k_dt_unquantized, dim3(grid_for((int64_t)n)), dim3(DT_NT), 0, stream_, img_, d_stats_, beta,
                       lambda, W, H, kde ? d_kde_ : nullptr, kde_count, d_gauss_, buf, buf + n, buf + 2 * n,
                       buf + 3 * n);
    hipError_t e = hipGetLastError();
    double *outs[4] = {h_tpos, h_tneg, h_right, h_down};
    for (int k = 0; k < 4 && e == hipSuccess; ++k)
        if (outs[k]) e = hipMemcpyAsync(outs[k], buf + k * n, n * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    (void)hipFree(buf);
    GCFG_HIP(e);
    return GCFG_OK;
}

}  // namespace gcfg

// ---------------------------------------------------------------------------
// C ABI for DataTerm.h.
struct gcfg_dataterm {
    hipStream_t stream = nullptr;
    gcfg::DtContext ctx;
    gcfg::StreamJoin join;
};

extern "C" {

int gcfg_dt_create(int32_t width, int32_t height, gcfg_dataterm **out) {
    if (!out) return GCFG_ERR_INVALID;
    *out = nullptr;
    gcfg_dataterm *dt = new (std::nothrow) gcfg_dataterm();
    if (!dt) return GCFG_ERR_NOMEM;
    if (hipStreamCreateWithFlags(&dt->stream, hipStreamNonBlocking) != hipSuccess) {
        gcfg::set_error("gcfg_dt_create: stream creation failed");
        delete dt;
        return GCFG_ERR_HIP;
    }
    int rc = dt->ctx.init(width, height, dt->stream, true);
    if (rc != GCFG_OK) {
        (void)hipStreamDestroy(dt->stream);
        delete dt;
        return rc;
    }
    *out = dt;
    return GCFG_OK;
}

int gcfg_dt_destroy(gcfg_dataterm *dt) {
    if (!dt) return GCFG_OK;
    hipStream_t s = dt->stream;
    (void)hipStreamSynchronize(s);
    delete dt;
    (void)hipStreamDestroy(s);
    return GCFG_OK;
}

int gcfg_dt_set_image(gcfg_dataterm *dt, const uint8_t *h_rgb) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.upload_image(h_rgb);
}

int gcfg_dt_set_image_device(gcfg_dataterm *dt, const uint8_t *d_rgb) {
    if (!dt || !d_rgb) return GCFG_ERR_INVALID;
    dt->ctx.use_image(d_rgb);
    dt->ctx.computed = false;
    return GCFG_OK;
}

int gcfg_dt_set_seeds(gcfg_dataterm *dt, const uint8_t *h_seeds) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.upload_seeds(h_seeds);
}

int gcfg_dt_add_seed_rect(gcfg_dataterm *dt, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t cls) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.add_seed_rect(x0, y0, w, h, cls);
}

int gcfg_dt_reset(gcfg_dataterm *dt) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.clear_seeds();
}

int gcfg_dt_set_lambda(gcfg_dataterm *dt, int32_t lambda) {
    if (!dt || lambda < 0 || lambda > gcfg::MAX_LAMBDA) {
        gcfg::set_error("gcfg_dt_set_lambda: lambda out of range");
        return GCFG_ERR_INVALID;
    }
    dt->ctx.lambda = lambda;
    dt->ctx.computed = false;
    return GCFG_OK;
}

int gcfg_dt_compute(gcfg_dataterm *dt) {
    if (!dt) return GCFG_ERR_INVALID;
    int rc = dt->join.before(dt->stream);
    if (rc == GCFG_OK) rc = dt->ctx.compute_owned();
    if (rc == GCFG_OK) rc = dt->join.after(dt->stream);
    if (rc == GCFG_OK && hipStreamSynchronize(dt->stream) != hipSuccess) return GCFG_ERR_HIP;
    return rc;
}

int gcfg_dt_set_stream(gcfg_dataterm *dt, void *hip_stream) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->join.set((hipStream_t)hip_stream);
}

int gcfg_dt_get_links(gcfg_dataterm *dt, int32_t **d_pos, int32_t **d_neg, int32_t **d_up, int32_t **d_down,
                      int32_t **d_left, int32_t **d_right) {
    if (!dt) return GCFG_ERR_INVALID;
    if (!dt->ctx.computed) {
        gcfg::set_error("gcfg_dt_get_links: call gcfg_dt_compute first");
        return GCFG_ERR_STATE;
    }
    if (d_pos) *d_pos = dt->ctx.pos_;
    if (d_neg) *d_neg = dt->ctx.neg_;
    if (d_up) *d_up = dt->ctx.up_;
    if (d_down) *d_down = dt->ctx.down_;
    if (d_left) *d_left = dt->ctx.left_;
    if (d_right) *d_right = dt->ctx.right_;
    return GCFG_OK;
}

int gcfg_dt_get_histograms(gcfg_dataterm *dt, int32_t *h_hist96, int32_t *h_counts2) {
    if (!dt) return GCFG_ERR_INVALID;
    if (!dt->ctx.computed) {
        gcfg::set_error("gcfg_dt_get_histograms: call gcfg_dt_compute first");
        return GCFG_ERR_STATE;
    }
    if (h_hist96) std::memcpy(h_hist96, &dt->ctx.stats.hist[0][0][0], 96 * sizeof(int32_t));
    if (h_counts2) std::memcpy(h_counts2, dt->ctx.stats.counts, 2 * sizeof(int32_t));
    return GCFG_OK;
}

int gcfg_dt_get_beta(gcfg_dataterm *dt, int64_t *sum, int64_t *pairs, double *beta) {
    if (!dt) return GCFG_ERR_INVALID;
    if (!dt->ctx.computed) {
        gcfg::set_error("gcfg_dt_get_beta: call gcfg_dt_compute first");
        return GCFG_ERR_STATE;
    }
    if (sum) *sum = (int64_t)dt->ctx.stats.beta_sum;
    if (pairs) *pairs = dt->ctx.pairs;
    if (beta) *beta = dt->ctx.beta;
    return GCFG_OK;
}

int gcfg_dt_unquantized(gcfg_dataterm *dt, double *h_tpos, double *h_tneg, double *h_right, double *h_down) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.unquantized(h_tpos, h_tneg, h_right, h_down);
}

int gcfg_dt_set_model(gcfg_dataterm *dt, int32_t model) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.set_model(model);
}

int gcfg_dt_set_sigma(gcfg_dataterm *dt, int32_t sigma) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.set_sigma(sigma);
}

int gcfg_dt_kde_select(gcfg_dataterm *dt, int32_t x, int32_t y) {
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.kde_select(x, y, 5);
}

int gcfg_dt_kde_count(gcfg_dataterm *dt, int32_t *count) {
    if (!dt || !count) return GCFG_ERR_INVALID;
    *count = dt->ctx.kde_count;
    return GCFG_OK;
}

}  // extern "C"
