// Colour data-term kernels and the DataTerm.h C ABI (DataTerm.cpp:12-165,
// DataTermKernels.cu:19-150 re-specified as histogram / -log-likelihood /
// contrast-weighted n-links; formulas and pass structure in dataterm.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "dataterm.hpp"
#include "engine.hpp"

namespace gcfg {

constexpr int DT_NT = 256;
constexpr int MAX_SQDIFF = 3 * 255 * 255;
constexpr int DT_ROUNDS = DT_CHUNK / (4 * DT_NT);  // quads of 4 pixels per lane in one chunk

struct Rgb {
    int c0, c1, c2;
};

__device__ __forceinline__ int sq(const Rgb &a, const Rgb &b) {
    const int d0 = a.c0 - b.c0, d1 = a.c1 - b.c1, d2 = a.c2 - b.c2;
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// Byte b of a little-endian word array (b is a constant after unrolling).
template <int N>
__device__ __forceinline__ int byte_at(const uint32_t (&w)[N], int b) {
    return (int)((w[b >> 2] >> (8 * (b & 3))) & 255u);
}
template <int N>
__device__ __forceinline__ Rgb px_at(const uint32_t (&w)[N], int b) {
    return Rgb{byte_at(w, b), byte_at(w, b + 1), byte_at(w, b + 2)};
}

// Bytes [o, o + 4N) of an image of `total` bytes as N words (o 4-byte
// aligned): dword loads when the range is inside the image, otherwise byte
// loads with the bytes outside it read as 0 (first and last lanes only).
template <int N>
__device__ __forceinline__ void load_words(const uint8_t *img, int64_t o, int64_t total, uint32_t (&w)[N]) {
    if (o >= 0 && o + 4 * N <= total) {
        const __attribute__((address_space(1))) uint32_t *p =
            (const __attribute__((address_space(1))) uint32_t *)(img + o);
#pragma unroll
        for (int k = 0; k < N; ++k) w[k] = p[k];
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int64_t q = o + 4 * k + b;
                if (q >= 0 && q < total) v |= (uint32_t)img[q] << (8 * b);
            }
            w[k] = v;
        }
    }
}

// 12 bytes (4 pixels) at any byte offset o: four aligned words realigned by
// o & 3, which is uniform across an image's lanes for a row offset.
__device__ __forceinline__ void load_px12(const uint8_t *img, int64_t o, int64_t total, uint32_t (&r)[3]) {
    const int64_t a = o & ~(int64_t)3;
    const unsigned s = (unsigned)(o - a);
    uint32_t w[4];
    load_words<4>(img, a, total, w);
    r[0] = __builtin_amdgcn_alignbyte(w[1], w[0], s);
    r[1] = __builtin_amdgcn_alignbyte(w[2], w[1], s);
    r[2] = __builtin_amdgcn_alignbyte(w[3], w[2], s);
}

// Work unit of workgroup b among n: units are dealt to the 8 XCDs (one L2
// each) in contiguous eighths, since workgroup b runs on XCD b % 8. A chunk
// reads the rows above and below it, which the neighbouring chunks own: kept
// on one XCD and in flight together, those reads hit its L2 instead of
// fetching every row twice (the stats pass's loads take half the time).
// Units past n (the grid is rounded up to whole eighths) return at once.
__device__ __forceinline__ int xcd_unit(int b, int n) {
    const int per = (n + NUM_XCD - 1) / NUM_XCD;
    return (b % NUM_XCD) * per + b / NUM_XCD;
}
__host__ __device__ inline int xcd_grid(int n) { return (n + NUM_XCD - 1) / NUM_XCD * NUM_XCD; }

// The job of a chunk: the single job, or the job table through the chunk map.
__device__ __forceinline__ void dt_job(const DtLaunch &L, int b, DtJob &J, int &c) {
    if (L.njobs == 0) {
        J = L.one;
        c = b;
        return;
    }
    J = L.jobs[L.cmap[b]];
    c = b - J.chunk0;
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

// The same weight from the closed form's f32 estimate k = rint(lambda *
// exp(-s / (2 beta))), confirmed against two thresholds: k is the weight
// exactly when thr[k] >= s > thr[k + 1] (thr[0] = +inf; past lambda, -inf).
// The f32 estimate misses only next to a rounding boundary of the f64
// formula; those take the binary search. Two independent LDS reads instead of
// a chain of log2(lambda) dependent ones.
__device__ __forceinline__ int nlink_weight_est(int s, const int32_t *thr, int lambda, float neg_inv_2beta) {
    int k = (int)__builtin_rintf((float)lambda * __expf((float)s * neg_inv_2beta));
    k = min(max(k, 0), lambda);
    const int a = thr[k], b = k < lambda ? thr[k + 1] : -1;
    return (a >= s && b < s) ? k : nlink_weight(s, thr, lambda);
}

// Stats pass: seed histograms and the chunk's sum of squared colour
// differences over right and down pairs (each pair once).
__global__ __launch_bounds__(DT_NT) void k_dt_stats(DtLaunch L) {
    __shared__ int hist[DT_HIST];
    __shared__ unsigned long long red[DT_NT / 64];
    const int unit = xcd_unit(blockIdx.x, L.units);
    if (unit >= L.units) return;
    DtJob J;
    int c;
    const int split = L.split;
    dt_job(L, unit / split, J, c);
    // This workgroup's share of the chunk's rounds.
    const int part = unit % split, r0 = part * (DT_ROUNDS / split), r1 = r0 + DT_ROUNDS / split;
    for (int i = threadIdx.x; i < DT_HIST; i += DT_NT) hist[i] = 0;
    __syncthreads();
    const int W = J.W, H = J.H, n = J.W * J.H;
    const int64_t total = 3 * (int64_t)n;
    unsigned long long bs = 0;
    int seen = 0;
    // Every round's loads are issued before any is used: four times the bytes
    // in flight per lane that a load-then-compute loop keeps.
    uint32_t w[DT_ROUNDS][5], b[DT_ROUNDS][3], sd[DT_ROUNDS];
#pragma unroll
    for (int r = 0; r < DT_ROUNDS; ++r) {
        const int i0 = c * DT_CHUNK + (r * DT_NT + (int)threadIdx.x) * 4;
        sd[r] = 0;
        if (r >= r0 && r < r1 && i0 < n) {
            load_words<5>(J.img, 3 * (int64_t)i0 - 4, total, w[r]);  // p(-1) .. p(4)
            load_px12(J.img, 3 * ((int64_t)i0 + W), total, b[r]);     // the four pixels below
            if (J.seeds) {
                if (i0 + 4 <= n) {
                    sd[r] = *(const __attribute__((address_space(1))) uint32_t *)(J.seeds + i0);
                } else {
                    for (int k = 0; i0 + k < n; ++k) sd[r] |= (uint32_t)J.seeds[i0 + k] << (8 * k);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < DT_ROUNDS; ++r) {
        const int i0 = c * DT_CHUNK + (r * DT_NT + (int)threadIdx.x) * 4;
        if (r < r0) continue;
        if (r >= r1 || i0 >= n) break;
        int x = i0 % W, y = i0 / W;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < n) {
                const Rgb pk = px_at(w[r], 4 + 3 * k);
                if (x + 1 < W) bs += (unsigned)sq(pk, px_at(w[r], 7 + 3 * k));
                if (y + 1 < H) bs += (unsigned)sq(pk, px_at(b[r], 3 * k));
                const int s = (int)((sd[r] >> (8 * k)) & 255u);
                if (s == GCFG_SEED_FG || s == GCFG_SEED_BG) {
                    const int cl = s - 1;
                    atomicAdd(&hist[cl * 48 + 0 * 16 + (pk.c0 >> 4)], 1);
                    atomicAdd(&hist[cl * 48 + 1 * 16 + (pk.c1 >> 4)], 1);
                    atomicAdd(&hist[cl * 48 + 2 * 16 + (pk.c2 >> 4)], 1);
                    atomicAdd(&hist[96 + cl], 1);
                    seen = 1;
                }
            }
            if (++x == W) {
                x = 0;
                ++y;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)bs, o, 64), hi = __shfl_xor((unsigned)(bs >> 32), o, 64);
        bs += ((unsigned long long)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bs;
    const int any_seed = __syncthreads_or(seen);
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < DT_NT / 64; ++k) t += red[k];
        J.beta_part[c * split + part] = t;
    }
    // Only the few chunks that hold seed pixels touch the image's counters.
    if (any_seed)
        for (int i = threadIdx.x; i < DT_HIST; i += DT_NT)
            if (hist[i]) atomicAdd(&J.hist_acc[i], hist[i]);
}

// n-link weight of a squared difference, evaluated as the oracle does.
__device__ __forceinline__ long long nlink_exact(int s, int lambda, double beta) {
    return beta > 0.0 ? llround((double)lambda * exp(-(double)s / (2.0 * beta))) : (long long)lambda;
}

// Table pass (L.tparts workgroups per image): reduce the partials,
// keep the statistics, build the t-link costs of the 16^3 colour bins of both
// classes and the n-link thresholds thr[k] = largest squared difference s
// with w(s) >= k (thr[0] = +inf, -1 where no s qualifies). Every part reduces
// the partials and histograms itself and evaluates its share of the costs and
// thresholds (f64 logarithms and exponentials: one image's table would keep a
// single workgroup busy for about 15 us); part 0 writes the statistics. The
// histogram accumulator is cleared by the links pass, after every part has
// read it.
__global__ __launch_bounds__(DT_NT) void k_dt_table(DtLaunch L) {
    __shared__ int hist[DT_HIST];
    __shared__ double rr[2][3][16];
    __shared__ unsigned long long red[DT_NT / 64];
    const int parts = L.tparts;
    const int part = blockIdx.x % parts;
    const DtJob J = L.njobs ? L.jobs[blockIdx.x / parts] : L.one;
    const int nch = dt_chunks(J.W, J.H) * L.split;  // beta partials
    unsigned long long bs = 0;
    for (int k = threadIdx.x; k < nch; k += DT_NT) bs += J.beta_part[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)bs, o, 64), hi = __shfl_xor((unsigned)(bs >> 32), o, 64);
        bs += ((unsigned long long)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bs;
    for (int i = threadIdx.x; i < DT_HIST; i += DT_NT)
        hist[i] = J.frozen ? (i < 96 ? (&J.frozen->hist[0][0][0])[i] : J.frozen->counts[i - 96]) : J.hist_acc[i];
    __syncthreads();
    unsigned long long tot = 0;
    for (int k = 0; k < DT_NT / 64; ++k) tot += red[k];
    if (part == 0) {
        for (int i = threadIdx.x; i < DT_HIST; i += DT_NT) {
            if (i < 96) (&J.stats->hist[0][0][0])[i] = hist[i];
            else J.stats->counts[i - 96] = hist[i];
        }
        if (threadIdx.x == 0) J.stats->beta_sum = tot;
    }
    if (threadIdx.x < 96) {
        const int k = threadIdx.x / 48;
        rr[k][(threadIdx.x / 16) % 3][threadIdx.x % 16] =
            (double)(hist[threadIdx.x] + 1) / (double)(hist[96 + k] + 16);
    }
    __syncthreads();
    if (J.model == GCFG_DT_MODEL_HISTOGRAM) {
        const int PER = DT_LUT / parts;
        for (int e = part * PER + threadIdx.x; e < (part + 1) * PER; e += DT_NT) {
            const int k = e >> 12, idx = e & 4095;
            double p = 1.0;
            p *= rr[k][0][(idx >> 8) & 15];
            p *= rr[k][1][(idx >> 4) & 15];
            p *= rr[k][2][idx & 15];
            J.tlut[e] = (uint16_t)llround(64.0 * -log(p));
        }
    }
    const int64_t pairs = (int64_t)(J.W - 1) * J.H + (int64_t)J.W * (J.H - 1);
    const double beta = pairs > 0 ? (double)tot / (double)pairs : 0.0;
    const int lambda = J.lambda;
    for (int k = part * DT_NT + threadIdx.x; k <= lambda; k += DT_NT * parts) {
        int t;
        if (k == 0) {
            t = 0x7fffffff;
        } else if (nlink_exact(MAX_SQDIFF, lambda, beta) >= k) {
            t = MAX_SQDIFF;
        } else if (nlink_exact(0, lambda, beta) < k) {
            t = -1;
        } else {
            // w(s) >= k <=> lambda exp(-s / 2 beta) >= k - 1/2 <=> s <= 2 beta ln(lambda / (k - 1/2)):
            // start there and step to the exact boundary (w(lo) >= k > w(lo + 1)) by evaluating w
            // itself; the binary search is kept for a start that is far off.
            const double est = 2.0 * beta * log((double)lambda / ((double)k - 0.5));
            int lo = (int)fmin(fmax(floor(est), 0.0), (double)(MAX_SQDIFF - 1));
            int steps = 0;
            while (steps < 8 && nlink_exact(lo, lambda, beta) < k) {
                --lo;
                ++steps;
            }
            while (steps < 8 && lo + 1 <= MAX_SQDIFF && nlink_exact(lo + 1, lambda, beta) >= k) {
                ++lo;
                ++steps;
            }
            if (steps >= 8 || lo < 0 || nlink_exact(lo, lambda, beta) < k ||
                (lo + 1 <= MAX_SQDIFF && nlink_exact(lo + 1, lambda, beta) >= k)) {
                lo = 0;
                int hi = MAX_SQDIFF;  // w(lo) >= k > w(hi)
                while (hi - lo > 1) {
                    const int mid = lo + (hi - lo) / 2;
                    if (nlink_exact(mid, lambda, beta) >= k) lo = mid; else hi = mid;
                }
            }
            t = lo;
        }
        J.thr[k] = t;
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
__device__ __forceinline__ double kde_density(const Rgb &a, const uint32_t *__restrict__ cols, int n,
                                              const float *g) {
    double p = 0.0;
    for (int j = 0; j < n; ++j) {
        const uint32_t c = cols[j];
        const int d0 = a.c0 - (int)(c >> 16), d1 = a.c1 - (int)((c >> 8) & 255u), d2 = a.c2 - (int)(c & 255u);
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

// Links pass: six planes for four pixels per lane. The weight of each
// horizontal pair is computed once (left of pixel k = right of pixel k - 1).
// The kernel-density mode is arithmetic-bound (seeds x pixels): its chunks are
// split over DT_KDE_SPLIT workgroups for more wavefronts in flight, whatever
// split the other passes use (a workgroup with a single round of its chunk
// runs the pass over the seed list for half a pair of rounds, twice the work
// per pixel, and each workgroup builds the kernel table).
constexpr int DT_KDE_SPLIT = 2;
template <bool KDE>
__global__ __launch_bounds__(DT_NT) void k_dt_links(DtLaunch L) {
    extern __shared__ int32_t thr[];
    __shared__ uint32_t lut2[KDE ? 1 : DT_LUT / 2];  // u16 costs, two per word
    // Kernel-density mode: the kernel value by squared distance,
    // gs[s] = g[floor(sqrt(s))] for s < 101^2 (no square root per seed and pixel).
    __shared__ float gs[KDE ? KDE_TAPS * KDE_TAPS : 1];
    DtJob J;
    int c;
    const int split = KDE ? DT_KDE_SPLIT : L.split;
    const int unit = xcd_unit(blockIdx.x, L.units);
    if (unit >= L.units) return;
    dt_job(L, unit / split, J, c);
    const int part = unit % split;  // this workgroup's rounds of the chunk
    const int r_beg = part * (DT_ROUNDS / split), r_end = r_beg + DT_ROUNDS / split;
    // The histogram accumulator is read by every table part; cleared here,
    // after the table pass and before the next compute's stats pass.
    if (c == 0 && part == 0)
        for (int i = threadIdx.x; i < DT_HIST; i += DT_NT) J.hist_acc[i] = 0;
    const int lambda = J.lambda;
    for (int i = threadIdx.x; i <= lambda; i += DT_NT) thr[i] = J.thr[i];
    if (KDE) {
        for (int i = threadIdx.x; i < KDE_TAPS * KDE_TAPS; i += DT_NT) gs[i] = J.gauss[isqrt24(i)];
    } else {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(J.tlut);
        for (int i = threadIdx.x; i < DT_LUT / 2; i += DT_NT) lut2[i] = src[i];
    }
    __syncthreads();
    const int W = J.W, H = J.H, n = J.W * J.H;
    const int64_t total = 3 * (int64_t)n;
    // beta as the table pass computed it (-1 / (2 beta) for the estimates; 0
    // when beta is 0: every weight is lambda).
    const int64_t pairs = (int64_t)(W - 1) * H + (int64_t)W * (H - 1);
    const double beta = pairs > 0 ? (double)J.stats->beta_sum / (double)pairs : 0.0;
    const float nib = beta > 0.0 ? (float)(-0.5 / beta) : 0.0f;
    // Loads of two rounds are in flight before either is used (the stores of
    // the six planes then overlap the next pair's loads); a workgroup with a
    // single round of the chunk runs half a pair.
    for (int r2 = r_beg; r2 < r_end; r2 += 2) {
    uint32_t ww[2][5], aa[2][3], bb[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i0 = c * DT_CHUNK + ((r2 + h) * DT_NT + (int)threadIdx.x) * 4;
        if (r2 + h < r_end && i0 < n) {
            load_words<5>(J.img, 3 * (int64_t)i0 - 4, total, ww[h]);  // p(-1) .. p(4)
            load_px12(J.img, 3 * ((int64_t)i0 - W), total, aa[h]);     // above
            load_px12(J.img, 3 * ((int64_t)i0 + W), total, bb[h]);     // below
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) ww[h][k] = 0;  // read by the density pass only
        }
    }
    // Kernel density of the eight pixels in one pass over the seed list (its
    // entries are uniform: scalar loads), eight independent f64 sums, each in
    // list order.
    double dens[2][4];
    if (KDE) {
        int px[2][4][3];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const Rgb q = px_at(ww[h], 4 + 3 * k);
                px[h][k][0] = q.c0;
                px[h][k][1] = q.c1;
                px[h][k][2] = q.c2;
                dens[h][k] = 0.0;
            }
        for (int j = 0; j < J.kde_n; ++j) {
            const uint32_t cj = J.kde[j];
            const int s0 = (int)(cj >> 16), s1 = (int)((cj >> 8) & 255u), s2 = (int)(cj & 255u);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int d0 = px[h][k][0] - s0, d1 = px[h][k][1] - s1, d2 = px[h][k][2] - s2;
                    const int sq2 = d0 * d0 + d1 * d1 + d2 * d2;
                    if (sq2 < KDE_TAPS * KDE_TAPS) dens[h][k] += (double)gs[sq2];
                }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i0 = c * DT_CHUNK + ((r2 + h) * DT_NT + (int)threadIdx.x) * 4;
        if (r2 + h >= r_end || i0 >= n) break;
        const uint32_t(&w)[5] = ww[h];
        const uint32_t(&a)[3] = aa[h];
        const uint32_t(&b)[3] = bb[h];
        int hw[5];  // weight of the pair (k - 1, k)
#pragma unroll
        for (int k = 0; k < 5; ++k)
            hw[k] = nlink_weight_est(sq(px_at(w, 1 + 3 * k), px_at(w, 4 + 3 * k)), thr, lambda, nib);
        int vp[4], vn[4], vu[4], vd[4], vl[4], vr[4];
        int x = i0 % W, y = i0 / W;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const Rgb pk = px_at(w, 4 + 3 * k);
            if (KDE) {
                const float pf = (float)dens[h][k];
                vp[k] = trunc_sat(300000.0 * (double)pf);
                vn[k] = trunc_sat(300000.0 * (double)(1.0f - pf));
            } else {
                const int idx = ((pk.c0 >> 4) << 8) | ((pk.c1 >> 4) << 4) | (pk.c2 >> 4);
                const int e0 = idx, e1 = 4096 + idx;
                vn[k] = (int)((lut2[e0 >> 1] >> (16 * (e0 & 1))) & 0xffffu);  // cost of the foreground label
                vp[k] = (int)((lut2[e1 >> 1] >> (16 * (e1 & 1))) & 0xffffu);  // cost of the background label
            }
            vl[k] = x > 0 ? hw[k] : 0;
            vr[k] = x + 1 < W ? hw[k + 1] : 0;
            vu[k] = y > 0 ? nlink_weight_est(sq(px_at(a, 3 * k), pk), thr, lambda, nib) : 0;
            vd[k] = y + 1 < H ? nlink_weight_est(sq(pk, px_at(b, 3 * k)), thr, lambda, nib) : 0;
            if (++x == W) {
                x = 0;
                ++y;
            }
        }
        const int *vals[6] = {vp, vn, vu, vd, vl, vr};
        if (J.aligned && i0 + 4 <= n) {
#pragma unroll
            for (int j = 0; j < 6; ++j)
                *reinterpret_cast<int4 *>(J.out[j] + i0) = make_int4(vals[j][0], vals[j][1], vals[j][2], vals[j][3]);
        } else {
#pragma unroll
            for (int j = 0; j < 6; ++j)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i0 + k < n) J.out[j][i0 + k] = vals[j][k];
        }
    }
    }
}

void dt_launch(const DtLaunch &L, int total_chunks, hipStream_t s) {
    const int jobs = L.njobs ? L.njobs : 1;
    const bool kde = L.njobs == 0 && L.one.model == GCFG_DT_MODEL_KDE;
    DtLaunch P = L;
    P.split = L.njobs == 0 && L.split > 1 ? std::min(L.split, DT_SPLIT_MAX) : 1;
    P.tparts = L.tparts == 16 || L.tparts == 32 ? L.tparts : 8;
    P.units = total_chunks * P.split;
    hipLaunchKernelGGL(k_dt_stats, dim3(xcd_grid(P.units)), dim3(DT_NT), 0, s, P);
    hipLaunchKernelGGL(k_dt_table, dim3(jobs * P.tparts), dim3(DT_NT), 0, s, P);
    const size_t lds = (size_t)(P.lds_lambda + 1) * sizeof(int32_t);
    P.units = total_chunks * P.split;
    if (kde) {
        P.units = total_chunks * DT_KDE_SPLIT;
        hipLaunchKernelGGL(k_dt_links<true>, dim3(xcd_grid(P.units)), dim3(DT_NT), lds, s, P);
    } else {
        hipLaunchKernelGGL(k_dt_links<false>, dim3(xcd_grid(P.units)), dim3(DT_NT), lds, s, P);
    }
}

__global__ void k_dt_freeze(const DtJob *jobs, int n) {
    const DtJob &J = jobs[blockIdx.x];
    if (!J.freeze_to) return;
    const int32_t *src = reinterpret_cast<const int32_t *>(J.stats);
    int32_t *dst = reinterpret_cast<int32_t *>(J.freeze_to);
    for (int i = threadIdx.x; i < (int)(sizeof(DtStats) / sizeof(int32_t)); i += blockDim.x) dst[i] = src[i];
}

void dt_freeze(const DtJob *d_jobs, int njobs, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_dt_freeze, dim3(njobs), dim3(64), 0, s, d_jobs, njobs);
}

__device__ __forceinline__ int sqdiff(const uint8_t *a, const uint8_t *b) {
    const int d0 = (int)a[0] - b[0], d1 = (int)a[1] - b[1], d2 = (int)a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
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
            const double p = kde_density(Rgb{a[0], a[1], a[2]}, kde_cols, kde_n, gauss);
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

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int grid_for(int64_t n) {
    const int64_t g = (n + DT_NT - 1) / DT_NT;
    return (int)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// ---------------------------------------------------------------------------
DtContext::~DtContext() {
    // Teardown: errors here have nowhere to go. Carved buffers belong to the
    // caller, who frees them once no work on them is queued.
    if (stream_ && !allocs_.empty()) (void)hipStreamSynchronize(stream_);
    for (void *p : allocs_) (void)hipFree(p);
    if (h_stats_) (void)hipHostFree(h_stats_);
}

// Sizes of the buffers init() carves, in order (256-byte aligned each).
static size_t carve_sizes(size_t n, size_t *sz) {
    size_t k = 0;
    sz[k++] = 3 * n;                                         // image
    sz[k++] = n;                                             // seeds
    sz[k++] = sizeof(DtStats);                               // stats
    sz[k++] = sizeof(DtStats);                               // frozen model
    sz[k++] = (n + DT_CHUNK - 1) / DT_CHUNK * DT_SPLIT_MAX * sizeof(unsigned long long);
    sz[k++] = DT_HIST * sizeof(int32_t);
    sz[k++] = DT_LUT * sizeof(uint16_t);
    sz[k++] = (DT_MAX_LAMBDA + 1) * sizeof(int32_t);
    sz[k++] = KDE_MAX_SEEDS * sizeof(uint32_t);
    sz[k++] = KDE_TAPS * sizeof(float);
    for (int i = 0; i < 6; ++i) sz[k++] = n * 4;             // output planes
    size_t tot = 0;
    for (size_t i = 0; i < k; ++i) tot += (sz[i] + 255) & ~(size_t)255;
    return tot;
}

size_t DtContext::block_bytes(size_t cap_px) {
    size_t sz[32];
    return carve_sizes(cap_px, sz);
}

int DtContext::init(int w, int h, hipStream_t stream, bool own_outputs, size_t cap_px, void *block) {
    if (w <= 0 || h <= 0 || (int64_t)w * h > ((int64_t)1 << 30) || cap_px > ((size_t)1 << 30)) {
        set_error("data term: shape must be positive and at most 2^30 pixels");
        return GCFG_ERR_INVALID;
    }
    W = w;
    H = h;
    stream_ = stream;
    const size_t n = std::max((size_t)W * H, cap_px);
    cap_px_ = n;
    if (block) {
        size_t sz[32];
        carve_sizes(n, sz);
        uint8_t *p = static_cast<uint8_t *>(block);
        size_t k = 0;
        auto take = [&](void **dst) {
            *dst = p;
            p += (sz[k] + 255) & ~(size_t)255;
            bytes_ += sz[k++];
        };
        take((void **)&own_img_);
        take((void **)&seeds_);
        take((void **)&d_stats_);
        take((void **)&d_frozen_);
        take((void **)&d_beta_part_);
        take((void **)&d_hist_acc_);
        take((void **)&d_tlut_);
        take((void **)&d_thr_);
        take((void **)&d_kde_);
        take((void **)&d_gauss_);
        take((void **)&pos_);
        take((void **)&neg_);
        take((void **)&up_);
        take((void **)&down_);
        take((void **)&left_);
        take((void **)&right_);
        carved_ = true;
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(DT_NT), 0, stream_, d_hist_acc_, (int64_t)DT_HIST, 0);
        GCFG_HIP(hipGetLastError());
        img_ = own_img_;
        return GCFG_OK;
    }
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
    rc |= dalloc((void **)&d_frozen_, sizeof(DtStats));
    rc |= dalloc((void **)&d_beta_part_, (n + DT_CHUNK - 1) / DT_CHUNK * DT_SPLIT_MAX * sizeof(unsigned long long));
    rc |= dalloc((void **)&d_hist_acc_, DT_HIST * sizeof(int32_t));
    rc |= dalloc((void **)&d_tlut_, DT_LUT * sizeof(uint16_t));
    rc |= dalloc((void **)&d_thr_, (DT_MAX_LAMBDA + 1) * sizeof(int32_t));
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
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(DT_NT), 0, stream_, d_hist_acc_, (int64_t)DT_HIST, 0);
    GCFG_HIP(hipGetLastError());
    img_ = own_img_;
    return clear_seeds();
}

int DtContext::reshape(int w, int h) {
    if (w <= 0 || h <= 0 || (size_t)w * h > cap_px_) {
        set_error("data term: shape beyond the buffers' capacity");
        return GCFG_ERR_INVALID;
    }
    W = w;
    H = h;
    frozen = false;
    kde_count = 0;
    computed = false;
    img_ = own_img_;
    return GCFG_OK;
}

int DtContext::upload_on(hipStream_t s, const uint8_t *h_rgb, const uint8_t *h_seeds) {
    if (!h_rgb) {
        set_error("data term: null image");
        return GCFG_ERR_INVALID;
    }
    GCFG_HIP(hipMemcpyAsync(own_img_, h_rgb, (size_t)3 * W * H, hipMemcpyHostToDevice, s));
    if (h_seeds) GCFG_HIP(hipMemcpyAsync(seeds_, h_seeds, (size_t)W * H, hipMemcpyHostToDevice, s));
    img_ = own_img_;
    computed = false;
    return GCFG_OK;
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

int DtContext::use_image(const uint8_t *d_rgb) {
    if (!d_rgb || ((uintptr_t)d_rgb & 3)) {
        set_error("data term: the device image must be 4-byte aligned");
        return GCFG_ERR_INVALID;
    }
    img_ = d_rgb;
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
    if (gauss_sigma_ != sigma) {
        // The host table is pageable memory: the copy has read it when the call returns.
        GCFG_HIP(hipMemcpy(d_gauss_, gauss, sizeof(gauss), hipMemcpyHostToDevice));
        gauss_sigma_ = sigma;
    }
    return GCFG_OK;
}

DtJob DtContext::job(int32_t *const *outs) const {
    DtJob j{};
    j.img = img_;
    j.seeds = seeds_;
    bool al = true;
    for (int k = 0; k < 6; ++k) {
        j.out[k] = outs[k];
        al &= aligned16(outs[k]);
    }
    j.W = W;
    j.H = H;
    j.chunk0 = 0;
    j.lambda = lambda;
    j.model = model;
    j.kde_n = kde_count;
    j.aligned = al ? 1 : 0;
    j.kde = d_kde_;
    j.gauss = d_gauss_;
    j.beta_part = d_beta_part_;
    j.hist_acc = d_hist_acc_;
    j.stats = d_stats_;
    j.frozen = frozen ? d_frozen_ : nullptr;
    j.tlut = d_tlut_;
    j.thr = d_thr_;
    return j;
}

int DtContext::compute(int32_t *pos, int32_t *neg, int32_t *up, int32_t *down, int32_t *left, int32_t *right) {
    if (lambda < 0 || lambda > DT_MAX_LAMBDA) {
        set_error("data term: lambda out of range");
        return GCFG_ERR_INVALID;
    }
    if (model == GCFG_DT_MODEL_KDE) {
        int rc = build_gauss();
        if (rc != GCFG_OK) return rc;
    }
    int32_t *outs[6] = {pos, neg, up, down, left, right};
    DtLaunch L{};
    L.one = job(outs);
    L.lds_lambda = lambda;
    // One image: a 1024^2 image has 256 chunks, one workgroup per CU; split
    // its chunks' rounds over more workgroups (up to four) and its table over
    // 32 so the passes have the wavefronts in flight that a batch has.
    const int nch = dt_chunks(W, H);
    L.split = nch <= 512 ? 4 : nch <= 2048 ? 2 : 1;
    L.tparts = 32;
    dt_launch(L, nch, stream_);
    GCFG_HIP(hipGetLastError());
    computed = true;
    return GCFG_OK;
}

DtJob DtContext::batch_job() const {
    int32_t *outs[6] = {pos_, neg_, up_, down_, left_, right_};
    DtJob j = job(outs);
    j.freeze_to = d_frozen_;
    return j;
}

int DtContext::freeze_model() {
    GCFG_HIP(hipMemcpyAsync(d_frozen_, d_stats_, sizeof(DtStats), hipMemcpyDeviceToDevice, stream_));
    frozen = true;
    return GCFG_OK;
}

int DtContext::compute_swap() {
    if (!pos_) {
        set_error("data term: no owned output planes");
        return GCFG_ERR_STATE;
    }
    const size_t n = cap_px_;  // the planes serve any later shape within the capacity
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

int DtContext::fetch_stats() {
    if (!computed) {
        set_error("data term: compute first");
        return GCFG_ERR_STATE;
    }
    if (!h_stats_) GCFG_HIP(hipHostMalloc((void **)&h_stats_, sizeof(DtStats), hipHostMallocDefault));
    GCFG_HIP(hipMemcpyAsync(h_stats_, d_stats_, sizeof(DtStats), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    stats = *h_stats_;
    pairs = (int64_t)(W - 1) * H + (int64_t)W * (H - 1);
    beta = pairs > 0 ? (double)stats.beta_sum / (double)pairs : 0.0;
    return GCFG_OK;
}

int DtContext::unquantized(double *h_tpos, double *h_tneg, double *h_right, double *h_down) {
    int rc = fetch_stats();
    if (rc != GCFG_OK) return rc;
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

// ---------------------------------------------------------------------------
// Fixed scratch of gcfg_dt_compute_noalloc: the per-chunk beta partials, the
// histogram counters, the statistics, the t-link table and the n-link
// thresholds of one image of at most GCFG_DT_NOALLOC_MAX_PIXELS pixels
// (161 KiB). Every GPU the library runs on holds its own copy until the
// process ends; calls that overlap on one GPU share it and race.
constexpr size_t NA_CHUNKS = ((size_t)GCFG_DT_NOALLOC_MAX_PIXELS + DT_CHUNK - 1) / DT_CHUNK;
constexpr size_t NA_OFF_BETA = 0;
constexpr size_t NA_OFF_HIST = NA_OFF_BETA + ((NA_CHUNKS * sizeof(unsigned long long) + 255) & ~(size_t)255);
constexpr size_t NA_OFF_STATS = NA_OFF_HIST + ((DT_HIST * sizeof(int32_t) + 255) & ~(size_t)255);
constexpr size_t NA_OFF_TLUT = NA_OFF_STATS + ((sizeof(DtStats) + 255) & ~(size_t)255);
constexpr size_t NA_OFF_THR = NA_OFF_TLUT + ((DT_LUT * sizeof(uint16_t) + 255) & ~(size_t)255);
constexpr size_t NA_BYTES = NA_OFF_THR + (((DT_MAX_LAMBDA + 1) * sizeof(int32_t) + 255) & ~(size_t)255);

static __device__ __align__(256) uint8_t scratch[NA_BYTES];

// Device address of `scratch` on the current device, looked up once per
// device (host code must never pass the symbol itself to a kernel).
constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static void *g_na_ptr[NA_MAX_DEVICES];

static int noalloc_scratch(uint8_t **out) {
    int dev = 0;
    GCFG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        set_error("gcfg_dt_compute_noalloc: device index beyond the supported range");
        return GCFG_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        GCFG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_ptr[dev] = p;
    }
    *out = static_cast<uint8_t *>(g_na_ptr[dev]);
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

// Batched histogram data term: every image's stats, table and links passes
// in one launch each.
struct gcfg_dt_batch {
    hipStream_t stream = nullptr;
    gcfg::StreamJoin join;
    int count = 0, total_chunks = 0;
    int32_t lambda = 50;
    gcfg::DtJob *d_jobs = nullptr;
    int32_t *d_cmap = nullptr;
    std::vector<void *> allocs;
    ~gcfg_dt_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
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
    if (!dt) return GCFG_ERR_INVALID;
    return dt->ctx.use_image(d_rgb);
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
    if (!dt || lambda < 0 || lambda > gcfg::DT_MAX_LAMBDA) {
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
    const int rc = dt->ctx.fetch_stats();
    if (rc != GCFG_OK) return rc;
    if (h_hist96) std::memcpy(h_hist96, &dt->ctx.stats.hist[0][0][0], 96 * sizeof(int32_t));
    if (h_counts2) std::memcpy(h_counts2, dt->ctx.stats.counts, 2 * sizeof(int32_t));
    return GCFG_OK;
}

int gcfg_dt_get_beta(gcfg_dataterm *dt, int64_t *sum, int64_t *pairs, double *beta) {
    if (!dt) return GCFG_ERR_INVALID;
    const int rc = dt->ctx.fetch_stats();
    if (rc != GCFG_OK) return rc;
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

int gcfg_dt_compute_noalloc(int32_t width, int32_t height, const uint8_t *d_rgb, const uint8_t *d_seeds,
                            int32_t lambda, int32_t *d_pos, int32_t *d_neg, int32_t *d_up, int32_t *d_down,
                            int32_t *d_left, int32_t *d_right, void *hip_stream) {
    using namespace gcfg;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t)GCFG_DT_NOALLOC_MAX_PIXELS) {
        set_error("gcfg_dt_compute_noalloc: the image must have 1 .. GCFG_DT_NOALLOC_MAX_PIXELS (2^26) pixels");
        return GCFG_ERR_INVALID;
    }
    if (lambda < 0 || lambda > DT_MAX_LAMBDA) {
        set_error("gcfg_dt_compute_noalloc: lambda out of range (0..4096)");
        return GCFG_ERR_INVALID;
    }
    if (!d_rgb || ((uintptr_t)d_rgb & 3) || ((uintptr_t)d_seeds & 3) || !d_pos || !d_neg || !d_up || !d_down ||
        !d_left || !d_right) {
        set_error("gcfg_dt_compute_noalloc: needs a 4-byte aligned image (and seed map) and six output planes");
        return GCFG_ERR_INVALID;
    }
    uint8_t *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc != GCFG_OK) return rc;
    hipStream_t st = (hipStream_t)hip_stream;
    DtJob j{};
    j.img = d_rgb;
    j.seeds = d_seeds;
    int32_t *o[6] = {d_pos, d_neg, d_up, d_down, d_left, d_right};
    bool al = true;
    for (int k = 0; k < 6; ++k) {
        j.out[k] = o[k];
        al &= aligned16(o[k]);
    }
    j.aligned = al ? 1 : 0;
    j.W = width;
    j.H = height;
    j.lambda = lambda;
    j.model = GCFG_DT_MODEL_HISTOGRAM;
    j.beta_part = reinterpret_cast<unsigned long long *>(base + NA_OFF_BETA);
    j.hist_acc = reinterpret_cast<int32_t *>(base + NA_OFF_HIST);
    j.stats = reinterpret_cast<DtStats *>(base + NA_OFF_STATS);
    j.tlut = reinterpret_cast<uint16_t *>(base + NA_OFF_TLUT);
    j.thr = reinterpret_cast<int32_t *>(base + NA_OFF_THR);
    // The histogram counters must start at zero: a kernel clears them (no
    // memset node, so the call can be captured into a graph).
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(DT_NT), 0, st, j.hist_acc, (int64_t)DT_HIST, 0);
    DtLaunch L{};
    L.one = j;
    L.lds_lambda = lambda;
    dt_launch(L, dt_chunks(width, height), st);
    GCFG_HIP(hipGetLastError());
    return GCFG_OK;
}

int gcfg_dt_batch_create(int32_t count, const gcfg_dt_image *images, int32_t lambda, gcfg_dt_batch **out) {
    using namespace gcfg;
    if (!out || count <= 0 || !images || lambda < 0 || lambda > DT_MAX_LAMBDA) {
        set_error("gcfg_dt_batch_create: invalid arguments (count > 0, lambda in 0..4096)");
        return GCFG_ERR_INVALID;
    }
    *out = nullptr;
    int64_t chunks = 0;
    for (int i = 0; i < count; ++i) {
        const gcfg_dt_image &im = images[i];
        const bool outs = im.d_pos && im.d_neg && im.d_up && im.d_down && im.d_left && im.d_right;
        if (im.width <= 0 || im.height <= 0 || (int64_t)im.width * im.height > ((int64_t)1 << 30) || !im.d_rgb ||
            ((uintptr_t)im.d_rgb & 3) || ((uintptr_t)im.d_seeds & 3) || !outs) {
            set_error("gcfg_dt_batch_create: image " + std::to_string(i) +
                      " needs a positive shape of at most 2^30 pixels, a 4-byte aligned image (and seed map) "
                      "and six output planes");
            return GCFG_ERR_INVALID;
        }
        chunks += dt_chunks(im.width, im.height);
        if (chunks >= ((int64_t)1 << 31) - 1) {
            set_error("gcfg_dt_batch_create: too many pixels in one batch");
            return GCFG_ERR_INVALID;
        }
    }
    gcfg_dt_batch *b = new (std::nothrow) gcfg_dt_batch();
    if (!b) return GCFG_ERR_NOMEM;
    b->count = count;
    b->lambda = lambda;
    b->total_chunks = (int)chunks;
    auto dalloc = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) {
            *p = nullptr;
            set_error("gcfg_dt_batch_create: device allocation failed");
            return GCFG_ERR_NOMEM;
        }
        b->allocs.push_back(*p);
        return GCFG_OK;
    };
    unsigned long long *beta_part = nullptr;
    int32_t *hist_acc = nullptr, *thr = nullptr;
    DtStats *stats = nullptr;
    uint16_t *tlut = nullptr;
    int rc = GCFG_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("gcfg_dt_batch_create: stream creation failed");
        rc = GCFG_ERR_HIP;
    }
    if (rc == GCFG_OK) rc = dalloc((void **)&b->d_jobs, (size_t)count * sizeof(DtJob));
    if (rc == GCFG_OK) rc = dalloc((void **)&b->d_cmap, (size_t)chunks * sizeof(int32_t));
    if (rc == GCFG_OK) rc = dalloc((void **)&beta_part, (size_t)chunks * sizeof(unsigned long long));
    if (rc == GCFG_OK) rc = dalloc((void **)&hist_acc, (size_t)count * DT_HIST * sizeof(int32_t));
    if (rc == GCFG_OK) rc = dalloc((void **)&stats, (size_t)count * sizeof(DtStats));
    if (rc == GCFG_OK) rc = dalloc((void **)&tlut, (size_t)count * DT_LUT * sizeof(uint16_t));
    if (rc == GCFG_OK) rc = dalloc((void **)&thr, (size_t)count * (lambda + 1) * sizeof(int32_t));
    if (rc != GCFG_OK) {
        delete b;
        return rc;
    }
    std::vector<DtJob> jobs(count);
    std::vector<int32_t> cmap((size_t)chunks);
    int c0 = 0;
    for (int i = 0; i < count; ++i) {
        const gcfg_dt_image &im = images[i];
        DtJob &j = jobs[i];
        j = DtJob{};
        j.img = im.d_rgb;
        j.seeds = im.d_seeds;
        int32_t *o[6] = {im.d_pos, im.d_neg, im.d_up, im.d_down, im.d_left, im.d_right};
        bool al = true;
        for (int k = 0; k < 6; ++k) {
            j.out[k] = o[k];
            al &= aligned16(o[k]);
        }
        j.aligned = al ? 1 : 0;
        j.W = im.width;
        j.H = im.height;
        j.chunk0 = c0;
        j.lambda = lambda;
        j.model = GCFG_DT_MODEL_HISTOGRAM;
        j.beta_part = beta_part + c0;
        j.hist_acc = hist_acc + (size_t)i * DT_HIST;
        j.stats = stats + i;
        j.tlut = tlut + (size_t)i * DT_LUT;
        j.thr = thr + (size_t)i * (lambda + 1);
        const int nc = dt_chunks(im.width, im.height);
        for (int k = 0; k < nc; ++k) cmap[(size_t)c0 + k] = i;
        c0 += nc;
    }
    hipError_t e = hipMemcpy(b->d_jobs, jobs.data(), jobs.size() * sizeof(DtJob), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_cmap, cmap.data(), cmap.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for((int64_t)count * DT_HIST)), dim3(DT_NT), 0, b->stream, hist_acc,
                           (int64_t)count * DT_HIST, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "gcfg_dt_batch_create");
    }
    *out = b;
    return GCFG_OK;
}

int gcfg_dt_batch_set_stream(gcfg_dt_batch *b, void *hip_stream) {
    if (!b) return GCFG_ERR_INVALID;
    return b->join.set((hipStream_t)hip_stream);
}

int gcfg_dt_batch_compute(gcfg_dt_batch *b) {
    using namespace gcfg;
    if (!b) return GCFG_ERR_INVALID;
    int rc = b->join.before(b->stream);
    if (rc != GCFG_OK) return rc;
    DtLaunch L{};
    L.jobs = b->d_jobs;
    L.cmap = b->d_cmap;
    L.njobs = b->count;
    L.lds_lambda = b->lambda;
    dt_launch(L, b->total_chunks, b->stream);
    GCFG_HIP(hipGetLastError());
    rc = b->join.after(b->stream);
    if (rc != GCFG_OK) return rc;
    GCFG_HIP(hipStreamSynchronize(b->stream));
    return GCFG_OK;
}

int gcfg_dt_batch_destroy(gcfg_dt_batch *b) {
    delete b;
    return GCFG_OK;
}

}  // extern "C"
