/*
 * Seeded synthetic-input generators for the binary-MRF min-cut path.
 *
 * Every generator is a pure host function of its arguments (no GPU work), so
 * the CPU oracle and the HIP path can be fed byte-identical inputs on any box.
 * Randomness: splitmix64 seeding + xoshiro256** stream; normals by the basic
 * Box-Muller transform (cosine branch only, one normal per two uniforms).
 *
 * Workloads generated here:
 *  - ellipse image: W x H RGB u8, axis-aligned ellipse centred at (W/2, H/2)
 *    with semi-axes (0.3 W, 0.35 H); inside ~ N((200,80,60), 20), outside
 *    ~ N((60,120,180), 20), each channel clip(llround(.), 0, 255);
 *  - seed map: 20x20 foreground square at the centre, 20x20 background
 *    squares in the four corners (1 = foreground, 2 = background);
 *  - stress terminal links: v = min(llround(32 * U^(-1/1.5)), 2^20),
 *    U ~ Uniform(0, 1], a fair coin puts v on pos or neg (the other is 0);
 *  - Perlin gray image: 4-octave gradient noise (base cell 64 px, lacunarity
 *    2, gain 0.5) replicated into three channels;
 *  - log-uniform batch shapes: area log-uniform on [amin, amax], aspect
 *    log-uniform on [1/2, 2], W = llround(sqrt(A r)), H = llround(sqrt(A / r)).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "gcfg.h"

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* Uniform on (0, 1]. */
static double rng_uniform_oc(rng_t *r) { return (double)((rng_next(r) >> 11) + 1) * 0x1.0p-53; }

/* Uniform on [0, 1). */
static double rng_uniform_co(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static double rng_normal(rng_t *r) {
    const double u1 = rng_uniform_oc(r);
    const double u2 = rng_uniform_co(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static uint8_t clip_u8(double v) {
    long long q = llround(v);
    if (q < 0) q = 0;
    if (q > 255) q = 255;
    return (uint8_t)q;
}

int gcfg_gen_ellipse_image(int width, int height, uint64_t seed, uint8_t *rgb) {
    if (width <= 0 || height <= 0 || !rgb) return GCFG_ERR_INVALID;
    static const double mu_fg[3] = {200.0, 80.0, 60.0};
    static const double mu_bg[3] = {60.0, 120.0, 180.0};
    const double sigma = 20.0;
    const double cx = width / 2.0, cy = height / 2.0;
    const double ax = 0.3 * width, ay = 0.35 * height;
    rng_t r;
    rng_seed(&r, seed);
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            const double dx = (x - cx) / ax, dy = (y - cy) / ay;
            const double *mu = (dx * dx + dy * dy <= 1.0) ? mu_fg : mu_bg;
            uint8_t *px = rgb + 3 * ((size_t)y * width + x);
            for (int c = 0; c < 3; ++c) px[c] = clip_u8(mu[c] + sigma * rng_normal(&r));
        }
    }
    return GCFG_OK;
}

static void fill_rect(uint8_t *map, int width, int height, int x0, int y0, int w, int h, uint8_t cls) {
    for (int y = y0; y < y0 + h; ++y) {
        if (y < 0 || y >= height) continue;
        for (int x = x0; x < x0 + w; ++x) {
            if (x < 0 || x >= width) continue;
            map[(size_t)y * width + x] = cls;
        }
    }
}

int gcfg_gen_seed_map(int width, int height, uint8_t *seeds) {
    if (width <= 0 || height <= 0 || !seeds) return GCFG_ERR_INVALID;
    const int s = 20;
    memset(seeds, 0, (size_t)width * height);
    fill_rect(seeds, width, height, 0, 0, s, s, GCFG_SEED_BG);
    fill_rect(seeds, width, height, width - s, 0, s, s, GCFG_SEED_BG);
    fill_rect(seeds, width, height, 0, height - s, s, s, GCFG_SEED_BG);
    fill_rect(seeds, width, height, width - s, height - s, s, s, GCFG_SEED_BG);
    fill_rect(seeds, width, height, width / 2 - s / 2, height / 2 - s / 2, s, s, GCFG_SEED_FG);
    return GCFG_OK;
}

int gcfg_gen_stress_tlinks(int width, int height, uint64_t seed, int32_t *pos, int32_t *neg) {
    if (width <= 0 || height <= 0 || !pos || !neg) return GCFG_ERR_INVALID;
    rng_t r;
    rng_seed(&r, seed);
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; ++i) {
        const double u = rng_uniform_oc(&r);
        double v = 32.0 * pow(u, -1.0 / 1.5);
        long long q = (v >= 1048576.0) ? 1048576 : llround(v);
        if (q > 1048576) q = 1048576;
        const int coin = (int)(rng_next(&r) >> 63);
        pos[i] = coin ? (int32_t)q : 0;
        neg[i] = coin ? 0 : (int32_t)q;
    }
    return GCFG_OK;
}

/* Classic gradient noise on a seeded 256-entry permutation (quintic fade). */
typedef struct {
    int perm[512];
} perlin_t;

static void perlin_init(perlin_t *p, rng_t *r) {
    int a[256];
    for (int i = 0; i < 256; ++i) a[i] = i;
    for (int i = 255; i > 0; --i) {
        const int j = (int)(rng_next(r) % (uint64_t)(i + 1));
        const int t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
    for (int i = 0; i < 512; ++i) p->perm[i] = a[i & 255];
}

static double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

static double grad2(int h, double x, double y) {
    switch (h & 7) {
        case 0: return x + y;
        case 1: return x - y;
        case 2: return -x + y;
        case 3: return -x - y;
        case 4: return x;
        case 5: return -x;
        case 6: return y;
        default: return -y;
    }
}

static double perlin2(const perlin_t *p, double x, double y) {
    const double fx = floor(x), fy = floor(y);
    const int xi = ((int)fx) & 255, yi = ((int)fy) & 255;
    const double xf = x - fx, yf = y - fy;
    const double u = fade(xf), v = fade(yf);
    const int aa = p->perm[p->perm[xi] + yi], ab = p->perm[p->perm[xi] + yi + 1];
    const int ba = p->perm[p->perm[xi + 1] + yi], bb = p->perm[p->perm[xi + 1] + yi + 1];
    const double x1 = grad2(aa, xf, yf) + u * (grad2(ba, xf - 1.0, yf) - grad2(aa, xf, yf));
    const double x2 = grad2(ab, xf, yf - 1.0) + u * (grad2(bb, xf - 1.0, yf - 1.0) - grad2(ab, xf, yf - 1.0));
    return x1 + v * (x2 - x1);
}

int gcfg_gen_perlin_image(int width, int height, uint64_t seed, uint8_t *rgb) {
    if (width <= 0 || height <= 0 || !rgb) return GCFG_ERR_INVALID;
    rng_t r;
    rng_seed(&r, seed);
    perlin_t p;
    perlin_init(&p, &r);
    const double ox = rng_uniform_co(&r) * 256.0, oy = rng_uniform_co(&r) * 256.0;
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            double sum = 0.0, amp = 1.0, freq = 1.0 / 64.0, norm = 0.0;
            for (int o = 0; o < 4; ++o) {
                sum += amp * perlin2(&p, ox + x * freq, oy + y * freq);
                norm += amp;
                amp *= 0.5;
                freq *= 2.0;
            }
            const uint8_t g = clip_u8(255.0 * (0.5 + 0.5 * sum / norm));
            uint8_t *px = rgb + 3 * ((size_t)y * width + x);
            px[0] = px[1] = px[2] = g;
        }
    }
    return GCFG_OK;
}

int gcfg_gen_batch_shapes(int count, uint64_t seed, double area_min, double area_max, int32_t *widths,
                          int32_t *heights) {
    if (count < 0 || !widths || !heights || area_min <= 0.0 || area_max < area_min) return GCFG_ERR_INVALID;
    rng_t r;
    rng_seed(&r, seed);
    const double la = log(area_min), lb = log(area_max);
    const double lr0 = log(0.5), lr1 = log(2.0);
    for (int i = 0; i < count; ++i) {
        const double a = exp(la + rng_uniform_co(&r) * (lb - la));
        const double asp = exp(lr0 + rng_uniform_co(&r) * (lr1 - lr0));
        long long w = llround(sqrt(a * asp)), h = llround(sqrt(a / asp));
        widths[i] = (int32_t)(w < 3 ? 3 : w);
        heights[i] = (int32_t)(h < 3 ? 3 : h);
    }
    return GCFG_OK;
}

uint64_t gcfg_gen_mix_seed(uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (stream + 1));
    return splitmix64(&x);
}
