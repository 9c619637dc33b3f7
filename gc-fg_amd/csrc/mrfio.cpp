// Host-side file formats of the reference's drivers: the Middlebury 2-label
// MRF text file (solveMRF.cu:10-65 reader, multiplyMRF.cpp:64-92 writer), the
// data-term file and the tiled constant-penalty MRF built from it
// (multiplyMRF.cpp:94-145, Application.cu:15-24) and the plain PPM/PGM images
// the drivers read and write (Application.cu:55-75, solveMRF.cu:113-127).
// Pure host code, no device work.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"

using gcfg::set_error;

namespace {

// Whitespace-separated integer reader over a whole file.
struct IntReader {
    std::vector<char> buf;
    size_t pos = 0;
    bool open(const char *path) {
        FILE *f = std::fopen(path, "rb");
        if (!f) return false;
        std::fseek(f, 0, SEEK_END);
        const long n = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buf.resize(n > 0 ? (size_t)n : 0);
        const size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
        std::fclose(f);
        buf.resize(got);
        return true;
    }
    void skip_space() {
        while (pos < buf.size()) {
            if (buf[pos] == '#') {  // PNM comment to end of line
                while (pos < buf.size() && buf[pos] != '\n') ++pos;
            } else if (std::isspace((unsigned char)buf[pos])) {
                ++pos;
            } else {
                break;
            }
        }
    }
    bool next_int(long long &v) {
        skip_space();
        if (pos >= buf.size()) return false;
        char *end = nullptr;
        const char *start = buf.data() + pos;
        // strtoll needs a terminator within the buffer: copy the token.
        char tok[32];
        size_t k = 0;
        while (pos + k < buf.size() && k < sizeof(tok) - 1 && !std::isspace((unsigned char)buf[pos + k])) {
            tok[k] = start[k];
            ++k;
        }
        tok[k] = 0;
        v = std::strtoll(tok, &end, 10);
        if (end == tok || *end) return false;
        pos += k;
        return true;
    }
    bool next_word(std::string &w) {
        skip_space();
        w.clear();
        while (pos < buf.size() && !std::isspace((unsigned char)buf[pos])) w.push_back(buf[pos++]);
        return !w.empty();
    }
};

bool read_ints(IntReader &r, int32_t *dst, size_t n) {
    long long v;
    for (size_t i = 0; i < n; ++i) {
        if (!r.next_int(v) || v < INT32_MIN || v > INT32_MAX) return false;
        if (dst) dst[i] = (int32_t)v;
    }
    return true;
}

int fail(const std::string &msg) {
    set_error(msg);
    return GCFG_ERR_INVALID;
}

}  // namespace

extern "C" {

int gcfg_mrf_info(const char *path, int32_t *width, int32_t *height, int32_t *n_labels) {
    IntReader r;
    if (!path || !r.open(path)) return fail(std::string("gcfg_mrf_info: cannot open ") + (path ? path : "(null)"));
    long long w, h, l;
    if (!r.next_int(w) || !r.next_int(h) || !r.next_int(l) || w <= 0 || h <= 0 || w * h > (1ll << 31))
        return fail("gcfg_mrf_info: bad header (expected 'width height labels')");
    if (width) *width = (int32_t)w;
    if (height) *height = (int32_t)h;
    if (n_labels) *n_labels = (int32_t)l;
    return GCFG_OK;
}

int gcfg_mrf_load(const char *path, int32_t width, int32_t height, int32_t *pos, int32_t *neg, int32_t *up,
                  int32_t *down, int32_t *left, int32_t *right) {
    IntReader r;
    if (!path || !r.open(path)) return fail(std::string("gcfg_mrf_load: cannot open ") + (path ? path : "(null)"));
    long long w, h, l;
    if (!r.next_int(w) || !r.next_int(h) || !r.next_int(l)) return fail("gcfg_mrf_load: bad header");
    if (w != width || h != height) return fail("gcfg_mrf_load: shape differs from gcfg_mrf_info");
    if (l != 2) return fail("gcfg_mrf_load: only 2-label files are supported (solveMRF.cu:24)");
    const size_t n = (size_t)w * (size_t)h;
    // ground truth (skipped), data terms, then the cues
    if (!read_ints(r, nullptr, n)) return fail("gcfg_mrf_load: truncated ground-truth block");
    if (!read_ints(r, pos, n) || !read_ints(r, neg, n)) return fail("gcfg_mrf_load: truncated data terms");
    std::vector<int32_t> hcue(n, 0), vcue(n, 0);
    for (long long y = 0; y < h; ++y)
        for (long long x = 0; x + 1 < w; ++x)
            if (!read_ints(r, &hcue[(size_t)(x + y * w)], 1)) return fail("gcfg_mrf_load: truncated horizontal cues");
    for (long long y = 0; y + 1 < h; ++y)
        for (long long x = 0; x < w; ++x)
            if (!read_ints(r, &vcue[(size_t)(y * w + x)], 1)) return fail("gcfg_mrf_load: truncated vertical cues");
    // The reference zeroes the last row of both cue planes after reading
    // (solveMRF.cu:60-63); its last column of hCue is never written, here 0.
    for (long long x = 0; x < w; ++x) {
        vcue[(size_t)((h - 1) * w + x)] = 0;
        hcue[(size_t)((h - 1) * w + x)] = 0;
    }
    // Cue to direction mapping of solveMRF.cu:87-90: up = down = vCue,
    // left = right = hCue.
    const size_t bytes = n * sizeof(int32_t);
    if (up) std::memcpy(up, vcue.data(), bytes);
    if (down) std::memcpy(down, vcue.data(), bytes);
    if (left) std::memcpy(left, hcue.data(), bytes);
    if (right) std::memcpy(right, hcue.data(), bytes);
    return GCFG_OK;
}

int gcfg_mrf_write(const char *path, int32_t width, int32_t height, int32_t mult, const int32_t *pos,
                   const int32_t *neg, const int32_t *hcue, const int32_t *vcue) {
    if (!path || width < 2 || height < 1 || mult < 1 || !pos || !neg || !hcue || !vcue)
        return fail("gcfg_mrf_write: invalid arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("gcfg_mrf_write: cannot create ") + path);
    // Layout of multiplyMRF.cpp:64-92: the image repeated `mult` times
    // horizontally, ground truth all 1, every value followed by a space.
    const long long W = (long long)width * mult;
    std::fprintf(f, "%lld %d %d ", W, height, 2);
    for (long long i = 0; i < W * height; ++i) std::fputs("1 ", f);
    for (int i = 0; i < height; ++i)
        for (long long j = 0; j < W; ++j) std::fprintf(f, "%d ", pos[(size_t)i * width + (size_t)(j % width)]);
    for (int i = 0; i < height; ++i)
        for (long long j = 0; j < W; ++j) std::fprintf(f, "%d ", neg[(size_t)i * width + (size_t)(j % width)]);
    for (int y = 0; y < height; ++y)
        for (long long x = 0; x < W - 1; ++x)
            std::fprintf(f, "%d ", hcue[(size_t)y * width + (size_t)(x % (width - 1))]);
    for (int y = 0; y + 1 < height; ++y)
        for (long long x = 0; x < W; ++x) std::fprintf(f, "%d ", vcue[(size_t)y * width + (size_t)(x % width)]);
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? GCFG_OK : fail("gcfg_mrf_write: write error");
}

int gcfg_dat_info(const char *path, int32_t *width, int32_t *height) {
    IntReader r;
    if (!path || !r.open(path)) return fail(std::string("gcfg_dat_info: cannot open ") + (path ? path : "(null)"));
    long long h, w;
    if (!r.next_int(h) || !r.next_int(w) || w <= 0 || h <= 0 || w * h > (1ll << 30))
        return fail("gcfg_dat_info: bad header (expected 'height width')");
    if (width) *width = (int32_t)w;
    if (height) *height = (int32_t)h;
    return GCFG_OK;
}

int gcfg_dat_load(const char *path, int32_t width, int32_t height, int32_t *pos, int32_t *neg) {
    IntReader r;
    if (!path || !r.open(path)) return fail(std::string("gcfg_dat_load: cannot open ") + (path ? path : "(null)"));
    long long h, w;
    if (!r.next_int(h) || !r.next_int(w)) return fail("gcfg_dat_load: bad header");
    if (w != width || h != height) return fail("gcfg_dat_load: shape differs from gcfg_dat_info");
    const size_t n = (size_t)w * (size_t)h;
    // Application.cu:18-23 / multiplyMRF.cpp:138-143: every neg value first, then every pos value.
    if (!read_ints(r, neg, n) || !read_ints(r, pos, n)) return fail("gcfg_dat_load: truncated data terms");
    return GCFG_OK;
}

int gcfg_dat_write(const char *path, int32_t width, int32_t height, const int32_t *pos, const int32_t *neg) {
    if (!path || width <= 0 || height <= 0 || !pos || !neg) return fail("gcfg_dat_write: invalid arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("gcfg_dat_write: cannot create ") + path);
    const size_t n = (size_t)width * height;
    std::fprintf(f, "%d %d\n", height, width);
    for (size_t i = 0; i < n; ++i) std::fprintf(f, "%d%c", neg[i], (i + 1) % (size_t)width ? ' ' : '\n');
    for (size_t i = 0; i < n; ++i) std::fprintf(f, "%d%c", pos[i], (i + 1) % (size_t)width ? ' ' : '\n');
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? GCFG_OK : fail("gcfg_dat_write: write error");
}

int gcfg_mrf_write_tiled(const char *path, int32_t out_w, int32_t out_h, int32_t width, int32_t height,
                         const int32_t *pos, const int32_t *neg, int32_t penalty) {
    if (!path || out_w < 1 || out_h < 1 || width < 1 || height < 1 || !pos || !neg)
        return fail("gcfg_mrf_write_tiled: invalid arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("gcfg_mrf_write_tiled: cannot create ") + path);
    // Layout of writeRihannaMRFData (multiplyMRF.cpp:94-122): header, ground
    // truth all 1, the data terms repeated over the output (source column
    // x mod width; source row y, which the reference indexes without a wrap
    // and so only defines for out_h <= height: rows past the source repeat it
    // here), then every horizontal and vertical cue equal to the penalty.
    std::fprintf(f, "%d %d 2 ", out_w, out_h);
    const long long N = (long long)out_w * out_h;
    for (long long i = 0; i < N; ++i) std::fputs("1 ", f);
    const int32_t *planes[2] = {pos, neg};
    for (const int32_t *pl : planes)
        for (int y = 0; y < out_h; ++y)
            for (int x = 0; x < out_w; ++x)
                std::fprintf(f, "%d ", pl[(size_t)(y % height) * width + (size_t)(x % width)]);
    for (long long i = 0; i < (long long)out_h * (out_w - 1); ++i) std::fprintf(f, "%d ", penalty);
    for (long long i = 0; i < (long long)(out_h - 1) * out_w; ++i) std::fprintf(f, "%d ", penalty);
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? GCFG_OK : fail("gcfg_mrf_write_tiled: write error");
}

int gcfg_ppm_info(const char *path, int32_t *width, int32_t *height) {
    IntReader r;
    if (!path || !r.open(path)) return fail(std::string("gcfg_ppm_info: cannot open ") + (path ? path : "(null)"));
    std::string magic;
    long long w, h, mx;
    if (!r.next_word(magic) || (magic != "P3" && magic != "P6") || !r.next_int(w) || !r.next_int(h) ||
        !r.next_int(mx) || w <= 0 || h <= 0 || mx != 255)
        return fail("gcfg_ppm_info: expected a P3 or P6 image with maxval 255");
    if (width) *width = (int32_t)w;
    if (height) *height = (int32_t)h;
    return GCFG_OK;
}

int gcfg_ppm_read(const char *path, int32_t width, int32_t height, uint8_t *rgb) {
    IntReader r;
    if (!path || !rgb || !r.open(path)) return fail("gcfg_ppm_read: cannot open file");
    std::string magic;
    long long w, h, mx;
    if (!r.next_word(magic) || !r.next_int(w) || !r.next_int(h) || !r.next_int(mx))
        return fail("gcfg_ppm_read: bad header");
    if (w != width || h != height || mx != 255) return fail("gcfg_ppm_read: shape or maxval mismatch");
    const size_t n = (size_t)w * (size_t)h * 3;
    if (magic == "P6") {
        // Exactly one whitespace byte separates the header from the raster.
        const size_t start = r.pos + 1;
        if (start + n > r.buf.size()) return fail("gcfg_ppm_read: truncated raster");
        std::memcpy(rgb, r.buf.data() + start, n);
        return GCFG_OK;
    }
    if (magic != "P3") return fail("gcfg_ppm_read: not a P3/P6 file");
    long long v;
    for (size_t i = 0; i < n; ++i) {
        if (!r.next_int(v) || v < 0 || v > 255) return fail("gcfg_ppm_read: bad or truncated raster");
        rgb[i] = (uint8_t)v;
    }
    return GCFG_OK;
}

int gcfg_write_label_ppm(const char *path, int32_t width, int32_t height, const int32_t *label, const uint8_t *rgb) {
    if (!path || width <= 0 || height <= 0 || !label) return fail("gcfg_write_label_ppm: invalid arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("gcfg_write_label_ppm: cannot create ") + path);
    // Text P3 as solveMRF.cu:113-127 (white where label is 1) or
    // Application.cu:61-75 (the input colour where label is 1); black elsewhere.
    std::fprintf(f, "P3 %d %d 255 \n", width, height);
    for (int i = 0; i < height; ++i) {
        for (int j = 0; j < width; ++j) {
            const size_t k = (size_t)i * width + j;
            if (label[k]) {
                if (rgb)
                    std::fprintf(f, "%d %d %d ", rgb[3 * k], rgb[3 * k + 1], rgb[3 * k + 2]);
                else
                    std::fputs("255 255 255 ", f);
            } else {
                std::fputs("0 0 0 ", f);
            }
        }
        std::fputc('\n', f);
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? GCFG_OK : fail("gcfg_write_label_ppm: write error");
}

int gcfg_write_label_pgm(const char *path, int32_t width, int32_t height, const int32_t *label) {
    if (!path || width <= 0 || height <= 0 || !label) return fail("gcfg_write_label_pgm: invalid arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("gcfg_write_label_pgm: cannot create ") + path);
    std::fprintf(f, "P5\n%d %d\n255\n", width, height);
    std::vector<uint8_t> row((size_t)width);
    for (int i = 0; i < height; ++i) {
        for (int j = 0; j < width; ++j) row[j] = label[(size_t)i * width + j] ? 255 : 0;
        std::fwrite(row.data(), 1, row.size(), f);
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok ? GCFG_OK : fail("gcfg_write_label_pgm: write error");
}

}  // extern "C"
