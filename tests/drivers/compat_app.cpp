// Drop-in check of include/GraphCut.h and include/DataTerm.h: the call
// sequence of the reference's interactive application
// (ComplexApplication.cu:105-135 init/process, :180-193 batch mode,
// :202-207 video loop), compiled against libgcfg.so unchanged except for
// hipMalloc in place of cudaMalloc.
//
// Input file (little-endian int32 header, then bytes):
//   W, H, penalty, sigma, model (-1: leave the default), nclicks, nframes,
//   nclicks x (x, y)            selectOnMap clicks on frame 0
//   nframes x W*H*3 u8          frames (BGR as stored by the caller)
// For every frame: setImage(frame), getDataTermsMixture(), then
// GC_SetDataterms(getPos(), getNeg()), GC_SetPenalty (penalty > 0) or
// GC_SetEdges(getUp(), ...) (penalty == 0), GC_SetGraph, GC_Optimize,
// setSelection(labels), drawSelection(frame, out).
// Output file per frame: int64 flow, int32 pos[W*H], neg[W*H], label[W*H],
// u8 out[W*H*3].
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "DataTerm.h"
#include "GraphCut.h"

static void die(const char *m) {
    std::fprintf(stderr, "compat_app: %s\n", m);
    std::exit(1);
}

int main(int argc, char **argv) {
    if (argc != 3) die("usage: compat_app <in> <out>");
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) die("cannot open input");
    int32_t hdr[7];
    if (std::fread(hdr, sizeof(int32_t), 7, in) != 7) die("short header");
    const int W = hdr[0], H = hdr[1], penalty = hdr[2], sigma = hdr[3], model = hdr[4], nclicks = hdr[5],
              nframes = hdr[6];
    if (W <= 0 || H <= 0 || nclicks < 0 || nframes <= 0) die("bad header");
    std::vector<int32_t> clicks(2 * (size_t)nclicks);
    if (nclicks && std::fread(clicks.data(), sizeof(int32_t), clicks.size(), in) != clicks.size())
        die("short clicks");
    const size_t n = (size_t)W * H;
    std::vector<unsigned char> frames(3 * n * nframes);
    if (std::fread(frames.data(), 1, frames.size(), in) != frames.size()) die("short frames");
    std::fclose(in);
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) die("cannot open output");

    GlobalWrapper gw = GC_Init(W, H);  // data terms and penalty set later (ComplexApplication.cu:105)
    // The reference's wrapper fields (GraphCut.h:60-128) read as its callers read them.
    if (gw.k.g.width != W || gw.k.g.height != H || gw.k.g.size != W * H || gw.k.g.width_ex < W ||
        gw.k.g.height_ex < H || gw.block_count != gw.k.block_x * gw.block_y ||
        gw.k.g.size_ex != gw.k.g.width_ex * gw.k.g.height_ex)
        die("GlobalWrapper geometry fields");
    if (model >= 0) DataTerm::setModel(model);  // -1: the drop-in's default model
    DataTerm::init_data_term(W, H);
    DataTerm::setSigmaGaussian(sigma);
    std::vector<int> labels(n), hpos(n), hneg(n);
    std::vector<unsigned char> shown(3 * n);

    DataTerm::setImage(frames.data());
    for (int c = 0; c < nclicks; ++c) DataTerm::selectOnMap(clicks[2 * c], clicks[2 * c + 1], 8);
    for (int f = 0; f < nframes; ++f) {
        unsigned char *frame = frames.data() + 3 * n * f;
        DataTerm::setImage(frame);
        DataTerm::getDataTermsMixture();
        GC_SetDataterms(&gw, DataTerm::getPos(), DataTerm::getNeg());
        if (penalty > 0)
            GC_SetPenalty(&gw, penalty);
        else
            GC_SetEdges(&gw, DataTerm::getUp(), DataTerm::getDown(), DataTerm::getLeft(), DataTerm::getRight());
        GC_SetGraph(gw);
        GC_Optimize(gw, labels.data());
        DataTerm::setSelection(labels.data());
        DataTerm::drawSelection(frame, shown.data());
        if (hipMemcpy(hpos.data(), DataTerm::getPos(), 4 * n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hneg.data(), DataTerm::getNeg(), 4 * n, hipMemcpyDeviceToHost) != hipSuccess)
            die("hipMemcpy");
        const long long flow = GC_GetFlow(gw);
        std::fwrite(&flow, sizeof(flow), 1, out);
        std::fwrite(hpos.data(), 4, n, out);
        std::fwrite(hneg.data(), 4, n, out);
        std::fwrite(labels.data(), 4, n, out);
        std::fwrite(shown.data(), 1, 3 * n, out);
    }
    std::fclose(out);
    DataTerm::free_data_term();
    GC_End(&gw);
    return 0;
}
