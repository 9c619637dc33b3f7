// Drop-in C++ layer: the reference's GC_* functions (GraphCut.cu:187-568)
// and DataTerm namespace (DataTerm.cpp:12-165) implemented over the C ABI.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "DataTerm.h"
#include "GraphCut.h"
#include "gcfg.h"

static void gc_check(int rc, const char *what) {
    if (rc != GCFG_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, gcfg_last_error());
        std::abort();
    }
}

// The 8-neighbourhood overloads (the arity GraphCut.h declares when the
// includer sets NEIGHBORHOOD to 8); this file is built with the default 4.
GlobalWrapper GC_Init(int width, int height, int *data_positive, int *data_negative, int penalty, int *up,
                      int *down, int *left, int *right, int *upleft, int *upright, int *downleft, int *downright);
void GC_SetEdges(GlobalWrapper *gw, int *up, int *down, int *left, int *right, int *upleft, int *upright,
                 int *downleft, int *downright);

static GlobalWrapper gc_init(int width, int height, int nbhd, int penalty, bool varying) {
    GlobalWrapper gw{};
    // Geometry of the tile grid (64 x 16 tiles stand where the reference's
    // 32 x 8 thread blocks did, GraphCut.cu:203-209).
    const int tw = 64, th = 16;
    gw.k.g.width = width;
    gw.k.g.height = height;
    gw.k.g.size = width * height;
    gw.k.block_x = width > 0 ? (width + tw - 1) / tw : 0;
    gw.block_y = height > 0 ? (height + th - 1) / th : 0;
    gw.block_count = gw.k.block_x * gw.block_y;
    gw.k.g.width_ex = gw.k.block_x * tw;
    gw.k.g.height_ex = gw.block_y * th;
    gw.k.g.size_ex = gw.k.g.width_ex * gw.k.g.height_ex;
    gw.penalty = penalty;
    gw.varying_edges = varying;
    gw.neighborhood = nbhd;
    gc_check(nbhd == 8 ? gcfg_graph_create_n8(width, height, &gw.graph) : gcfg_graph_create(width, height, &gw.graph),
             "GC_Init");
    gw.last_flow = new long long(0);
    return gw;
}

GlobalWrapper GC_Init(int width, int height, int *data_positive, int *data_negative, int penalty, int *up,
                      int *down, int *left, int *right) {
    GlobalWrapper gw = gc_init(width, height, 4, penalty, up != NULL);
    if (data_positive != NULL) {
        GC_SetDataterms(&gw, data_positive, data_negative);
        if (gw.varying_edges) GC_SetEdges(&gw, up, down, left, right);
        GC_SetGraph(gw);
    }
    return gw;
}

GlobalWrapper GC_Init(int width, int height, int *data_positive, int *data_negative, int penalty, int *up,
                      int *down, int *left, int *right, int *upleft, int *upright, int *downleft, int *downright) {
    GlobalWrapper gw = gc_init(width, height, 8, penalty, up != NULL);
    if (data_positive != NULL) {
        GC_SetDataterms(&gw, data_positive, data_negative);
        if (gw.varying_edges) GC_SetEdges(&gw, up, down, left, right, upleft, upright, downleft, downright);
        GC_SetGraph(gw);
    }
    return gw;
}

void GC_SetDataterms(GlobalWrapper *gw, int *data_positive, int *data_negative) {
    gw->data_positive = data_positive;
    gw->data_negative = data_negative;
}

void GC_SetEdges(GlobalWrapper *gw, int *up, int *down, int *left, int *right) {
    gw->varying_edges = true;
    gw->up = up;
    gw->down = down;
    gw->left = left;
    gw->right = right;
}

void GC_SetEdges(GlobalWrapper *gw, int *up, int *down, int *left, int *right, int *upleft, int *upright,
                 int *downleft, int *downright) {
    GC_SetEdges(gw, up, down, left, right);
    gw->upleft = upleft;
    gw->upright = upright;
    gw->downleft = downleft;
    gw->downright = downright;
}

void GC_SetPenalty(GlobalWrapper *gw, int p) {
    gw->varying_edges = false;
    gw->penalty = p;
}

void GC_SetGraph(GlobalWrapper gw) {
    gc_check(gcfg_graph_set_dataterms(gw.graph, gw.data_positive, gw.data_negative), "GC_SetGraph");
    if (gw.varying_edges && gw.neighborhood == 8)
        gc_check(gcfg_graph_set_edges8(gw.graph, gw.up, gw.down, gw.left, gw.right, gw.upleft, gw.upright,
                                       gw.downleft, gw.downright),
                 "GC_SetGraph");
    else if (gw.varying_edges)
        gc_check(gcfg_graph_set_edges(gw.graph, gw.up, gw.down, gw.left, gw.right), "GC_SetGraph");
    else
        gc_check(gcfg_graph_set_penalty(gw.graph, gw.penalty), "GC_SetGraph");
    gc_check(gcfg_graph_build(gw.graph), "GC_SetGraph");
}

void GC_Optimize(GlobalWrapper gw, int *label) {
    gcfg_stats st{};
    gc_check(gcfg_graph_optimize(gw.graph, label, &st), "GC_Optimize");
    if (gw.last_flow) *gw.last_flow = st.flow;
}

void GC_End(GlobalWrapper *gw) {
    if (gw->graph) gcfg_graph_destroy(gw->graph);
    delete gw->last_flow;
    *gw = GlobalWrapper{};
}

void GC_SetOptions(GlobalWrapper gw, const gcfg_options *opt) {
    gc_check(gcfg_graph_set_options(gw.graph, opt), "GC_SetOptions");
}

long long GC_GetFlow(GlobalWrapper gw) { return gw.last_flow ? *gw.last_flow : 0; }

// ---------------------------------------------------------------------------
namespace {
gcfg_dataterm *g_dt = nullptr;
int g_w = 0, g_h = 0;
// The reference's only colour model is the kernel density
// (DataTerm.cpp:105-137, DataTermKernels.cu:61-83): a driver that switches
// headers gets the same t-links unless it asks for the histogram model.
int g_model = GCFG_DT_MODEL_KDE;
std::vector<int> g_selection;

gcfg_dataterm *dt() {
    if (!g_dt) {
        std::fprintf(stderr, "DataTerm: init_data_term was not called\n");
        std::abort();
    }
    return g_dt;
}

int *link(int k) {
    int *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    gc_check(gcfg_dt_get_links(dt(), &p[0], &p[1], &p[2], &p[3], &p[4], &p[5]), "DataTerm::get");
    return p[k];
}
}  // namespace

namespace DataTerm {

int *getPos() { return link(0); }
int *getNeg() { return link(1); }
int *getUp() { return link(2); }
int *getDown() { return link(3); }
int *getLeft() { return link(4); }
int *getRight() { return link(5); }

void init_data_term(int w, int h) {
    if (g_dt) free_data_term();
    gc_check(gcfg_dt_create(w, h, &g_dt), "DataTerm::init_data_term");
    gc_check(gcfg_dt_set_model(g_dt, g_model), "DataTerm::init_data_term");
    g_w = w;
    g_h = h;
    g_selection.assign((size_t)w * h, 0);
}

void free_data_term() {
    if (g_dt) gcfg_dt_destroy(g_dt);
    g_dt = nullptr;
}

void setImage(unsigned char *h_img) { gc_check(gcfg_dt_set_image(dt(), h_img), "DataTerm::setImage"); }

void setSelection(int *h_select) {
    for (size_t i = 0; i < g_selection.size(); ++i) g_selection[i] = h_select[i];
}

void reset() {
    gc_check(gcfg_dt_reset(dt()), "DataTerm::reset");
    for (auto &v : g_selection) v = 0;
}

// DataTerm.cpp:105-127: the 5x5 square at (x..x+4, y..y+4); the sigma
// argument is unused there too (the kernel width is setSigmaGaussian's).
void selectOnMap(int x, int y, int sigma) {
    (void)sigma;
    if (g_model == GCFG_DT_MODEL_KDE)
        gc_check(gcfg_dt_kde_select(dt(), x, y), "DataTerm::selectOnMap");
    else
        gc_check(gcfg_dt_add_seed_rect(dt(), x, y, 5, 5, GCFG_SEED_FG), "DataTerm::selectOnMap");
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            const int xx = x + i, yy = y + j;
            if (xx >= 0 && yy >= 0 && xx < g_w && yy < g_h) g_selection[(size_t)yy * g_w + xx] = 1;
        }
}

void selectRect(int x0, int y0, int w, int h, int cls) {
    gc_check(gcfg_dt_add_seed_rect(dt(), x0, y0, w, h, cls), "DataTerm::selectRect");
}

void drawSelection(unsigned char *inp, unsigned char *out) {
    const size_t n = (size_t)g_w * g_h;
    for (size_t i = 0; i < n; ++i) {
        const bool sel = g_selection[i] == 1;
        out[3 * i + 0] = sel ? 0 : inp[3 * i + 0];
        out[3 * i + 1] = sel ? 0 : inp[3 * i + 1];
        out[3 * i + 2] = sel ? 255 : inp[3 * i + 2];
    }
}

void getDataTermsMixture() { gc_check(gcfg_dt_compute(dt()), "DataTerm::getDataTermsMixture"); }

void setSigmaGaussian(int sigma) { gc_check(gcfg_dt_set_sigma(dt(), sigma), "DataTerm::setSigmaGaussian"); }

// May precede init_data_term; the model then applies from init on.
void setModel(int model) {
    if (g_dt) gc_check(gcfg_dt_set_model(g_dt, model), "DataTerm::setModel");
    else if (model != GCFG_DT_MODEL_HISTOGRAM && model != GCFG_DT_MODEL_KDE)
        gc_check(GCFG_ERR_INVALID, "DataTerm::setModel");
    g_model = model;
}

void setLambda(int lambda) { gc_check(gcfg_dt_set_lambda(dt(), lambda), "DataTerm::setLambda"); }

}  // namespace DataTerm
