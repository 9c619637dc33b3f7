// C ABI: plain graph (GraphCut.cu GC_*), batched solve, options and errors.
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "engine.hpp"

namespace gcfg {
const char *last_error_cstr();
}

using gcfg::Engine;
using gcfg::InstState;
using gcfg::now_ms;
using gcfg::set_error;

struct gcfg_graph {
    Engine eng;
    int W = 0, H = 0, slot = -1;
    bool built = false;
    int32_t *d_label = nullptr;  // internal label buffer (W*H)
    int32_t *label_target = nullptr;
    const int32_t *pos = nullptr, *neg = nullptr, *up = nullptr, *down = nullptr, *left = nullptr,
                  *right = nullptr;
    int32_t penalty = 0;
    const int32_t *diag[4] = {nullptr, nullptr, nullptr, nullptr};  // 8-neighbourhood planes
    int nbhd = 4;
    gcfg_options opt{};
    gcfg::StreamJoin join;
};

struct gcfg_batch {
    Engine eng;
    std::vector<gcfg_instance> inst;
    std::vector<int> slots;
    int32_t *d_labels = nullptr;  // internal label arena
    std::vector<size_t> label_off;
    // Packed labels (gcfg_batch_solve_bits): device words of every instance
    // and their pinned host copy (coherent, written by the scheduler kernel
    // as each instance finishes; h_bits_dev is its device-visible address),
    // allocated with the batch.
    uint64_t *d_bits = nullptr, *h_bits = nullptr, *h_bits_dev = nullptr;
    std::vector<size_t> bits_off;
    size_t bits_words = 0;
    std::vector<InstState> states;
    gcfg::StreamJoin join;
};

extern "C" {

void gcfg_default_options(gcfg_options *o) {
    if (!o) return;
    o->sweeps_per_gr = 1;
    o->inner_iters = 8;
    o->steps_per_poll = 8;
    o->use_graph = 1;
    o->max_steps = 200000;
    o->verbose = 0;
    o->profile = 0;
    o->kernels = 1;
    o->gr_rounds = 48;
    o->gr_block_w = 1;
    o->gr_block_h = 1;
    o->gr_cap = 4;        // the build's relabel: capped short while flow moves locally
    o->gr_cap_late = 15;  // the next three capped rounds; exact after 4 rounds (see the scheduler)
    o->pr_waves_per_cu = 0;
    o->gr_waves_per_cu = 0;
    o->lazy_caps = 0;
}

const char *gcfg_last_error(void) { return gcfg::last_error_cstr(); }

int gcfg_device_info(char *name, size_t name_len, int32_t *cu_count, size_t *hbm_bytes) {
    int dev = 0;
    GCFG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GCFG_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return GCFG_OK;
}

// ---------------------------------------------------------------------------
static int graph_create(int32_t width, int32_t height, int nbhd, gcfg_graph **out) {
    if (!out || !Engine::shape_supported(width, height)) {
        set_error("gcfg_graph_create: invalid shape (positive, at most 2^30 pixels) or output pointer");
        return GCFG_ERR_INVALID;
    }
    *out = nullptr;
    gcfg_graph *g = new (std::nothrow) gcfg_graph();
    if (!g) return GCFG_ERR_NOMEM;
    gcfg_default_options(&g->opt);
    g->W = width;
    g->H = height;
    g->nbhd = nbhd;
    int rc = g->eng.init(g->eng.tiles_for(width, height), 1, g->opt, nbhd);
    if (rc == GCFG_OK && hipMalloc((void **)&g->d_label, (size_t)width * height * sizeof(int32_t)) != hipSuccess) {
        set_error("gcfg_graph_create: label buffer allocation failed");
        rc = GCFG_ERR_NOMEM;
    }
    if (rc == GCFG_OK) {
        g->slot = g->eng.add_instance(width, height, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                      g->d_label);
        if (g->slot < 0) rc = g->slot;
    }
    if (rc != GCFG_OK) {
        if (g->d_label) (void)hipFree(g->d_label);
        delete g;
        return rc;
    }
    g->label_target = g->d_label;
    *out = g;
    return GCFG_OK;
}

int gcfg_graph_create(int32_t width, int32_t height, gcfg_graph **out) {
    return graph_create(width, height, 4, out);
}

int gcfg_graph_create_n8(int32_t width, int32_t height, gcfg_graph **out) {
    return graph_create(width, height, 8, out);
}

static int graph_push_inputs(gcfg_graph *g) {
    return g->eng.set_inputs(g->slot, g->pos, g->neg, g->up, g->down, g->left, g->right, g->penalty,
                             g->label_target, g->diag[0] ? g->diag : nullptr);
}

int gcfg_graph_set_dataterms(gcfg_graph *g, const int32_t *d_pos, const int32_t *d_neg) {
    if (!g || !d_pos || !d_neg) {
        set_error("gcfg_graph_set_dataterms: null argument");
        return GCFG_ERR_INVALID;
    }
    g->pos = d_pos;
    g->neg = d_neg;
    return graph_push_inputs(g);
}

int gcfg_graph_set_edges(gcfg_graph *g, const int32_t *d_up, const int32_t *d_down, const int32_t *d_left,
                         const int32_t *d_right) {
    if (!g || !d_up || !d_down || !d_left || !d_right) {
        set_error("gcfg_graph_set_edges: null argument");
        return GCFG_ERR_INVALID;
    }
    g->up = d_up;
    g->down = d_down;
    g->left = d_left;
    g->right = d_right;
    for (auto &d : g->diag) d = nullptr;  // an 8-neighbourhood graph then needs gcfg_graph_set_edges8
    return graph_push_inputs(g);
}

int gcfg_graph_set_edges8(gcfg_graph *g, const int32_t *d_up, const int32_t *d_down, const int32_t *d_left,
                          const int32_t *d_right, const int32_t *d_upleft, const int32_t *d_upright,
                          const int32_t *d_downleft, const int32_t *d_downright) {
    if (!g || !d_up || !d_down || !d_left || !d_right || !d_upleft || !d_upright || !d_downleft || !d_downright) {
        set_error("gcfg_graph_set_edges8: null argument");
        return GCFG_ERR_INVALID;
    }
    if (g->nbhd != 8) {
        set_error("gcfg_graph_set_edges8: graph was created with the 4-neighbourhood");
        return GCFG_ERR_STATE;
    }
    g->up = d_up;
    g->down = d_down;
    g->left = d_left;
    g->right = d_right;
    g->diag[0] = d_upleft;
    g->diag[1] = d_upright;
    g->diag[2] = d_downleft;
    g->diag[3] = d_downright;
    return graph_push_inputs(g);
}

int gcfg_graph_set_penalty(gcfg_graph *g, int32_t penalty) {
    if (!g || penalty < 0) {
        set_error("gcfg_graph_set_penalty: null handle or negative penalty");
        return GCFG_ERR_INVALID;
    }
    g->up = g->down = g->left = g->right = nullptr;
    for (auto &d : g->diag) d = nullptr;
    g->penalty = penalty;
    return graph_push_inputs(g);
}

int gcfg_graph_set_stream(gcfg_graph *g, void *hip_stream) {
    if (!g) return GCFG_ERR_INVALID;
    return g->join.set((hipStream_t)hip_stream);
}

int gcfg_batch_set_stream(gcfg_batch *b, void *hip_stream) {
    if (!b) return GCFG_ERR_INVALID;
    return b->join.set((hipStream_t)hip_stream);
}

int gcfg_graph_set_options(gcfg_graph *g, const gcfg_options *opt) {
    if (!g || !opt) {
        set_error("gcfg_graph_set_options: null argument");
        return GCFG_ERR_INVALID;
    }
    if (!gcfg::options_valid(*opt)) return GCFG_ERR_INVALID;
    g->opt = *opt;
    g->eng.opt = *opt;
    return GCFG_OK;
}

int gcfg_graph_build(gcfg_graph *g) {
    if (!g) return GCFG_ERR_INVALID;
    if (!g->pos || !g->neg) {
        set_error("gcfg_graph_build: data terms not set");
        return GCFG_ERR_STATE;
    }
    if (g->opt.verbose) std::printf("pen %d\n", g->penalty);
    int rc = graph_push_inputs(g);
    if (rc == GCFG_OK) rc = g->join.before(g->eng.stream());
    if (rc == GCFG_OK) rc = g->eng.build({g->slot});
    g->built = rc == GCFG_OK;
    return rc;
}

static int graph_solve(gcfg_graph *g, int32_t *d_target, gcfg_stats *stats) {
    const double t0 = now_ms();
    if (!g->built) {
        set_error("gcfg_graph_optimize: graph not built (call gcfg_graph_build)");
        return GCFG_ERR_STATE;
    }
    if (d_target != g->label_target) {
        // The label destination is part of the uploaded descriptor.
        g->label_target = d_target;
        int rc = graph_push_inputs(g);
        if (rc == GCFG_OK) rc = g->eng.upload_desc();
        if (rc != GCFG_OK) return rc;
    }
    g->eng.reset_profile();
    const int64_t steps0 = g->eng.steps_run;
    int rc = g->join.before(g->eng.stream());
    if (rc == GCFG_OK) rc = g->eng.run(g->opt.max_steps);
    if (rc == GCFG_OK) rc = g->join.after(g->eng.stream());
    if (rc == GCFG_OK) rc = g->eng.fetch_states_pinned(g->slot, 1);
    if (rc != GCFG_OK) return rc;
    // The counters as of the final poll (nothing has run since).
    const InstState st = g->eng.pinned_states()[0];
    g->eng.steps_hint = st.steps;
    const gcfg::Ctl ctl = g->eng.polled_ctl();
    if (st.err) {
        set_error("invalid capacities in the inputs (|pos - neg| > 2^29, or an n-link outside [0, 2^28) "
                  "(4-neighbourhood) / [0, 2^27) (8-neighbourhood))");
        return GCFG_ERR_INVALID;
    }
    if (g->opt.verbose) std::printf("Final energy level: %lld\n", (long long)st.flow);
    if (stats) {
        stats->flow = st.flow;
        stats->pr_tiles = st.pr_tiles;
        stats->gr_tiles = st.gr_tiles;
        stats->gr_init_tiles = st.gr_init_tiles;
        stats->gr_exact_tiles = st.gr_exact_tiles;
        stats->fin_items = st.fin_items;
        stats->pr_iters = ctl.pr_iters;
        stats->gr_iters = ctl.gr_iters;
        stats->steps = st.steps;
        stats->gr_rounds = st.gr_rounds;
        stats->ms = now_ms() - t0;
        stats->pr_kernel_ms = g->eng.prof_pr_ms;
        stats->gr_kernel_ms = g->eng.prof_gr_ms;
        stats->sched_kernel_ms = g->eng.prof_sched_ms;
        stats->build_kernel_ms = g->eng.prof_build_ms;
        stats->launches = g->eng.steps_run - steps0;
    }
    return GCFG_OK;
}

int gcfg_graph_optimize_device(gcfg_graph *g, int32_t *d_label, gcfg_stats *stats) {
    if (!g || !d_label) {
        set_error("gcfg_graph_optimize_device: null argument");
        return GCFG_ERR_INVALID;
    }
    return graph_solve(g, d_label, stats);
}

int gcfg_graph_optimize(gcfg_graph *g, int32_t *h_label, gcfg_stats *stats) {
    if (!g || !h_label) {
        set_error("gcfg_graph_optimize: null argument");
        return GCFG_ERR_INVALID;
    }
    const double t0 = now_ms();
    int rc = graph_solve(g, g->d_label, stats);
    if (rc != GCFG_OK) return rc;
    GCFG_HIP(hipMemcpyAsync(h_label, g->d_label, (size_t)g->W * g->H * sizeof(int32_t), hipMemcpyDeviceToHost,
                            g->eng.stream()));
    GCFG_HIP(hipStreamSynchronize(g->eng.stream()));
    if (stats) stats->ms = now_ms() - t0;
    return GCFG_OK;
}

int gcfg_graph_read_state(gcfg_graph *g, int32_t *h_caps, int32_t *h_excess) {
    if (!g || !h_caps || !h_excess) {
        set_error("gcfg_graph_read_state: null argument");
        return GCFG_ERR_INVALID;
    }
    if (!g->built) {
        set_error("gcfg_graph_read_state: graph not built");
        return GCFG_ERR_STATE;
    }
    return g->eng.read_state(g->slot, h_caps, h_excess);
}

int gcfg_graph_read_diag_state(gcfg_graph *g, int32_t *h_caps8) {
    if (!g || !h_caps8) {
        set_error("gcfg_graph_read_diag_state: null argument");
        return GCFG_ERR_INVALID;
    }
    if (!g->built) {
        set_error("gcfg_graph_read_diag_state: graph not built");
        return GCFG_ERR_STATE;
    }
    return g->eng.read_state(g->slot, nullptr, nullptr, nullptr, h_caps8);
}

int gcfg_graph_neighborhood(gcfg_graph *g, int32_t *nbhd) {
    if (!g || !nbhd) return GCFG_ERR_INVALID;
    *nbhd = g->nbhd;
    return GCFG_OK;
}

int gcfg_graph_read_distances(gcfg_graph *g, int32_t *h_dist) {
    if (!g || !h_dist) {
        set_error("gcfg_graph_read_distances: null argument");
        return GCFG_ERR_INVALID;
    }
    if (!g->built) {
        set_error("gcfg_graph_read_distances: graph not built");
        return GCFG_ERR_STATE;
    }
    return g->eng.read_state(g->slot, nullptr, nullptr, h_dist);
}

int gcfg_graph_destroy(gcfg_graph *g) {
    if (!g) return GCFG_OK;
    g->eng.sync();
    if (g->d_label) (void)hipFree(g->d_label);
    delete g;
    return GCFG_OK;
}

// ---------------------------------------------------------------------------
static int alloc_label_bits(gcfg_batch *b);

int gcfg_batch_create(int32_t count, const gcfg_instance *inst, const gcfg_options *opt, gcfg_batch **out) {
    if (!out || count <= 0 || !inst) {
        set_error("gcfg_batch_create: invalid arguments");
        return GCFG_ERR_INVALID;
    }
    *out = nullptr;
    gcfg_options o;
    if (opt)
        o = *opt;
    else
        gcfg_default_options(&o);
    gcfg_batch *b = new (std::nothrow) gcfg_batch();
    if (!b) return GCFG_ERR_NOMEM;
    b->inst.assign(inst, inst + count);
    int64_t tiles = 0;
    size_t labels = 0;
    for (int i = 0; i < count; ++i) {
        const gcfg_instance &in = inst[i];
        if (!Engine::shape_supported(in.width, in.height) || !in.d_pos || !in.d_neg) {
            set_error("gcfg_batch_create: instance " + std::to_string(i) +
                      " has no data terms or an invalid shape (positive, at most 2^30 pixels)");
            delete b;
            return GCFG_ERR_INVALID;
        }
        tiles += b->eng.tiles_for(in.width, in.height);
        if (tiles >= (int64_t)1 << 28) {
            set_error("gcfg_batch_create: more than 2^28 tiles in one batch");
            delete b;
            return GCFG_ERR_INVALID;
        }
        b->label_off.push_back(labels);
        labels += (size_t)in.width * in.height;
    }
    int rc = b->eng.init((int)tiles, count, o);
    if (rc == GCFG_OK && hipMalloc((void **)&b->d_labels, labels * sizeof(int32_t)) != hipSuccess) {
        set_error("gcfg_batch_create: label arena allocation failed");
        rc = GCFG_ERR_NOMEM;
    }
    if (rc == GCFG_OK) rc = alloc_label_bits(b);
    for (int i = 0; rc == GCFG_OK && i < count; ++i) {
        const gcfg_instance &in = inst[i];
        const int s = b->eng.add_instance(in.width, in.height, in.d_pos, in.d_neg, in.d_up, in.d_down, in.d_left,
                                          in.d_right, in.penalty, b->d_labels + b->label_off[i]);
        if (s < 0) rc = s;
        b->slots.push_back(s);
    }
    if (rc != GCFG_OK) {
        if (b->d_labels) (void)hipFree(b->d_labels);
        if (b->d_bits) (void)hipFree(b->d_bits);
        if (b->h_bits) (void)hipHostFree(b->h_bits);
        delete b;
        return rc;
    }
    b->states.resize(count);
    *out = b;
    return GCFG_OK;
}

// Cold solve of every instance. Labels go to d_targets (device int32), or
// with bits to the packed arena, from which the scheduler copies each
// instance's words to the pinned host buffer the step it finishes (so the
// copies overlap the solve of the others and none is left after it);
// otherwise to the internal int32 arena.
static int batch_run(gcfg_batch *b, int32_t **d_targets, bool bits, int64_t *flows, gcfg_stats *stats) {
    const double t0 = now_ms();
    const int n = (int)b->slots.size();
    for (int i = 0; i < n; ++i) {
        const gcfg_instance &in = b->inst[i];
        int32_t *lab = bits ? nullptr : d_targets ? d_targets[i] : b->d_labels + b->label_off[i];
        b->eng.set_inputs(b->slots[i], in.d_pos, in.d_neg, in.d_up, in.d_down, in.d_left, in.d_right, in.penalty,
                          lab);
        b->eng.set_label_bits(b->slots[i], bits ? b->d_bits + b->bits_off[i] : nullptr);
        b->eng.set_label_host(b->slots[i], bits ? b->h_bits_dev + b->bits_off[i] : nullptr);
    }
    b->eng.reset_profile();
    const int64_t steps0 = b->eng.steps_run;
    int rc = b->join.before(b->eng.stream());
    if (rc == GCFG_OK) rc = b->eng.build(b->slots);
    if (rc == GCFG_OK) rc = b->eng.run(b->eng.opt.max_steps);
    if (rc == GCFG_OK) rc = b->join.after(b->eng.stream());
    // Instance slots were assigned 0..n-1 in order at creation: one copy.
    // (Copying them with every completion poll instead cost a one-off stall
    // of several ms a few solves into a process.)
    if (rc == GCFG_OK) rc = b->eng.fetch_states_pinned(0, n);
    if (rc != GCFG_OK) return rc;
    std::memcpy(b->states.data(), b->eng.pinned_states(), (size_t)n * sizeof(InstState));
    // The counters as of the final poll (nothing has run since).
    const gcfg::Ctl ctl = b->eng.polled_ctl();
    gcfg_stats agg{};
    agg.pr_iters = ctl.pr_iters;
    agg.gr_iters = ctl.gr_iters;
    for (int i = 0; i < n; ++i) {
        const InstState &st = b->states[i];
        if (st.err) {
            set_error("invalid capacities in instance " + std::to_string(i) +
                      " (|pos - neg| > 2^29, or an n-link outside [0, 2^28))");
            return GCFG_ERR_INVALID;
        }
        if (flows) flows[i] = st.flow;
        agg.flow += st.flow;
        agg.pr_tiles += st.pr_tiles;
        agg.gr_tiles += st.gr_tiles;
        agg.gr_init_tiles += st.gr_init_tiles;
        agg.gr_exact_tiles += st.gr_exact_tiles;
        agg.fin_items += st.fin_items;
        agg.steps = st.steps > agg.steps ? st.steps : agg.steps;
        agg.gr_rounds += st.gr_rounds;
    }
    agg.ms = now_ms() - t0;
    b->eng.steps_hint = agg.steps;
    agg.pr_kernel_ms = b->eng.prof_pr_ms;
    agg.gr_kernel_ms = b->eng.prof_gr_ms;
    agg.sched_kernel_ms = b->eng.prof_sched_ms;
    agg.build_kernel_ms = b->eng.prof_build_ms;
    agg.launches = b->eng.steps_run - steps0;
    if (stats) *stats = agg;
    return GCFG_OK;
}

int gcfg_batch_solve_device(gcfg_batch *b, int32_t **d_labels, int64_t *flows, gcfg_stats *stats) {
    if (!b || !d_labels) {
        set_error("gcfg_batch_solve_device: null argument");
        return GCFG_ERR_INVALID;
    }
    return batch_run(b, d_labels, false, flows, stats);
}

int gcfg_batch_solve(gcfg_batch *b, int32_t **h_labels, int64_t *flows, gcfg_stats *stats) {
    if (!b) return GCFG_ERR_INVALID;
    const double t0 = now_ms();
    int rc = batch_run(b, nullptr, false, flows, stats);
    if (rc != GCFG_OK) return rc;
    if (h_labels) {
        for (size_t i = 0; i < b->inst.size(); ++i) {
            if (!h_labels[i]) continue;
            GCFG_HIP(hipMemcpyAsync(h_labels[i], b->d_labels + b->label_off[i],
                                    (size_t)b->inst[i].width * b->inst[i].height * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, b->eng.stream()));
        }
        GCFG_HIP(hipStreamSynchronize(b->eng.stream()));
    }
    if (stats) stats->ms = now_ms() - t0;
    return GCFG_OK;
}

// Device and pinned host buffers of a batch's packed labels (1/32 of the
// int32 label arena each), allocated with the batch so that no solve pays
// for them. Each instance starts on a 128-byte boundary, so the completion
// copies are whole 16-byte lanes and full lines. The host buffer is
// coherent: the scheduler's stores reach host memory without a cache flush.
static int alloc_label_bits(gcfg_batch *b) {
    size_t words = 0;
    b->bits_off.clear();
    for (const gcfg_instance &in : b->inst) {
        b->bits_off.push_back(words);
        words += (Engine::label_words(in.width, in.height) + 15) & ~(size_t)15;
    }
    const size_t bytes = std::max<size_t>(words, 1) * sizeof(uint64_t);
    void *dev = nullptr;
    if (hipMalloc((void **)&b->d_bits, bytes) != hipSuccess ||
        hipHostMalloc((void **)&b->h_bits, bytes, hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(&dev, b->h_bits, 0) != hipSuccess || !dev) {
        if (b->d_bits) (void)hipFree(b->d_bits);
        if (b->h_bits) (void)hipHostFree(b->h_bits);
        b->d_bits = nullptr;
        b->h_bits = nullptr;
        set_error("packed label allocation failed");
        return GCFG_ERR_NOMEM;
    }
    b->h_bits_dev = static_cast<uint64_t *>(dev);
    b->bits_words = words;
    return GCFG_OK;
}

int gcfg_batch_solve_bits(gcfg_batch *b, const uint64_t **h_bits, int64_t *flows, gcfg_stats *stats) {
    if (!b) return GCFG_ERR_INVALID;
    const double t0 = now_ms();
    if (!b->d_bits) return GCFG_ERR_STATE;
    const int rc = batch_run(b, nullptr, true, flows, stats);
    if (rc != GCFG_OK) return rc;
    if (h_bits)
        for (size_t i = 0; i < b->inst.size(); ++i) h_bits[i] = b->h_bits + b->bits_off[i];
    if (stats) stats->ms = now_ms() - t0;
    return GCFG_OK;
}

int gcfg_batch_destroy(gcfg_batch *b) {
    if (!b) return GCFG_OK;
    b->eng.sync();
    if (b->d_bits) (void)hipFree(b->d_bits);
    if (b->h_bits) (void)hipHostFree(b->h_bits);
    if (b->d_labels) (void)hipFree(b->d_labels);
    delete b;
    return GCFG_OK;
}

// ---------------------------------------------------------------------------
int gcfg_graph_autotune(gcfg_graph *g, int32_t trials, gcfg_options *best, double *best_ms) {
    if (!g) return GCFG_ERR_INVALID;
    const gcfg_options start = g->opt;
    auto solve = [&](const gcfg_options &o) -> double {
        if (gcfg_graph_set_options(g, &o) != GCFG_OK) return -1.0;
        const double t0 = now_ms();
        gcfg_stats st;
        if (gcfg_graph_build(g) != GCFG_OK || graph_solve(g, g->label_target, &st) != GCFG_OK) return -1.0;
        return now_ms() - t0;
    };
    gcfg_options win;
    int rc = gcfg::coordinate_search(start, trials, solve, &win, best_ms);
    if (rc != GCFG_OK) {
        gcfg_graph_set_options(g, &start);
        return rc;
    }
    if (best) *best = win;
    return gcfg_graph_set_options(g, &win);
}

int gcfg_batch_autotune(gcfg_batch *b, int32_t trials, gcfg_options *best, double *best_ms) {
    if (!b) return GCFG_ERR_INVALID;
    const gcfg_options start = b->eng.opt;
    auto solve = [&](const gcfg_options &o) -> double {
        b->eng.opt = o;
        const double t0 = now_ms();
        if (batch_run(b, nullptr, false, nullptr, nullptr) != GCFG_OK) return -1.0;
        return now_ms() - t0;
    };
    gcfg_options win;
    int rc = gcfg::coordinate_search(start, trials, solve, &win, best_ms);
    b->eng.opt = rc == GCFG_OK ? win : start;
    if (rc == GCFG_OK && best) *best = win;
    return rc;
}

}  // extern "C"
