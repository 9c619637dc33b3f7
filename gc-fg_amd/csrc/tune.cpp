// Autotuner: picks the solver knobs (inner push/relabel iterations per tile
// pass, colour-pass pairs between global relabels, relaxation rounds per BFS
// tile sweep, steps per host poll) for a given workload by timing cold solves,
// and keeps the choices per workload class in a small text file. The reference
// fixes these as compile-time macros (GraphCut.h:130-171).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"

using gcfg::now_ms;
using gcfg::set_error;

namespace gcfg {

// Coordinate search: each knob in turn over its candidate values, keeping the
// best so far (gains under 2 % are treated as noise); two passes at most.
// `solve` runs one cold solve with the given options and returns its wall
// time in ms, or a negative value on error.
int coordinate_search(gcfg_options start, int trials, const std::function<double(const gcfg_options &)> &solve,
                      gcfg_options *best, double *best_ms) {
    if (trials < 1) trials = 1;
    auto measure = [&](const gcfg_options &o) -> double {
        double m = 1e300;
        for (int i = 0; i < trials; ++i) {
            const double t = solve(o);
            if (t < 0) return -1.0;
            if (t < m) m = t;
        }
        return m;
    };
    gcfg_options cur = start;
    cur.profile = 0;
    double cur_ms = measure(cur);
    if (cur_ms < 0) return GCFG_ERR_HIP;
    struct Knob {
        int32_t gcfg_options::*field;
        std::vector<int32_t> values;
    };
    const std::vector<Knob> knobs = {
        {&gcfg_options::inner_iters, {1, 2, 4, 8, 12, 16, 32}},
        {&gcfg_options::sweeps_per_gr, {1, 2, 4, 8}},
        {&gcfg_options::gr_rounds, {2, 4, 8, 48}},
        {&gcfg_options::steps_per_poll, {4, 8, 16, 32}},
        {&gcfg_options::gr_cap, {0, 2, 3, 4, 6, 8}},
        {&gcfg_options::gr_cap_late, {4, 8, 12, 15, 16, 32}},
        {&gcfg_options::gr_waves_per_cu, {8, 16, 28, 56}},
        {&gcfg_options::pr_waves_per_cu, {8, 16, 32}},
    };
    // Global-relabel block shapes (width, height in tiles) are one knob.
    const std::vector<std::pair<int32_t, int32_t>> shapes = {{1, 1}, {1, 2}, {1, 4}, {1, 8}, {1, 16},
                                                             {2, 2}, {2, 4}, {2, 8}, {4, 4}};
    // The later rounds' relabel cap and the push iterations per pass pay off
    // together (a larger cap lets flow go further per round, more iterations
    // let it get there): they are tried as pairs before one at a time. On a
    // 4096^2 image with heavy-tailed t-links cap 12 alone and 12 iterations
    // alone each gained under 2 %, together 15 %.
    const std::vector<std::pair<int32_t, int32_t>> cap_iters = {{4, 2}, {8, 8}, {12, 12}, {15, 12}, {16, 16}};
    for (int pass = 0; pass < 2; ++pass) {
        bool improved = false;
        for (const auto &ci : cap_iters) {
            if (cur.gr_cap_late == ci.first && cur.inner_iters == ci.second) continue;
            gcfg_options cand = cur;
            cand.gr_cap_late = ci.first;
            cand.inner_iters = ci.second;
            const double ms = measure(cand);
            if (ms < 0) return GCFG_ERR_HIP;
            if (ms < cur_ms * 0.98) {
                cur = cand;
                cur_ms = ms;
                improved = true;
            }
        }
        for (const auto &sh : shapes) {
            if (cur.gr_block_w == sh.first && cur.gr_block_h == sh.second) continue;
            gcfg_options cand = cur;
            cand.gr_block_w = sh.first;
            cand.gr_block_h = sh.second;
            const double ms = measure(cand);
            if (ms < 0) return GCFG_ERR_HIP;
            if (ms < cur_ms * 0.98) {
                cur = cand;
                cur_ms = ms;
                improved = true;
            }
        }
        for (const Knob &k : knobs) {
            for (int32_t v : k.values) {
                if (cur.*(k.field) == v) continue;
                gcfg_options cand = cur;
                cand.*(k.field) = v;
                const double ms = measure(cand);
                if (ms < 0) return GCFG_ERR_HIP;
                if (ms < cur_ms * 0.98) {
                    cur = cand;
                    cur_ms = ms;
                    improved = true;
                }
            }
        }
        if (!improved) break;
    }
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return GCFG_OK;
}

}  // namespace gcfg

extern "C" {

static const char *kFields[] = {"sweeps_per_gr", "inner_iters", "steps_per_poll", "use_graph",
                                "max_steps",     "verbose",     "profile",        "kernels",
                                "gr_rounds",     "gr_block_w",  "gr_block_h",     "gr_cap",
                                "pr_waves_per_cu", "gr_waves_per_cu", "lazy_caps", "gr_cap_late"};

static std::vector<int32_t *> option_fields(gcfg_options *o) {
    return {&o->sweeps_per_gr, &o->inner_iters, &o->steps_per_poll, &o->use_graph,
            &o->max_steps,     &o->verbose,     &o->profile,        &o->kernels,
            &o->gr_rounds,     &o->gr_block_w,  &o->gr_block_h,     &o->gr_cap,
            &o->pr_waves_per_cu, &o->gr_waves_per_cu, &o->lazy_caps, &o->gr_cap_late};
}

int gcfg_tune_store(const char *path, const char *key, const gcfg_options *opt) {
    if (!path || !key || !opt || std::strpbrk(key, " \t\n")) {
        set_error("gcfg_tune_store: invalid arguments (the key must not contain whitespace)");
        return GCFG_ERR_INVALID;
    }
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            ls >> k;
            if (!k.empty() && k != key) keep.push_back(line);
        }
    }
    gcfg_options o = *opt;
    std::ostringstream os;
    os << key;
    auto f = option_fields(&o);
    for (size_t i = 0; i < f.size(); ++i) os << ' ' << kFields[i] << '=' << *f[i];
    keep.push_back(os.str());
    std::ofstream out(path, std::ios::trunc);
    if (!out) {
        set_error(std::string("gcfg_tune_store: cannot write ") + path);
        return GCFG_ERR_INVALID;
    }
    for (auto &l : keep) out << l << '\n';
    return GCFG_OK;
}

int gcfg_tune_lookup(const char *path, const char *key, gcfg_options *opt) {
    if (!path || !key || !opt) return GCFG_ERR_INVALID;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string k;
        ls >> k;
        if (k != key) continue;
        gcfg_options o;
        gcfg_default_options(&o);
        auto f = option_fields(&o);
        std::string tok;
        while (ls >> tok) {
            const size_t eq = tok.find('=');
            if (eq == std::string::npos) continue;
            const std::string name = tok.substr(0, eq);
            for (size_t i = 0; i < f.size(); ++i)
                if (name == kFields[i]) *f[i] = (int32_t)std::strtol(tok.c_str() + eq + 1, nullptr, 10);
        }
        *opt = o;
        return GCFG_OK;
    }
    set_error(std::string("gcfg_tune_lookup: no entry for ") + key);
    return GCFG_ERR_STATE;
}

}  // extern "C"
