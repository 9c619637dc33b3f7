// Host side of the solver engine: arena allocation, instance table, step
// loop. Replaces the host loop of GraphCut.cu:385-534 (which synchronised
// with the device every 10 iterations and after every GlobalRelabel and
// SpreadLabels launch) by chunks of device-scheduled steps and one pinned
// poll per chunk.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>

#include "engine.hpp"

namespace gcfg {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error_cstr() { return g_last_error.c_str(); }

int hip_fail(hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return GCFG_ERR_HIP;
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    for (hipEvent_t e : build_ev_)
        if (e) (void)hipEventDestroy(e);
    for (void *p : allocs_) (void)hipFree(p);
    if (ring_) (void)hipHostFree(ring_);
    if (pin_ctl_) (void)hipHostFree(pin_ctl_);
    if (pin_states_) (void)hipHostFree(pin_states_);
    if (pin_qtrace_) (void)hipHostFree(pin_qtrace_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int StreamJoin::set(hipStream_t s) {
    if (s && !ev_in) {
        GCFG_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
        GCFG_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    }
    user = s;
    return GCFG_OK;
}

int StreamJoin::before(hipStream_t own) {
    if (!user) return GCFG_OK;
    GCFG_HIP(hipEventRecord(ev_in, user));
    GCFG_HIP(hipStreamWaitEvent(own, ev_in, 0));
    return GCFG_OK;
}

int StreamJoin::after(hipStream_t own) {
    if (!user) return GCFG_OK;
    GCFG_HIP(hipEventRecord(ev_out, own));
    GCFG_HIP(hipStreamWaitEvent(user, ev_out, 0));
    return GCFG_OK;
}

StreamJoin::~StreamJoin() {
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
}

int Engine::init(int cap_tiles, int cap_slots, const gcfg_options &o, int nbhd) {
    opt = o;
    if (!options_valid(opt)) return GCFG_ERR_INVALID;
    if (nbhd != 4 && nbhd != 8) {
        set_error("neighbourhood must be 4 or 8");
        return GCFG_ERR_INVALID;
    }
    nbhd_ = nbhd;
    kp_.nbhd = nbhd;
    cap_tiles = std::max(cap_tiles, 1);
    cap_slots = std::max(cap_slots, 1);
    // Relabel queue items hold a tile id shifted by 3 bits (an op in the low
    // bits), push items a flag in bit 30: tile ids stay below 2^28 (275 G
    // nodes, far beyond one GPU's memory).
    if (cap_tiles >= (1 << 28)) {
        set_error("arena of more than 2^28 tiles");
        return GCFG_ERR_INVALID;
    }
    cap_tiles_ = cap_tiles;
    cap_slots_ = cap_slots;
    GCFG_HIP(hipGetDevice(&dev_));
    GCFG_HIP(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, dev_));
    GCFG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    auto dalloc = [&](void **p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 16);
        if (e != hipSuccess) {
            *p = nullptr;
            set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
            return GCFG_ERR_NOMEM;
        }
        allocs_.push_back(*p);
        return GCFG_OK;
    };
    const size_t nodes = (size_t)cap_tiles * TN;
    int rc = GCFG_OK;
    rc |= dalloc((void **)&kp_.caps, nodes * sizeof(int4));
    rc |= dalloc((void **)&kp_.exc, nodes * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.dist, nodes * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.mask, nodes);
    rc |= dalloc((void **)&kp_.tiles, (size_t)cap_tiles * sizeof(TileInfo));
    rc |= dalloc((void **)&kp_.act_flag, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.dirty, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.live, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.stale, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.tsum, (size_t)cap_tiles * 2 * sizeof(int64_t));
    rc |= dalloc((void **)&kp_.queue_pr, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.queue_gr, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.bmap, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.mat, (size_t)cap_tiles * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.inflow, (size_t)cap_tiles * INFLOW * sizeof(int32_t));
    rc |= dalloc((void **)&kp_.desc, (size_t)cap_slots * sizeof(InstDesc));
    rc |= dalloc((void **)&kp_.state, (size_t)cap_slots * sizeof(InstState));
    rc |= dalloc((void **)&kp_.ctl, sizeof(Ctl));
    rc |= dalloc((void **)&d_slots_, (size_t)(2 * cap_slots + 1) * sizeof(int32_t));
    if (nbhd == 8) {
        rc |= dalloc((void **)&kp_.caps8, nodes * sizeof(int4));
        rc |= dalloc((void **)&kp_.mask8, nodes);
        rc |= dalloc((void **)&kp_.nbr8, (size_t)cap_tiles * sizeof(int4));
    }
    if (rc != GCFG_OK) return GCFG_ERR_NOMEM;
    ring_cap_ = std::max<size_t>((size_t)4 << 20, 2 * ((size_t)cap_slots * sizeof(InstDesc) + 4096));
    GCFG_HIP(hipHostMalloc((void **)&ring_, ring_cap_, hipHostMallocDefault));
    ring_off_ = 0;
    GCFG_HIP(hipHostMalloc((void **)&pin_ctl_, sizeof(Ctl), hipHostMallocDefault));
    GCFG_HIP(hipHostMalloc((void **)&pin_states_, (size_t)cap_slots * sizeof(InstState), hipHostMallocDefault));
    GCFG_HIP(hipMemsetAsync(kp_.state, 0, (size_t)cap_slots * sizeof(InstState), stream_));
    GCFG_HIP(hipMemsetAsync(kp_.desc, 0, (size_t)cap_slots * sizeof(InstDesc), stream_));
    dev_desc_.assign(cap_slots, InstDesc{});
    GCFG_HIP(hipMemsetAsync(kp_.ctl, 0, sizeof(Ctl), stream_));
    GCFG_HIP(hipMemsetAsync(kp_.dirty, 0, (size_t)cap_tiles * sizeof(int32_t), stream_));
    GCFG_HIP(hipMemsetAsync(kp_.act_flag, 0, (size_t)cap_tiles * sizeof(int32_t), stream_));
    GCFG_HIP(hipMemsetAsync(kp_.live, 0, (size_t)cap_tiles * sizeof(int32_t), stream_));
    GCFG_HIP(hipMemsetAsync(kp_.stale, 0, (size_t)cap_tiles * sizeof(int32_t), stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    kp_.n_slots = cap_slots;
    apply_options();
    h_desc_.assign(cap_slots, InstDesc{});
    slot_used_.assign(cap_slots, 0);
    free_ranges_.clear();
    free_ranges_.push_back({0, cap_tiles});
    std::memset(pin_ctl_, 0, sizeof(Ctl));
    return GCFG_OK;
}

bool options_valid(const gcfg_options &o) {
    if (o.sweeps_per_gr <= 0 || o.inner_iters <= 0 || o.steps_per_poll <= 0) {
        set_error("options: sweeps_per_gr, inner_iters and steps_per_poll must be positive");
        return false;
    }
    if (o.pr_waves_per_cu < 0 || o.gr_waves_per_cu < 0 || o.pr_waves_per_cu > 1024 || o.gr_waves_per_cu > 1024) {
        set_error("options: pr_waves_per_cu and gr_waves_per_cu must be in [0, 1024]");
        return false;
    }
    if (o.lazy_caps < 0 || o.lazy_caps > 2) {
        set_error("options: lazy_caps must be 0, 1 or 2");
        return false;
    }
    if (o.gr_cap < 0 || o.gr_cap_late < 0) {
        set_error("options: gr_cap and gr_cap_late must be >= 0");
        return false;
    }
    if ((o.gr_block_w || o.gr_block_h) && !gr_block_supported(o.gr_block_w, o.gr_block_h)) {
        set_error("options: unsupported gr_block_w x gr_block_h (1x1, 1x2, 1x4, 1x8, 1x16, 2x2, 2x4, 2x8, 4x4)");
        return false;
    }
    return true;
}

// Kernel parameters and grid sizes follow the current options; a captured
// step graph with stale parameters is dropped.
void Engine::apply_options() {
    const int cu = std::max(cu_count_, 1);
    const int variant = opt.kernels ? 1 : 0;
    int gbx = opt.gr_block_w > 0 ? opt.gr_block_w : 1, gby = opt.gr_block_h > 0 ? opt.gr_block_h : 1;
    // 8-neighbourhood relabels run one-tile blocks (a block's inner exchange
    // would need its corner nodes too).
    if (!variant || nbhd_ == 8 || !gr_block_supported(gbx, gby)) gbx = gby = 1;
    // Push launches default to the wavefronts the kernel keeps resident per
    // CU (its register use sets it: two to five per SIMD); more would wait for
    // a slot with their contiguous run of the queue. Relabel launches default
    // to 48 per CU, well beyond residency: their items are short and uneven,
    // queued workgroups take the place of finished ones, and a queue longer
    // than the grid would give workgroups runs of two or more items in a row
    // (a 4096^2 image's relabel sweeps have up to 16 k items).
    KParams pk = kp_;
    pk.variant = variant;
    pk.inner_iters = opt.inner_iters;
    const int wpr = opt.pr_waves_per_cu > 0 ? opt.pr_waves_per_cu : (variant ? pr_resident_waves_per_cu(pk) : 16);
    const int wgr = opt.gr_waves_per_cu > 0 ? opt.gr_waves_per_cu : 48;
    const int gpr = variant ? cu * wpr : cu * 4;
    const int ggr = variant ? std::max(cu, cu * wgr / (gbx * gby)) : cu * 8;
    const int grr = opt.gr_rounds > 0 ? opt.gr_rounds : 48;
    // Iteration counters add one same-address atomic per workgroup and launch,
    // which slows the kernels being timed, so only profile >= 2 turns them on.
    const int prof = opt.profile >= 2 ? 1 : 0;
    const int cap = variant && opt.gr_cap > 0 ? opt.gr_cap : 0;
    // Later capped rounds never use a smaller cap than the first: a
    // reachability round reads distances below the last cap as witnesses,
    // and a block skipped since an earlier round holds distances below that
    // round's cap.
    const int cap_late = cap ? std::max(cap, opt.gr_cap_late) : 0;
    if (kp_.inner_iters != opt.inner_iters || kp_.variant != variant || grid_pr_ != gpr || grid_gr_ != ggr ||
        kp_.gr_rounds != grr || kp_.profile != prof || kp_.gbx != gbx || kp_.gby != gby || kp_.gr_cap != cap ||
        kp_.gr_cap_late != cap_late) {
        kp_.gr_cap = cap;
        kp_.gr_cap_late = cap_late;
        kp_.inner_iters = opt.inner_iters;
        kp_.gr_rounds = grr;
        kp_.profile = prof;
        kp_.variant = variant;
        kp_.gbx = gbx;
        kp_.gby = gby;
        grid_pr_ = gpr;
        grid_gr_ = ggr;
        drop_graphs();
    }
}

size_t Engine::tiles_free() const {
    size_t f = 0;
    for (auto &r : free_ranges_) f += r.second;
    return f;
}

int Engine::alloc_tiles(int n) {
    for (size_t i = 0; i < free_ranges_.size(); ++i) {
        if (free_ranges_[i].second >= n) {
            const int start = free_ranges_[i].first;
            free_ranges_[i].first += n;
            free_ranges_[i].second -= n;
            if (free_ranges_[i].second == 0) free_ranges_.erase(free_ranges_.begin() + i);
            return start;
        }
    }
    return -1;
}

void Engine::free_tiles(int start, int n) {
    free_ranges_.push_back({start, n});
    std::sort(free_ranges_.begin(), free_ranges_.end());
    std::vector<std::pair<int, int>> merged;
    for (auto &r : free_ranges_) {
        if (!merged.empty() && merged.back().first + merged.back().second == r.first)
            merged.back().second += r.second;
        else
            merged.push_back(r);
    }
    free_ranges_.swap(merged);
}

int Engine::add_instance(int W, int H, const int32_t *pos, const int32_t *neg, const int32_t *up,
                         const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                         int32_t *label_out) {
    if (!shape_supported(W, H)) {
        set_error("instance shape must be positive and at most 2^30 pixels");
        return GCFG_ERR_INVALID;
    }
    int slot = -1;
    for (int i = 0; i < cap_slots_; ++i)
        if (!slot_used_[i]) {
            slot = i;
            break;
        }
    if (slot < 0) {
        set_error("no free instance slot");
        return GCFG_ERR_BUDGET;
    }
    const int tx = (W + TW - 1) / TW, ty = (H + TH - 1) / TH;
    const int start = alloc_tiles(tx * ty);
    if (start < 0) {
        set_error("arena has no free tile range of the requested size");
        return GCFG_ERR_BUDGET;
    }
    slot_used_[slot] = 1;
    InstDesc &d = h_desc_[slot];
    d = InstDesc{};
    d.W = W;
    d.H = H;
    d.tiles_x = tx;
    d.tiles_y = ty;
    d.tile0 = start;
    d.ntiles = tx * ty;
    d.n_interior = (W > 2 && H > 2) ? (W - 2) * (H - 2) : 0;
    set_inputs(slot, pos, neg, up, down, left, right, penalty, label_out);
    return slot;
}

int Engine::set_inputs(int slot, const int32_t *pos, const int32_t *neg, const int32_t *up, const int32_t *down,
                       const int32_t *left, const int32_t *right, int32_t penalty, int32_t *label_out,
                       const int32_t *const *diag) {
    InstDesc &d = h_desc_[slot];
    const bool d8 = nbhd_ == 8 && up && diag;
    d.upleft = d8 ? diag[0] : nullptr;
    d.upright = d8 ? diag[1] : nullptr;
    d.downleft = d8 ? diag[2] : nullptr;
    d.downright = d8 ? diag[3] : nullptr;
    d.pos = pos;
    d.neg = neg;
    d.up = up;
    d.down = up ? down : nullptr;
    d.left = up ? left : nullptr;
    d.right = up ? right : nullptr;
    d.penalty = penalty;
    d.label_out = label_out;
    return GCFG_OK;
}

void Engine::remove_instance(int slot) {
    if (slot < 0 || slot >= cap_slots_ || !slot_used_[slot]) return;
    free_tiles(h_desc_[slot].tile0, h_desc_[slot].ntiles);
    slot_used_[slot] = 0;
    h_desc_[slot] = InstDesc{};
    update_sched_width();
}

int Engine::build(const std::vector<int> &slots) {
    if (slots.empty()) return GCFG_OK;
    for (int s : slots) {
        const InstDesc &d = h_desc_[s];
        if (!d.pos || !d.neg) {
            set_error("data terms not set");
            return GCFG_ERR_STATE;
        }
        if (nbhd_ == 8 && d.up && !(d.upleft && d.upright && d.downleft && d.downright)) {
            set_error("8-neighbourhood: per-arc capacities need all eight direction planes");
            return GCFG_ERR_STATE;
        }
        if (!d.label_out && !d.label_bits) {
            set_error("label buffer not set");
            return GCFG_ERR_STATE;
        }
    }
    const int n = (int)slots.size();
    int total = 0;
    for (int i = 0; i < n; ++i) total += h_desc_[slots[i]].ntiles;
    apply_options();
    update_sched_width();
    // Lazy capacities pay off for large builds solved in few steps: the build
    // skips 16 B/node of capacity writes (about 3.7 ns per tile at its
    // bandwidth) while the tiles of the push passes read their capacities
    // from the caller's planes instead of the arena. Taken as: at least 1024
    // tiles per step the previous solve needed, or without one 32 k tiles
    // (32 M nodes). Large batches qualify; single images and long solves do
    // not. The descriptors carry the choice to the kernels.
    const int64_t need = steps_hint > 0 ? (int64_t)steps_hint * 1024 : 32768;
    // (8-neighbourhood: never; diagonal reverse arcs go to the capacities.)
    const int lazy =
        kp_.variant && nbhd_ == 4 && (opt.lazy_caps == 2 || (opt.lazy_caps == 0 && (int64_t)total >= need)) ? 1 : 0;
    for (int s : slots) h_desc_[s].lazy = lazy;
    // Build list: slots, then each one's tile offset. The descriptor table and
    // the list are uploaded only when they differ from what the device holds
    // (a repeated solve of the same batch uploads nothing).
    std::vector<int32_t> list(2 * n);
    for (int i = 0, acc = 0; i < n; ++i) {
        list[i] = slots[i];
        list[n + i] = acc;
        acc += h_desc_[slots[i]].ntiles;
    }
    {
        const int rc = upload_desc_slots(slots);
        if (rc != GCFG_OK) return rc;
    }
    if (dev_list_ != list) {
        const int rc = stage(d_slots_, list.data(), (size_t)(2 * n) * sizeof(int32_t));
        if (rc != GCFG_OK) return rc;
        dev_list_ = list;
    }
    launch_arm(kp_, d_slots_, n, 0, opt.sweeps_per_gr, stream_, kp_.variant ? total : 0);
    LaunchEvents ev;
    if (opt.profile) {
        for (auto &e : build_ev_)
            if (!e) GCFG_HIP(hipEventCreate(&e));
        ev = LaunchEvents{build_ev_[0], build_ev_[1]};
        build_timed_ = true;
    }
    launch_build(kp_, d_slots_, n, total, stream_, ev);
    GCFG_HIP(hipGetLastError());
    return GCFG_OK;
}

void Engine::update_sched_width() {
    int wide = 0;
    for (int i = 0; i < cap_slots_; ++i) wide |= slot_used_[i] && h_desc_[i].ntiles > 2048;
    kp_.sched_wide = wide;  // captured chunks are kept per width (StepGraph::wide)
}

bool Engine::any_lazy() const {
    for (int i = 0; i < cap_slots_; ++i)
        if (slot_used_[i] && h_desc_[i].lazy) return true;
    return false;
}

int Engine::materialize(const std::vector<int> &slots) {
    bool any = false;
    for (int s : slots) {
        if (s < 0 || s >= cap_slots_ || !slot_used_[s] || !h_desc_[s].lazy) continue;
        launch_materialize(kp_, s, h_desc_[s].ntiles, stream_);
        h_desc_[s].lazy = 0;  // every capacity written, no inflow pending
        any = true;
    }
    GCFG_HIP(hipGetLastError());
    if (!any) return GCFG_OK;
    return upload_desc();  // ordered after the materialising kernels on the stream
}

int Engine::stage(void *d_dst, const void *h_src, size_t bytes) {
    if (bytes == 0) return GCFG_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > ring_cap_) {  // larger than the ring: a synchronous copy
        GCFG_HIP(hipStreamSynchronize(stream_));
        GCFG_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
        return GCFG_OK;
    }
    if (ring_off_ + need > ring_cap_) {
        GCFG_HIP(hipStreamSynchronize(stream_));
        ring_off_ = 0;
    }
    std::memcpy(ring_ + ring_off_, h_src, bytes);
    GCFG_HIP(hipMemcpyAsync(d_dst, ring_ + ring_off_, bytes, hipMemcpyHostToDevice, stream_));
    ring_off_ += need;
    return GCFG_OK;
}

int Engine::upload_desc_slot(int slot) {
    if (std::memcmp(&dev_desc_[slot], &h_desc_[slot], sizeof(InstDesc)) == 0) return GCFG_OK;
    const int rc = stage(kp_.desc + slot, &h_desc_[slot], sizeof(InstDesc));
    if (rc == GCFG_OK) dev_desc_[slot] = h_desc_[slot];
    return rc;
}

// Each staged copy is a host call and a small copy on the stream, so a new
// batch of 256 instances uploaded one descriptor at a time paid 256 of them
// before its first step; consecutive changed slots go as one block.
int Engine::upload_desc_slots(const std::vector<int> &slots) {
    std::vector<int> ch;
    for (int s : slots)
        if (s >= 0 && s < cap_slots_ && std::memcmp(&dev_desc_[s], &h_desc_[s], sizeof(InstDesc)) != 0)
            ch.push_back(s);
    std::sort(ch.begin(), ch.end());
    ch.erase(std::unique(ch.begin(), ch.end()), ch.end());
    for (size_t i = 0; i < ch.size();) {
        size_t j = i + 1;
        while (j < ch.size() && ch[j] == ch[j - 1] + 1) ++j;
        const int a = ch[i], n = ch[j - 1] - a + 1;
        const int rc = stage(kp_.desc + a, &h_desc_[a], (size_t)n * sizeof(InstDesc));
        if (rc != GCFG_OK) return rc;
        for (int k = a; k < a + n; ++k) dev_desc_[k] = h_desc_[k];
        i = j;
    }
    return GCFG_OK;
}

// Descriptors that differ from the device's, queued after the work already
// on the stream.
int Engine::upload_desc() {
    std::vector<int> all(cap_slots_);
    for (int s = 0; s < cap_slots_; ++s) all[s] = s;
    return upload_desc_slots(all);
}

int Engine::snapshot_states(InstState *h_dst, int n, hipEvent_t ev) {
    if (n < 0 || n > cap_slots_) {
        set_error("snapshot_states: range outside the instance slots");
        return GCFG_ERR_INVALID;
    }
    if (n > 0)
        GCFG_HIP(hipMemcpyAsync(h_dst, kp_.state, (size_t)n * sizeof(InstState), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipEventRecord(ev, stream_));
    return GCFG_OK;
}

int Engine::rearm(const std::vector<int> &slots) {
    if (slots.empty()) return GCFG_OK;
    const int n = (int)slots.size();
    const int rc = stage(d_slots_, slots.data(), (size_t)n * sizeof(int32_t));
    if (rc != GCFG_OK) return rc;
    dev_list_.clear();  // the build list on the device is overwritten
    apply_options();
    launch_arm(kp_, d_slots_, n, 1, opt.sweeps_per_gr, stream_, 0);
    GCFG_HIP(hipGetLastError());
    return GCFG_OK;
}

int Engine::edit(int slot, const std::vector<int32_t> &strokes, int32_t delta) {
    const int n = (int)strokes.size() / 5;
    if (n <= 0) return GCFG_OK;
    if ((int)strokes.size() > strokes_cap_) {
        GCFG_HIP(hipStreamSynchronize(stream_));
        if (d_strokes_) {
            (void)hipFree(d_strokes_);
            allocs_.erase(std::remove(allocs_.begin(), allocs_.end(), (void *)d_strokes_), allocs_.end());
        }
        strokes_cap_ = std::max<int>((int)strokes.size(), 5 * 64);
        GCFG_HIP(hipMalloc((void **)&d_strokes_, (size_t)strokes_cap_ * sizeof(int32_t)));
        allocs_.push_back(d_strokes_);
    }
    const int rc = stage(d_strokes_, strokes.data(), strokes.size() * sizeof(int32_t));
    if (rc != GCFG_OK) return rc;
    launch_edit(kp_, slot, d_strokes_, n, delta, stream_);
    GCFG_HIP(hipGetLastError());
    return GCFG_OK;
}

// Steps alternate queue-counter sets by parity, so a captured chunk holds an
// even number of steps starting at parity 0 and always ends where it began.

void Engine::drop_graphs() {
    for (StepGraph &g : graphs_) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    graphs_.clear();
}

hipGraphExec_t Engine::cached_graph(int steps, int parity) const {
    for (const StepGraph &g : graphs_)
        if (g.steps == steps && g.parity == parity && g.wide == kp_.sched_wide) return g.exec;
    return nullptr;
}

int Engine::step_graph(int steps, int parity, hipGraphExec_t *out) {
    if ((*out = cached_graph(steps, parity)) != nullptr) return GCFG_OK;
    if (graphs_.size() >= 16) {  // keep a handful of chunk lengths, parities and widths
        (void)hipGraphExecDestroy(graphs_.front().exec);
        (void)hipGraphDestroy(graphs_.front().graph);
        graphs_.erase(graphs_.begin());
    }
    StepGraph g{steps, parity, kp_.sched_wide, nullptr, nullptr};
    GCFG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i) launch_step(kp_, grid_pr_, grid_gr_, (parity + i) & 1, stream_);
    GCFG_HIP(hipStreamEndCapture(stream_, &g.graph));
    hipError_t e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g.graph);
        return hip_fail(e, "hipGraphInstantiate");
    }
    graphs_.push_back(g);
    *out = g.exec;
    return GCFG_OK;
}

int Engine::launch_steps(int n) {
    apply_options();
    if (opt.profile) {
        if ((int)events_.size() < 6 * n) {
            for (int i = (int)events_.size(); i < 6 * n; ++i) {
                hipEvent_t e;
                GCFG_HIP(hipEventCreate(&e));
                events_.push_back(e);
            }
        }
        const bool trace = opt.verbose >= 2;
        if (trace && qtrace_cap_ < n) {
            if (pin_qtrace_) (void)hipHostFree(pin_qtrace_);
            pin_qtrace_ = nullptr;
            GCFG_HIP(hipHostMalloc((void **)&pin_qtrace_, (size_t)n * sizeof(QCtl), hipHostMallocDefault));
            qtrace_cap_ = n;
        }
        for (int i = 0; i < n; ++i) {
            hipEvent_t *ev = &events_[6 * i];
            launch_sched(kp_, parity_, stream_, LaunchEvents{ev[0], ev[1]});
            launch_pr(kp_, grid_pr_, parity_, stream_, LaunchEvents{ev[2], ev[3]});
            launch_gr(kp_, grid_gr_, parity_, stream_, LaunchEvents{ev[4], ev[5]});
            // This step's queue lengths, before the next scheduler clears them.
            if (trace)
                GCFG_HIP(hipMemcpyAsync(pin_qtrace_ + i, &kp_.ctl->q[parity_], sizeof(QCtl), hipMemcpyDeviceToHost,
                                        stream_));
            parity_ ^= 1;
        }
        events_pending_ = n;
        GCFG_HIP(hipGetLastError());
        steps_run += n;
        return GCFG_OK;
    }
    if (opt.use_graph) {
        const int steps = std::max(n, 1);
        hipGraphExec_t exec = cached_graph(steps, parity_);
        if (exec) {
            GCFG_HIP(hipGraphLaunch(exec, stream_));
        } else {
            // A chunk length not captured yet: launch it directly, then
            // capture, instantiate and upload its graph while the GPU runs
            // these steps (instantiating and the first launch of a new graph
            // cost about 0.2 ms each on the host, the whole of a short solve).
            for (int i = 0; i < steps; ++i) launch_step(kp_, grid_pr_, grid_gr_, (parity_ + i) & 1, stream_);
            GCFG_HIP(hipGetLastError());
            int rc = step_graph(steps, parity_, &exec);
            if (rc != GCFG_OK) return rc;
            GCFG_HIP(hipGraphUpload(exec, stream_));
        }
        parity_ ^= steps & 1;
        steps_run += steps;
    } else {
        for (int i = 0; i < n; ++i) {
            launch_step(kp_, grid_pr_, grid_gr_, parity_, stream_);
            parity_ ^= 1;
        }
        steps_run += n;
    }
    GCFG_HIP(hipGetLastError());
    return GCFG_OK;
}

int Engine::run(int64_t max_steps) {
    const int64_t start = steps_run;
    apply_options();
    if (!kp_.variant && any_lazy()) {
        // The workgroup kernels read capacities directly: write out every
        // lazy instance's (the options switched kernels after a wavefront build).
        std::vector<int> all;
        for (int i = 0; i < cap_slots_; ++i)
            if (slot_used_[i]) all.push_back(i);
        int rc = materialize(all);
        if (rc != GCFG_OK) return rc;
    }
    int chunk = steps_hint > 0 ? std::min(steps_hint, 4096) : opt.steps_per_poll;
    for (;;) {
        int rc = launch_steps(chunk);
        chunk = opt.steps_per_poll;
        if (rc != GCFG_OK) return rc;
        GCFG_HIP(hipMemcpyAsync(pin_ctl_, kp_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
        GCFG_HIP(hipStreamSynchronize(stream_));
        if (events_pending_) {
            rc = collect_profile();
            if (rc != GCFG_OK) return rc;
        }
        if (pin_ctl_->n_live <= 0) return GCFG_OK;
        if (max_steps > 0 && steps_run - start > max_steps) {
            set_error("solver did not finish within max_steps");
            return GCFG_ERR_NOT_CONVERGED;
        }
    }
}

int Engine::collect_profile() {
    if (build_timed_) {
        float b = 0;
        GCFG_HIP(hipEventElapsedTime(&b, build_ev_[0], build_ev_[1]));
        prof_build_ms += b;
        build_timed_ = false;
    }
    for (int i = 0; i < events_pending_; ++i) {
        float a = 0, b = 0, c = 0;
        GCFG_HIP(hipEventElapsedTime(&a, events_[6 * i], events_[6 * i + 1]));
        GCFG_HIP(hipEventElapsedTime(&b, events_[6 * i + 2], events_[6 * i + 3]));
        GCFG_HIP(hipEventElapsedTime(&c, events_[6 * i + 4], events_[6 * i + 5]));
        prof_sched_ms += a;
        prof_pr_ms += b;
        prof_gr_ms += c;
        if (opt.verbose >= 2 && pin_qtrace_ && i < qtrace_cap_)
            std::printf("step %lld: pr %d items %.1f us, gr %d items %.1f us, sched %.1f us\n",
                        (long long)prof_steps, pin_qtrace_[i].qlen_pr, b * 1e3, pin_qtrace_[i].qlen_gr, c * 1e3,
                        a * 1e3);
        prof_steps++;
    }
    events_pending_ = 0;
    return GCFG_OK;
}

int Engine::pump(int steps) {
    if (events_pending_) {
        GCFG_HIP(hipStreamSynchronize(stream_));
        int rc = collect_profile();
        if (rc != GCFG_OK) return rc;
    }
    return launch_steps(steps);
}

int Engine::live_count(int32_t *out) {
    GCFG_HIP(hipMemcpyAsync(pin_ctl_, kp_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    *out = pin_ctl_->n_live;
    return GCFG_OK;
}

int Engine::fetch_state(int slot, InstState *out) {
    GCFG_HIP(hipMemcpyAsync(out, kp_.state + slot, sizeof(InstState), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    return GCFG_OK;
}

int Engine::fetch_states(InstState *out, int n) {
    GCFG_HIP(hipMemcpyAsync(out, kp_.state, (size_t)n * sizeof(InstState), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    return GCFG_OK;
}

int Engine::fetch_states_pinned(int first, int n) {
    if (first < 0 || n < 0 || first + n > cap_slots_) {
        set_error("fetch_states_pinned: range outside the instance slots");
        return GCFG_ERR_INVALID;
    }
    GCFG_HIP(hipMemcpyAsync(pin_states_, kp_.state + first, (size_t)n * sizeof(InstState), hipMemcpyDeviceToHost,
                            stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    return GCFG_OK;
}

int Engine::fetch_ctl(Ctl *out) {
    GCFG_HIP(hipMemcpyAsync(pin_ctl_, kp_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
    GCFG_HIP(hipStreamSynchronize(stream_));
    *out = *pin_ctl_;
    return GCFG_OK;
}

int Engine::reframe(int slot, const int32_t *pos, const int32_t *neg, const int32_t *up, const int32_t *down,
                    const int32_t *left, const int32_t *right, int32_t penalty) {
    if (slot < 0 || slot >= cap_slots_ || !slot_used_[slot]) {
        set_error("reframe: unknown instance");
        return GCFG_ERR_INVALID;
    }
    if (nbhd_ != 4) {
        set_error("reframe: 4-neighbourhood instances only");
        return GCFG_ERR_STATE;
    }
    InstDesc &d = h_desc_[slot];
    if (!pos || !neg || !d.pos || !d.neg) {
        set_error("reframe: data terms not set");
        return GCFG_ERR_STATE;
    }
    if (d.W < 3 || d.H < 3) return set_inputs(slot, pos, neg, up, down, left, right, penalty, d.label_out);
    const int32_t *in1[6] = {pos, neg, up, up ? down : nullptr, up ? left : nullptr, up ? right : nullptr};
    int rc0 = materialize({slot});  // the warm start reads and writes residuals directly
    if (rc0 != GCFG_OK) return rc0;
    launch_reframe(kp_, slot, d.ntiles, in1, penalty, stream_);
    GCFG_HIP(hipGetLastError());
    int rc = set_inputs(slot, pos, neg, up, down, left, right, penalty, d.label_out);
    if (rc == GCFG_OK) rc = upload_desc();  // waits for the reframe kernels, which read the old inputs
    return rc;
}

int Engine::read_state(int slot, int32_t *h_caps, int32_t *h_exc, int32_t *h_dist, int32_t *h_caps8) {
    const InstDesc &d = h_desc_[slot];
    const size_t n = (size_t)d.W * d.H;
    if (h_caps8 && nbhd_ != 8) {
        set_error("diagonal capacities exist only in an 8-neighbourhood graph");
        return GCFG_ERR_STATE;
    }
    int32_t *dc = nullptr, *de = nullptr, *dd = nullptr, *d8 = nullptr;
    hipError_t e0 = hipSuccess;
    if (h_caps && e0 == hipSuccess) e0 = hipMalloc((void **)&dc, n * 4 * sizeof(int32_t));
    if (h_exc && e0 == hipSuccess) e0 = hipMalloc((void **)&de, n * sizeof(int32_t));
    if (h_dist && e0 == hipSuccess) e0 = hipMalloc((void **)&dd, n * sizeof(int32_t));
    if (h_caps8 && e0 == hipSuccess) e0 = hipMalloc((void **)&d8, n * 4 * sizeof(int32_t));
    hipError_t e1 = hipSuccess, e2 = hipSuccess, e3 = hipSuccess, e4 = hipSuccess, e5 = hipSuccess;
    if (e0 == hipSuccess) {
        launch_read_state(kp_, slot, dc, de, dd, d8, stream_);
        if (dc) e1 = hipMemcpyAsync(h_caps, dc, n * 4 * sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
        if (de) e2 = hipMemcpyAsync(h_exc, de, n * sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
        if (dd) e3 = hipMemcpyAsync(h_dist, dd, n * sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
        if (d8) e5 = hipMemcpyAsync(h_caps8, d8, n * 4 * sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
        e4 = hipStreamSynchronize(stream_);
    }
    if (dc) (void)hipFree(dc);
    if (de) (void)hipFree(de);
    if (dd) (void)hipFree(dd);
    if (d8) (void)hipFree(d8);
    GCFG_HIP(e0);
    GCFG_HIP(e1);
    GCFG_HIP(e2);
    GCFG_HIP(e3);
    GCFG_HIP(e5);
    GCFG_HIP(e4);
    return GCFG_OK;
}

int Engine::sync() {
    GCFG_HIP(hipStreamSynchronize(stream_));
    return GCFG_OK;
}

}  // namespace gcfg
