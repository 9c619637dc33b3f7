// Serving: continuous batching of segmentation requests on one engine.
//
// The reference re-initialises the whole graph for every frame
// (ComplexApplication.cu:127-135 -> GC_SetGraph, GraphCut.cu:301-319). Here a
// session keeps its image, data term and residual graph resident in HBM; a new
// image is admitted into the running engine as soon as its memory fits the
// budget (other instances keep solving), and an edit (seed strokes) changes
// the excess of the touched nodes and re-arms the instance from its current
// residual graph (warm start: global relabel, then push-relabel).
//
// The host never waits for the device on the request path. Three streams:
//  * the engine stream runs data terms, builds and chunks of solver steps;
//    after each chunk the instance states are copied into one of two pinned
//    snapshots and an event is recorded, so the next chunk runs while the
//    host reads the previous snapshot;
//  * an upload stream copies request images and seed maps; the engine stream
//    waits for them with an event;
//  * packed labels (one bit per pixel) reach pinned staging without a copy
//    of their own: the request's staging is assigned when its solve starts,
//    and the scheduler kernel writes the words there in the step the
//    instance finishes, so the snapshot that shows it done completes it;
//  * a download stream copies a finished instance's int32 labels into pinned
//    staging; the request completes when that copy's event has passed.
//  Either way the labels then go to the caller's buffer, or are lent.
// Small uploads (descriptors, slot lists, strokes) go through the engine's
// pinned ring. Session buffers (image, seed map, data term, labels) return to
// a pool when a session closes and serve later sessions of at most as many
// pixels, so admission rarely allocates.
#include <algorithm>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <new>
#include <vector>

#include "dataterm.hpp"
#include "engine.hpp"

using gcfg::DtContext;
using gcfg::DtJob;
using gcfg::DtLaunch;
using gcfg::dt_chunks;
using gcfg::dt_freeze;
using gcfg::dt_launch;
using gcfg::Engine;
using gcfg::InstDesc;
using gcfg::InstState;
using gcfg::now_ms;
using gcfg::set_error;

namespace {

enum { REQ_IMAGE = 0, REQ_EDIT = 1, REQ_FRAME = 2 };

// One device allocation made with the server and carved into session
// blocks (first fit, neighbouring free ranges merged), so that admitting a
// session allocates nothing.
class DevHeap {
   public:
    ~DevHeap() {
        if (base_) (void)hipFree(base_);
    }
    int init(size_t bytes) {
        bytes &= ~(size_t)255;
        if (bytes && hipMalloc((void **)&base_, bytes) != hipSuccess) {
            base_ = nullptr;
            set_error("server: session heap allocation failed");
            return GCFG_ERR_NOMEM;
        }
        if (bytes) free_[0] = bytes;
        return GCFG_OK;
    }
    void *alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, rem = it->second - bytes;
            free_.erase(it);
            if (rem) free_[off + bytes] = rem;
            return base_ + off;
        }
        return nullptr;
    }
    void release(void *p, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        size_t off = (size_t)(static_cast<uint8_t *>(p) - base_);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }

   private:
    uint8_t *base_ = nullptr;
    std::map<size_t, size_t> free_;  // offset -> bytes
};

// Device buffers of one session: the data term and the label map, carved
// from one heap block.
struct Buffers {
    DtContext dt;
    void *block = nullptr;
    size_t block_bytes = 0;
    void *d_label = nullptr;  // int32 labels or packed words
    size_t label_bytes = 0;
    size_t dt_bytes0 = 0;  // the data term's bytes right after init (all in the block)
    // Block plus what the data term allocated itself (video frame planes).
    size_t bytes() const { return block_bytes + (dt.device_bytes() - dt_bytes0); }
};

// Pinned host memory that a finished request's labels are copied into
// (coherent: the scheduler kernel writes packed labels there directly; d is
// its device-visible address).
struct Staging {
    void *h = nullptr;
    void *d = nullptr;
    size_t bytes = 0;
};

struct Request {
    int64_t id = 0;
    int kind = REQ_IMAGE;
    int32_t session = 0;
    double t_submit = 0, t_admit = 0;
    int W = 0, H = 0;
    const uint8_t *rgb = nullptr, *seeds = nullptr;
    void *h_label = nullptr;
    int32_t format = GCFG_LABELS_INT32;
    std::vector<int32_t> strokes;
    int32_t delta = 0;
    int64_t first_snapshot = 0;  // snapshots from this sequence number on see the admitted instance
    int64_t flow = 0;
    Staging stg;
    size_t label_bytes = 0;
    struct CopyGroup *grp = nullptr;  // the label copies queued with this one (one completion event)
};

// The label copies queued from one snapshot: one wait on the snapshot, the
// copies, one event after them (instead of three calls per request).
struct CopyGroup {
    hipEvent_t ev = nullptr;
    int pending = 0;      // requests of the group not finished yet
    int state = 0;        // 0 not known done, 1 done, -1 failed (as of the current pass)
};

struct Session {
    int32_t id = 0;
    int slot = -1;
    int W = 0, H = 0;
    Buffers *buf = nullptr;
    size_t tile_bytes = 0;
    bool busy = false;
    double last_used = 0;
    // Sum over the session's edits of delta x strokes: a bound on how far the
    // edits have moved any node's pos - neg. Kept at most 2^29 so that the
    // excess of every node stays inside int32 (solver.hip, tlinks_bad).
    int64_t edit_mass = 0;
};
constexpr int64_t EDIT_MASS_MAX = (int64_t)1 << 29;

struct Snapshot {
    InstState *h = nullptr;  // pinned
    hipEvent_t ev = nullptr;
    int64_t seq = 0;
    bool pending = false;
};
constexpr int SNAPSHOTS = 2;  // chunks of steps in flight

size_t label_bytes_for(int W, int H, int32_t format) {
    return format == GCFG_LABELS_BITS ? Engine::label_words(W, H) * sizeof(uint64_t)
                                      : (size_t)W * H * sizeof(int32_t);
}

}  // namespace

struct gcfg_server {
    Engine eng;
    gcfg_options opt{};
    size_t budget = 0, used = 0;
    int cap_slots = 0;
    int64_t next_id = 1;
    int32_t next_session = 1;
    int32_t label_format = GCFG_LABELS_INT32;
    hipStream_t up = nullptr, down = nullptr;
    hipEvent_t up_ev = nullptr;
    std::deque<Request *> queue;
    std::vector<Request *> running, copying;
    std::map<int32_t, Session *> sessions;
    std::map<int, Session *> by_slot;
    DevHeap heap;                     // session buffers
    std::vector<Buffers *> pool;      // idle session buffers (counted in used)
    std::vector<Staging> staging;     // idle pinned label staging
    std::vector<Staging> lent;        // staging whose labels the last poll lent to the caller
    std::vector<hipEvent_t> events;   // idle events
    Snapshot snaps[SNAPSHOTS];
    int64_t snap_seq = 0;             // sequence number of the next snapshot
    int hi_slot = 0;                  // slots [0, hi_slot) may be in use
    std::deque<gcfg_result> ready;
    // Batched data term of the images admitted in one pass: the job table
    // (one per slot at most) and the job of every chunk, uploaded through the
    // engine's staging ring.
    DtJob *d_jobs = nullptr;
    int32_t *d_cmap = nullptr;
    size_t cmap_cap = 0;
    std::vector<DtJob> h_jobs;
    std::vector<int32_t> h_cmap;
};

static size_t tile_bytes_for(const Engine &eng, int W, int H) {
    return (size_t)eng.tiles_for(W, H) * Engine::bytes_per_tile();
}

static void free_buffers(gcfg_server *s, Buffers *b) {
    s->used -= std::min(s->used, b->bytes());
    if (b->block) s->heap.release(b->block, b->block_bytes);
    delete b;
}

static void free_pool_entry(gcfg_server *s, size_t i) {
    Buffers *b = s->pool[i];
    s->pool.erase(s->pool.begin() + (long)i);
    free_buffers(s, b);
}

static void destroy_session(gcfg_server *s, Session *ss) {
    if (ss->slot >= 0) {
        s->eng.remove_instance(ss->slot);
        s->by_slot.erase(ss->slot);
    }
    s->used -= std::min(s->used, ss->tile_bytes);
    if (ss->buf) s->pool.push_back(ss->buf);  // its bytes stay counted until the pool entry is freed
    s->sessions.erase(ss->id);
    delete ss;
}

static void push_result(gcfg_server *s, Request *r, int status, int64_t flow) {
    gcfg_result res{};
    res.request_id = r->id;
    res.session = r->session;
    res.status = status;
    res.flow = flow;
    const double t = now_ms();
    res.latency_ms = t - r->t_submit;
    res.solve_ms = r->t_admit > 0 ? t - r->t_admit : 0.0;
    res.done_ms = t;
    s->ready.push_back(res);
}

static bool session_has_queued(gcfg_server *s, int32_t id) {
    for (Request *q : s->queue)
        if ((q->kind == REQ_EDIT || q->kind == REQ_FRAME) && q->session == id) return true;
    return false;
}

// Free pooled buffers (except `keep`, which the pending request will use),
// then evict idle sessions least recently used first, until `need` more
// bytes fit.
static bool make_room(gcfg_server *s, size_t need, const Buffers *keep) {
    while (s->used + need > s->budget) {
        int big = -1;
        for (size_t i = 0; i < s->pool.size(); ++i)
            if (s->pool[i] != keep && (big < 0 || s->pool[i]->bytes() > s->pool[big]->bytes())) big = (int)i;
        if (big >= 0) {
            free_pool_entry(s, (size_t)big);
            continue;
        }
        Session *victim = nullptr;
        for (auto &kv : s->sessions) {
            Session *ss = kv.second;
            if (ss->busy || session_has_queued(s, ss->id)) continue;
            if (!victim || ss->last_used < victim->last_used) victim = ss;
        }
        if (!victim) return false;
        destroy_session(s, victim);
    }
    return true;
}

// Label bytes a W x H session needs in either layout.
static size_t label_need(int W, int H) {
    return std::max(label_bytes_for(W, H, GCFG_LABELS_INT32), label_bytes_for(W, H, GCFG_LABELS_BITS));
}

// The smallest pooled buffer set that holds a W x H session and is not more
// than twice its size, or -1.
static int pooled_for(const gcfg_server *s, int W, int H) {
    const size_t n = (size_t)W * H;
    int best = -1;
    for (size_t i = 0; i < s->pool.size(); ++i) {
        const size_t cap = s->pool[i]->dt.capacity();
        if (cap >= n && cap <= 2 * n && s->pool[i]->label_bytes >= label_need(W, H) &&
            (best < 0 || cap < s->pool[best]->dt.capacity()))
            best = (int)i;
    }
    return best;
}

// Buffers for a W x H session: a pooled set (pooled_for) or a new block
// from the heap (GCFG_ERR_BUDGET if the heap has no room for it now).
static int take_buffers(gcfg_server *s, int W, int H, Buffers **out) {
    const size_t n = (size_t)W * H;
    const int best = pooled_for(s, W, H);
    Buffers *b = nullptr;
    if (best >= 0) {
        b = s->pool[best];
        s->pool.erase(s->pool.begin() + best);
        s->used -= std::min(s->used, b->bytes());
        const int rc = b->dt.reshape(W, H);
        if (rc != GCFG_OK) {
            s->used += b->bytes();
            free_buffers(s, b);
            return rc;
        }
    } else {
        const size_t dtb = DtContext::block_bytes(n), lb = label_need(W, H);
        void *block = s->heap.alloc(dtb + lb);
        for (size_t i = 0; !block && !s->pool.empty();) {  // pooled sets that do not fit make room
            (void)i;
            free_pool_entry(s, 0);
            block = s->heap.alloc(dtb + lb);
        }
        if (!block) {
            set_error("server: no heap block for the session (fragmented or held by live sessions)");
            return GCFG_ERR_BUDGET;
        }
        b = new (std::nothrow) Buffers();
        if (!b) {
            s->heap.release(block, dtb + lb);
            return GCFG_ERR_NOMEM;
        }
        b->block = block;
        b->block_bytes = dtb + lb;
        b->d_label = static_cast<uint8_t *>(block) + dtb;
        b->label_bytes = lb;
        const int rc = b->dt.init(W, H, s->eng.stream(), true, n, block);
        if (rc != GCFG_OK) {
            s->heap.release(block, dtb + lb);
            delete b;
            return rc;
        }
        b->dt_bytes0 = b->dt.device_bytes();
    }
    s->used += b->bytes();
    *out = b;
    return GCFG_OK;
}

static int stage_for_device(gcfg_server *s, Request *r, int W, int H);
static void drop_staging(gcfg_server *s, Request *r);

// An image request's session: buffers, the image and seed map copied on
// the upload stream, an engine slot pointing at the buffers' planes. The
// data term and the build are queued for all images of an admission pass
// together (admit_finish).
static int admit_image(gcfg_server *s, Request *r, Session **out) {
    Session *ss = new (std::nothrow) Session();
    if (!ss) return GCFG_ERR_NOMEM;
    ss->id = r->session;
    ss->W = r->W;
    ss->H = r->H;
    int rc = take_buffers(s, r->W, r->H, &ss->buf);
    if (rc != GCFG_OK) {
        delete ss;
        return rc;
    }
    DtContext &dt = ss->buf->dt;
    rc = dt.upload_on(s->up, r->rgb, r->seeds);
    if (rc == GCFG_OK) {
        const bool bits = r->format == GCFG_LABELS_BITS;
        ss->slot = s->eng.add_instance(r->W, r->H, dt.pos_, dt.neg_, dt.up_, dt.down_, dt.left_, dt.right_, 0,
                                       bits ? nullptr : (int32_t *)ss->buf->d_label);
        if (ss->slot < 0) rc = ss->slot;
        else s->eng.set_label_bits(ss->slot, bits ? (uint64_t *)ss->buf->d_label : nullptr);
    }
    if (rc == GCFG_OK) rc = stage_for_device(s, r, r->W, r->H);
    if (rc == GCFG_OK) s->eng.set_label_host(ss->slot, (uint64_t *)r->stg.d);
    if (rc != GCFG_OK) {
        drop_staging(s, r);
        if (ss->slot >= 0) s->eng.remove_instance(ss->slot);
        (void)hipStreamSynchronize(s->up);  // the upload ends before the buffers serve another session
        free_buffers(s, ss->buf);
        delete ss;
        return rc;
    }
    ss->tile_bytes = tile_bytes_for(s->eng, r->W, r->H);
    s->used += ss->tile_bytes;
    s->sessions[ss->id] = ss;
    s->by_slot[ss->slot] = ss;
    s->hi_slot = std::max(s->hi_slot, ss->slot + 1);
    ss->busy = true;
    ss->last_used = now_ms();
    *out = ss;
    return GCFG_OK;
}

// The GPU work of the images admitted in one pass, queued once for all of
// them: one wait for their uploads, one batched data term (stats, table and
// links launches over every chunk of every image; the per-image launch
// latencies were most of a small request's GPU time), one launch that keeps
// each image's statistics as its session's colour model, one build.
static int admit_finish(gcfg_server *s, const std::vector<Session *> &fresh) {
    if (fresh.empty()) return GCFG_OK;
    if (hipEventRecord(s->up_ev, s->up) != hipSuccess || hipStreamWaitEvent(s->eng.stream(), s->up_ev, 0) != hipSuccess) {
        set_error("server: ordering the uploads failed");
        return GCFG_ERR_HIP;
    }
    s->h_jobs.clear();
    s->h_cmap.clear();
    int lambda = 0;
    std::vector<int> slots;
    for (Session *ss : fresh) {
        DtJob j = ss->buf->dt.batch_job();
        j.chunk0 = (int32_t)s->h_cmap.size();
        const int nc = dt_chunks(j.W, j.H);
        s->h_cmap.insert(s->h_cmap.end(), (size_t)nc, (int32_t)s->h_jobs.size());
        lambda = std::max(lambda, (int)j.lambda);
        s->h_jobs.push_back(j);
        slots.push_back(ss->slot);
    }
    if (s->h_cmap.size() > s->cmap_cap || s->h_jobs.size() > (size_t)s->cap_slots) {
        set_error("server: admission pass larger than its job tables");
        return GCFG_ERR_STATE;
    }
    int rc = s->eng.stage(s->d_jobs, s->h_jobs.data(), s->h_jobs.size() * sizeof(DtJob));
    if (rc == GCFG_OK) rc = s->eng.stage(s->d_cmap, s->h_cmap.data(), s->h_cmap.size() * sizeof(int32_t));
    if (rc != GCFG_OK) return rc;
    DtLaunch L{};
    L.jobs = s->d_jobs;
    L.cmap = s->d_cmap;
    L.njobs = (int32_t)s->h_jobs.size();
    L.lds_lambda = lambda;
    dt_launch(L, (int)s->h_cmap.size(), s->eng.stream());
    dt_freeze(s->d_jobs, L.njobs, s->eng.stream());
    if (hipGetLastError() != hipSuccess) {
        set_error("server: data-term launch failed");
        return GCFG_ERR_HIP;
    }
    for (Session *ss : fresh) ss->buf->dt.batch_queued();  // later frames reuse this colour model
    return s->eng.build(slots);
}

// Label destination of a session's next solve in the request's format
// (and the staging the device writes packed labels to).
static void point_labels(gcfg_server *s, Session *ss, const Request *r) {
    const bool bits = r->format == GCFG_LABELS_BITS;
    const InstDesc &d = s->eng.desc(ss->slot);
    s->eng.set_inputs(ss->slot, d.pos, d.neg, d.up, d.down, d.left, d.right, d.penalty,
                      bits ? nullptr : (int32_t *)ss->buf->d_label);
    s->eng.set_label_bits(ss->slot, bits ? (uint64_t *)ss->buf->d_label : nullptr);
    s->eng.set_label_host(ss->slot, (uint64_t *)r->stg.d);
}

static hipEvent_t take_event(gcfg_server *s) {
    if (!s->events.empty()) {
        hipEvent_t e = s->events.back();
        s->events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Pinned staging in power-of-two sizes from 4 KiB: a freed buffer serves any
// later request of its size class, so the steady state allocates none (an
// allocation of pinned memory costs tens of microseconds on the request
// path).
static size_t staging_class(size_t bytes) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

static int take_staging(gcfg_server *s, size_t bytes, Staging *out) {
    const size_t want = staging_class(bytes);
    int best = -1;
    for (size_t i = 0; i < s->staging.size(); ++i)
        if (s->staging[i].bytes >= want && (best < 0 || s->staging[i].bytes < s->staging[best].bytes))
            best = (int)i;
    if (best >= 0) {
        *out = s->staging[best];
        s->staging[best] = s->staging.back();
        s->staging.pop_back();
        return GCFG_OK;
    }
    Staging st;
    st.bytes = want;
    if (hipHostMalloc(&st.h, st.bytes, hipHostMallocCoherent) != hipSuccess) {
        set_error("server: pinned staging allocation failed");
        return GCFG_ERR_NOMEM;
    }
    if (hipHostGetDevicePointer(&st.d, st.h, 0) != hipSuccess || !st.d) {
        (void)hipHostFree(st.h);
        set_error("server: pinned staging has no device address");
        return GCFG_ERR_HIP;
    }
    *out = st;
    return GCFG_OK;
}

// Packed labels are written to the request's staging by the device when
// its instance finishes (Engine::set_label_host); int32 labels are copied
// after it.
static bool labels_by_device(const Request *r) { return r->format == GCFG_LABELS_BITS; }

// Staging for a request whose labels the device writes, taken as its solve
// starts.
static int stage_for_device(gcfg_server *s, Request *r, int W, int H) {
    if (!labels_by_device(r)) return GCFG_OK;
    r->label_bytes = label_bytes_for(W, H, r->format);
    return take_staging(s, r->label_bytes, &r->stg);
}

// A request that ends before its solve could write the staging: the
// staging goes back to the pool.
static void drop_staging(gcfg_server *s, Request *r) {
    if (r->stg.h) s->staging.push_back(r->stg);
    r->stg = Staging{};
}

static void finish(gcfg_server *s, Request *r, int status) {
    auto it = s->sessions.find(r->session);
    if (it != s->sessions.end()) {
        it->second->busy = false;
        it->second->last_used = now_ms();
    }
    push_result(s, r, status, status == GCFG_OK ? r->flow : 0);
    if (r->stg.h) {
        if (status == GCFG_OK && !r->h_label) {
            // Lent: the result points at the staging until the next poll.
            s->ready.back().labels = r->stg.h;
            s->ready.back().label_bytes = r->label_bytes;
            s->lent.push_back(r->stg);
        } else {
            s->staging.push_back(r->stg);
        }
    }
    if (r->grp && --r->grp->pending == 0) {
        s->events.push_back(r->grp->ev);
        delete r->grp;
    }
    delete r;
}

// Requests whose instance a completed snapshot shows done: queue the label
// copy on the download stream.
static int process_snapshot(gcfg_server *s, const Snapshot &sn) {
    std::vector<Request *> still, done;
    std::vector<const void *> src;  // the device labels of each request in `done`
    for (Request *r : s->running) {
        auto it = s->sessions.find(r->session);
        if (it == s->sessions.end()) {
            set_error("server: session vanished while solving");
            drop_staging(s, r);
            push_result(s, r, GCFG_ERR_STATE, 0);
            delete r;
            continue;
        }
        Session *ss = it->second;
        const InstState &st = sn.h[ss->slot];
        if (sn.seq < r->first_snapshot || st.phase != gcfg::PH_DONE) {
            still.push_back(r);
            continue;
        }
        if (st.err) {
            set_error("server: request inputs outside the supported range");
            finish(s, r, GCFG_ERR_INVALID);
            continue;
        }
        r->flow = st.flow;
        if (labels_by_device(r)) {
            // The scheduler wrote the packed labels to the staging in the
            // step the instance finished, before this snapshot.
            if (r->h_label) std::memcpy(r->h_label, r->stg.h, r->label_bytes);
            finish(s, r, GCFG_OK);
            continue;
        }
        r->label_bytes = label_bytes_for(ss->W, ss->H, r->format);
        const int rc = take_staging(s, r->label_bytes, &r->stg);
        if (rc != GCFG_OK) {
            finish(s, r, rc);
            continue;
        }
        done.push_back(r);
        src.push_back(ss->buf->d_label);
    }
    s->running.swap(still);
    if (done.empty()) return GCFG_OK;
    CopyGroup *g = new (std::nothrow) CopyGroup();
    if (g) g->ev = take_event(s);
    bool ok = g && g->ev && hipStreamWaitEvent(s->down, sn.ev, 0) == hipSuccess;
    for (size_t i = 0; ok && i < done.size(); ++i) {
        Request *r = done[i];
        ok = hipMemcpyAsync(r->stg.h, src[i], r->label_bytes, hipMemcpyDeviceToHost, s->down) == hipSuccess;
    }
    ok = ok && hipEventRecord(g->ev, s->down) == hipSuccess;
    if (!ok) {
        set_error("server: label copy failed");
        (void)hipStreamSynchronize(s->down);  // nothing queued may still write the staging
        if (g) {
            if (g->ev) s->events.push_back(g->ev);
            delete g;
        }
        for (Request *r : done) finish(s, r, GCFG_ERR_HIP);
        return GCFG_OK;
    }
    g->pending = (int)done.size();
    for (Request *r : done) {
        r->grp = g;
        s->copying.push_back(r);
    }
    return GCFG_OK;
}

// Label copies that have landed: hand the labels to the caller.
static bool finish_copies(gcfg_server *s, bool block_first) {
    bool any = false;
    std::vector<Request *> still;
    for (Request *r : s->copying) r->grp->state = 0;
    for (size_t i = 0; i < s->copying.size(); ++i) {
        Request *r = s->copying[i];
        CopyGroup *g = r->grp;
        if (g->state == 0) {
            const hipError_t q = (block_first && i == 0) ? hipEventSynchronize(g->ev) : hipEventQuery(g->ev);
            g->state = q == hipSuccess ? 1 : q == hipErrorNotReady ? 2 : -1;
        }
        if (g->state == 2) {
            still.push_back(r);
            continue;
        }
        if (g->state < 0) {
            set_error("server: label copy failed");
            finish(s, r, GCFG_ERR_HIP);
            any = true;
            continue;
        }
        if (r->h_label) std::memcpy(r->h_label, r->stg.h, r->label_bytes);
        finish(s, r, GCFG_OK);
        any = true;
    }
    s->copying.swap(still);
    return any;
}

// Oldest pending snapshot index, or -1.
static int oldest_snapshot(gcfg_server *s) {
    int best = -1;
    for (int i = 0; i < SNAPSHOTS; ++i)
        if (s->snaps[i].pending && (best < 0 || s->snaps[i].seq < s->snaps[best].seq)) best = i;
    return best;
}

static int drain_snapshots(gcfg_server *s, bool block, bool *progress) {
    for (;;) {
        const int i = oldest_snapshot(s);
        if (i < 0) return GCFG_OK;
        Snapshot &sn = s->snaps[i];
        const hipError_t q = block ? hipEventSynchronize(sn.ev) : hipEventQuery(sn.ev);
        if (q == hipErrorNotReady) return GCFG_OK;
        if (q != hipSuccess) return gcfg::hip_fail(q, "server snapshot");
        sn.pending = false;
        block = false;
        const int rc = process_snapshot(s, sn);
        if (rc != GCFG_OK) return rc;
        *progress = true;
    }
}

// Launch a chunk of steps followed by a snapshot, while fewer than
// SNAPSHOTS chunks are in flight.
// A snapshot slot is free, so feed() can launch the next chunk.
static bool can_pump(const gcfg_server *s) {
    for (int i = 0; i < SNAPSHOTS; ++i)
        if (!s->snaps[i].pending) return true;
    return false;
}

static int feed(gcfg_server *s) {
    if (s->running.empty()) return GCFG_OK;
    for (int i = 0; i < SNAPSHOTS; ++i) {
        Snapshot &sn = s->snaps[i];
        if (sn.pending) continue;
        int rc = s->eng.pump(s->opt.steps_per_poll);
        if (rc == GCFG_OK) rc = s->eng.snapshot_states(sn.h, s->hi_slot, sn.ev);
        if (rc != GCFG_OK) return rc;
        sn.seq = s->snap_seq++;
        sn.pending = true;
        return GCFG_OK;  // one chunk per call: admissions in between join the next chunk
    }
    return GCFG_OK;
}

static int admit(gcfg_server *s) {
    std::deque<Request *> keep;
    std::vector<Session *> fresh;  // images admitted in this pass (GPU work queued together at the end)
    std::vector<Request *> fresh_req;
    bool blocked = false;  // an image that does not fit blocks later images (FIFO fairness)
    while (!s->queue.empty()) {
        Request *r = s->queue.front();
        s->queue.pop_front();
        if (r->kind == REQ_EDIT || r->kind == REQ_FRAME) {
            auto it = s->sessions.find(r->session);
            if (it == s->sessions.end()) {
                bool image_pending = false;  // the session's image is still queued
                for (Request *q : keep) image_pending |= q->kind == REQ_IMAGE && q->session == r->session;
                for (Request *q : s->queue) image_pending |= q->kind == REQ_IMAGE && q->session == r->session;
                if (image_pending) {
                    keep.push_back(r);
                    continue;
                }
                set_error("server: edit of an unknown or evicted session");
                push_result(s, r, GCFG_ERR_STATE, 0);
                delete r;
                continue;
            }
            Session *ss = it->second;
            if (ss->busy) {
                keep.push_back(r);
                continue;
            }
            int rc = GCFG_OK;
            if (r->kind == REQ_EDIT) {
                const int64_t mass = (int64_t)r->delta * (int64_t)(r->strokes.size() / 5);
                if (ss->edit_mass + mass > EDIT_MASS_MAX) {
                    set_error("server: the session's edits would move pos - neg by more than 2^29 "
                              "(sum of delta x strokes over its edits)");
                    rc = GCFG_ERR_INVALID;
                } else {
                    ss->edit_mass += mass;
                    rc = s->eng.edit(ss->slot, r->strokes, r->delta);
                }
            } else if (r->W != ss->W || r->H != ss->H) {
                set_error("server: frame size differs from the session's");
                rc = GCFG_ERR_INVALID;
            } else {
                // Next frame: data term of the new image with the session's
                // colour model, then the residual graph reparametrised onto it.
                DtContext &dt = ss->buf->dt;
                const size_t before = ss->buf->bytes();
                rc = dt.upload_on(s->up, r->rgb, nullptr);
                if (rc == GCFG_OK && (hipEventRecord(s->up_ev, s->up) != hipSuccess ||
                                      hipStreamWaitEvent(s->eng.stream(), s->up_ev, 0) != hipSuccess)) {
                    set_error("server: ordering the upload failed");
                    rc = GCFG_ERR_HIP;
                }
                if (rc == GCFG_OK) rc = dt.compute_swap();
                s->used += ss->buf->bytes() - before;  // frame planes allocated on the first frame
                if (rc == GCFG_OK)
                    rc = s->eng.reframe(ss->slot, dt.pos_, dt.neg_, dt.up_, dt.down_, dt.left_, dt.right_, 0);
            }
            if (rc == GCFG_OK) rc = stage_for_device(s, r, ss->W, ss->H);
            if (rc == GCFG_OK) point_labels(s, ss, r);
            if (rc == GCFG_OK) rc = s->eng.upload_desc_slot(ss->slot);
            if (rc == GCFG_OK) rc = s->eng.rearm({ss->slot});
            if (rc != GCFG_OK) {
                drop_staging(s, r);
                push_result(s, r, rc, 0);
                delete r;
                continue;
            }
            r->t_admit = now_ms();
            r->first_snapshot = s->snap_seq;
            ss->busy = true;
            ss->last_used = r->t_admit;
            s->running.push_back(r);
            continue;
        }
        const size_t full = tile_bytes_for(s->eng, r->W, r->H) + DtContext::block_bytes((size_t)r->W * r->H) +
                            label_need(r->W, r->H);
        if (full > s->budget || s->eng.tiles_for(r->W, r->H) > s->eng.cap_tiles()) {
            set_error("server: request larger than the HBM budget");
            push_result(s, r, GCFG_ERR_BUDGET, 0);
            delete r;
            continue;
        }
        // A pooled buffer set is already counted: only the tiles are new then.
        const int pb = blocked ? -1 : pooled_for(s, r->W, r->H);
        const Buffers *reuse = pb >= 0 ? s->pool[pb] : nullptr;
        const size_t need = reuse ? tile_bytes_for(s->eng, r->W, r->H) : full;
        if (blocked || !make_room(s, need, reuse) || (size_t)s->eng.tiles_for(r->W, r->H) > s->eng.tiles_free()) {
            blocked = true;
            keep.push_back(r);
            continue;
        }
        r->t_admit = now_ms();
        r->first_snapshot = s->snap_seq;
        Session *ss = nullptr;
        int rc = admit_image(s, r, &ss);
        if (rc == GCFG_ERR_BUDGET) {  // tile range fragmented: wait for frees
            blocked = true;
            r->t_admit = 0;
            keep.push_back(r);
            continue;
        }
        if (rc != GCFG_OK) {
            push_result(s, r, rc, 0);
            delete r;
            continue;
        }
        fresh.push_back(ss);
        fresh_req.push_back(r);
    }
    s->queue.swap(keep);
    const int rc = admit_finish(s, fresh);
    if (rc != GCFG_OK) {
        // Nothing of the pass runs: its requests fail, and their buffers
        // return to the pool once the queued work on them has ended.
        (void)s->eng.sync();
        for (size_t i = 0; i < fresh.size(); ++i) {
            drop_staging(s, fresh_req[i]);
            push_result(s, fresh_req[i], rc, 0);
            delete fresh_req[i];
            destroy_session(s, fresh[i]);
        }
        return rc;
    }
    s->running.insert(s->running.end(), fresh_req.begin(), fresh_req.end());
    return GCFG_OK;
}

extern "C" {

double gcfg_now_ms(void) { return now_ms(); }

int gcfg_server_create(size_t hbm_budget_bytes, const gcfg_options *opt, gcfg_server **out) {
    if (!out || hbm_budget_bytes < ((size_t)1 << 20)) {
        set_error("gcfg_server_create: budget must be at least 1 MiB");
        return GCFG_ERR_INVALID;
    }
    *out = nullptr;
    gcfg_server *s = new (std::nothrow) gcfg_server();
    if (!s) return GCFG_ERR_NOMEM;
    if (opt)
        s->opt = *opt;
    else
        gcfg_default_options(&s->opt);
    s->budget = hbm_budget_bytes;
    s->cap_slots = 1024;
    // The budget is split between the tile arena and the session heap in the
    // proportion one session needs them: per 1024 pixels a tile and about
    // 32 bytes per pixel of data term, inputs and labels.
    const size_t per_tile = Engine::bytes_per_tile() + (size_t)gcfg::TN * 32;
    const int tiles = (int)std::max<size_t>(1, hbm_budget_bytes / per_tile);
    int rc = s->eng.init(tiles, s->cap_slots, s->opt);
    if (rc == GCFG_OK) {
        const size_t arena = (size_t)tiles * Engine::bytes_per_tile();
        rc = s->heap.init(hbm_budget_bytes > arena ? hbm_budget_bytes - arena : 0);
    }
    if (rc == GCFG_OK && (hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess ||
                          hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking) != hipSuccess ||
                          hipEventCreateWithFlags(&s->up_ev, hipEventDisableTiming) != hipSuccess)) {
        set_error("gcfg_server_create: stream or event creation failed");
        rc = GCFG_ERR_HIP;
    }
    if (rc == GCFG_OK) {
        // An image has at most a quarter as many data-term chunks as tiles,
        // plus one, and every admitted image fits in the arena.
        s->cmap_cap = (size_t)tiles / 4 + (size_t)s->cap_slots;
        if (hipMalloc((void **)&s->d_jobs, (size_t)s->cap_slots * sizeof(DtJob)) != hipSuccess ||
            hipMalloc((void **)&s->d_cmap, s->cmap_cap * sizeof(int32_t)) != hipSuccess) {
            set_error("gcfg_server_create: job table allocation failed");
            rc = GCFG_ERR_NOMEM;
        }
    }
    for (int i = 0; rc == GCFG_OK && i < SNAPSHOTS; ++i) {
        if (hipHostMalloc((void **)&s->snaps[i].h, (size_t)s->cap_slots * sizeof(InstState),
                          hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&s->snaps[i].ev, hipEventDisableTiming) != hipSuccess) {
            set_error("gcfg_server_create: snapshot allocation failed");
            rc = GCFG_ERR_NOMEM;
        }
    }
    if (rc != GCFG_OK) {
        gcfg_server_destroy(s);
        return rc;
    }
    *out = s;
    return GCFG_OK;
}

int gcfg_server_set_label_format(gcfg_server *s, int32_t format) {
    if (!s || (format != GCFG_LABELS_INT32 && format != GCFG_LABELS_BITS)) {
        set_error("gcfg_server_set_label_format: format must be GCFG_LABELS_INT32 or GCFG_LABELS_BITS");
        return GCFG_ERR_INVALID;
    }
    s->label_format = format;
    return GCFG_OK;
}

int gcfg_server_submit_image(gcfg_server *s, int32_t width, int32_t height, const uint8_t *h_rgb,
                             const uint8_t *h_seeds, int32_t *h_label, int64_t *request_id, int32_t *session) {
    if (!s || !Engine::shape_supported(width, height) || !h_rgb || !h_seeds) {
        set_error("gcfg_server_submit_image: invalid arguments (shape positive, at most 2^30 pixels)");
        return GCFG_ERR_INVALID;
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return GCFG_ERR_NOMEM;
    r->id = s->next_id++;
    r->kind = REQ_IMAGE;
    r->session = s->next_session++;
    r->t_submit = now_ms();
    r->W = width;
    r->H = height;
    r->rgb = h_rgb;
    r->seeds = h_seeds;
    r->h_label = h_label;
    r->format = s->label_format;
    s->queue.push_back(r);
    if (request_id) *request_id = r->id;
    if (session) *session = r->session;
    return GCFG_OK;
}

int gcfg_server_submit_edit(gcfg_server *s, int32_t session, int32_t n_strokes, const int32_t *strokes,
                            int32_t delta, int32_t *h_label, int64_t *request_id) {
    if (!s || n_strokes < 0 || (n_strokes > 0 && !strokes) || delta < 0) {
        set_error("gcfg_server_submit_edit: invalid arguments");
        return GCFG_ERR_INVALID;
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return GCFG_ERR_NOMEM;
    r->id = s->next_id++;
    r->kind = REQ_EDIT;
    r->session = session;
    r->t_submit = now_ms();
    r->strokes.assign(strokes, strokes + 5 * (size_t)n_strokes);
    r->delta = delta;
    r->h_label = h_label;
    r->format = s->label_format;
    s->queue.push_back(r);
    if (request_id) *request_id = r->id;
    return GCFG_OK;
}

int gcfg_server_submit_frame(gcfg_server *s, int32_t session, int32_t width, int32_t height, const uint8_t *h_rgb,
                             int32_t *h_label, int64_t *request_id) {
    if (!s || width <= 0 || height <= 0 || !h_rgb) {
        set_error("gcfg_server_submit_frame: invalid arguments");
        return GCFG_ERR_INVALID;
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return GCFG_ERR_NOMEM;
    r->id = s->next_id++;
    r->kind = REQ_FRAME;
    r->session = session;
    r->t_submit = now_ms();
    r->W = width;
    r->H = height;
    r->rgb = h_rgb;
    r->h_label = h_label;
    r->format = s->label_format;
    s->queue.push_back(r);
    if (request_id) *request_id = r->id;
    return GCFG_OK;
}

int gcfg_server_poll(gcfg_server *s, int32_t wait, int32_t max_results, gcfg_result *results,
                     int32_t *n_results) {
    if (!s || (max_results > 0 && !results)) return GCFG_ERR_INVALID;
    // Labels lent by the previous poll expire now.
    for (const Staging &st : s->lent) s->staging.push_back(st);
    s->lent.clear();
    int rc = GCFG_OK;
    for (;;) {
        bool progress = finish_copies(s, false);
        rc = drain_snapshots(s, false, &progress);
        if (rc != GCFG_OK) break;
        // Admission right before a chunk is pumped: an admitted request's
        // work is queued on the engine stream behind the chunks in flight and
        // starts with the next chunk either way, so requests that arrived
        // since the last chunk are admitted together (one batched data term
        // and build), and the queue is not rescanned on every poll.
        if (can_pump(s)) {
            rc = admit(s);
            if (rc != GCFG_OK) break;
            rc = feed(s);
            if (rc != GCFG_OK) break;
        }
        const bool idle = s->running.empty() && s->queue.empty() && s->copying.empty();
        if (!wait || !s->ready.empty() || idle) break;
        if (!progress) {
            // Nothing moved: wait for the oldest chunk in flight, else the oldest label copy.
            if (oldest_snapshot(s) >= 0) {
                rc = drain_snapshots(s, true, &progress);
                if (rc != GCFG_OK) break;
            } else if (!s->copying.empty()) {
                finish_copies(s, true);
            } else if (!s->queue.empty()) {
                set_error("server: queued requests cannot be admitted (budget held by busy sessions)");
                rc = GCFG_ERR_BUDGET;
                break;
            }
        }
    }
    int n = 0;
    while (n < max_results && !s->ready.empty()) {
        results[n++] = s->ready.front();
        s->ready.pop_front();
    }
    if (n_results) *n_results = n;
    return rc;
}

int gcfg_server_close_session(gcfg_server *s, int32_t session) {
    if (!s) return GCFG_ERR_INVALID;
    auto it = s->sessions.find(session);
    if (it == s->sessions.end()) return GCFG_OK;
    if (it->second->busy) {
        set_error("gcfg_server_close_session: session has a request in flight");
        return GCFG_ERR_STATE;
    }
    destroy_session(s, it->second);
    return GCFG_OK;
}

int gcfg_server_load(gcfg_server *s, int32_t *in_flight, int32_t *queued, size_t *bytes_in_use) {
    if (!s) return GCFG_ERR_INVALID;
    if (in_flight) *in_flight = (int32_t)(s->running.size() + s->copying.size());
    if (queued) *queued = (int32_t)s->queue.size();
    if (bytes_in_use) *bytes_in_use = s->used;
    return GCFG_OK;
}

int gcfg_server_destroy(gcfg_server *s) {
    if (!s) return GCFG_OK;
    (void)s->eng.sync();
    if (s->up) (void)hipStreamSynchronize(s->up);
    if (s->down) (void)hipStreamSynchronize(s->down);
    for (Request *r : s->queue) delete r;
    for (Request *r : s->running) {
        if (r->stg.h) (void)hipHostFree(r->stg.h);
        delete r;
    }
    for (Request *r : s->copying) {
        if (r->stg.h) (void)hipHostFree(r->stg.h);
        if (r->grp && --r->grp->pending == 0) {
            (void)hipEventDestroy(r->grp->ev);
            delete r->grp;
        }
        delete r;
    }
    for (Staging &st : s->lent) (void)hipHostFree(st.h);
    std::vector<Session *> all;
    for (auto &kv : s->sessions) all.push_back(kv.second);
    for (Session *ss : all) destroy_session(s, ss);
    while (!s->pool.empty()) free_pool_entry(s, s->pool.size() - 1);
    for (Staging &st : s->staging) (void)hipHostFree(st.h);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    for (Snapshot &sn : s->snaps) {
        if (sn.h) (void)hipHostFree(sn.h);
        if (sn.ev) (void)hipEventDestroy(sn.ev);
    }
    if (s->up_ev) (void)hipEventDestroy(s->up_ev);
    if (s->d_jobs) (void)hipFree(s->d_jobs);
    if (s->d_cmap) (void)hipFree(s->d_cmap);
    if (s->up) (void)hipStreamDestroy(s->up);
    if (s->down) (void)hipStreamDestroy(s->down);
    delete s;
    return GCFG_OK;
}

}  // extern "C"
