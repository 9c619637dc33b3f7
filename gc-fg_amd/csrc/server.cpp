// Serving: continuous batching of segmentation requests on one engine.
//
// The reference re-initialises the whole graph for every frame
// (ComplexApplication.cu:127-135 -> GC_SetGraph, GraphCut.cu:301-319). Here a
// session keeps its image, data term and residual graph resident in HBM; a new
// image is admitted into the running engine as soon as its memory fits the
// budget (other instances keep solving), and an edit (seed strokes) changes
// the excess of the touched nodes and re-arms the instance from its current
// residual graph (warm start: global relabel, then push-relabel).
#include <algorithm>
#include <deque>
#include <map>
#include <new>
#include <vector>

#include "dataterm.hpp"
#include "engine.hpp"

using gcfg::DtContext;
using gcfg::Engine;
using gcfg::InstState;
using gcfg::now_ms;
using gcfg::set_error;

namespace {

enum { REQ_IMAGE = 0, REQ_EDIT = 1, REQ_FRAME = 2 };

struct Request {
    int64_t id = 0;
    int kind = REQ_IMAGE;
    int32_t session = 0;
    double t_submit = 0, t_admit = 0;
    int W = 0, H = 0;
    const uint8_t *rgb = nullptr, *seeds = nullptr;
    int32_t *h_label = nullptr;
    std::vector<int32_t> strokes;
    int32_t delta = 0;
};

struct Session {
    int32_t id = 0;
    int slot = -1;
    int W = 0, H = 0;
    DtContext *dt = nullptr;
    int32_t *d_label = nullptr;
    size_t bytes = 0;
    bool busy = false;
    double last_used = 0;
    // Sum over the session's edits of delta x strokes: a bound on how far the
    // edits have moved any node's pos - neg. Kept at most 2^29 so that the
    // excess of every node stays inside int32 (solver.hip, tlinks_bad).
    int64_t edit_mass = 0;
};
constexpr int64_t EDIT_MASS_MAX = (int64_t)1 << 29;

}  // namespace

struct gcfg_server {
    Engine eng;
    gcfg_options opt{};
    size_t budget = 0, used = 0;
    int cap_slots = 0;
    int64_t next_id = 1;
    int32_t next_session = 1;
    std::deque<Request *> queue;
    std::vector<Request *> running;
    std::map<int32_t, Session *> sessions;
    std::map<int, Session *> by_slot;
    std::vector<InstState> states;
    std::deque<gcfg_result> ready;
};

static size_t session_cost(const Engine &eng, int W, int H) {
    const size_t n = (size_t)W * H;
    return (size_t)eng.tiles_for(W, H) * Engine::bytes_per_tile() + n * (3 + 1 + 24 + 4) + 64 * 1024;
}

static void destroy_session(gcfg_server *s, Session *ss) {
    if (ss->slot >= 0) {
        s->eng.sync();
        s->eng.remove_instance(ss->slot);
        s->by_slot.erase(ss->slot);
    }
    delete ss->dt;
    if (ss->d_label) (void)hipFree(ss->d_label);
    s->used -= std::min(s->used, ss->bytes);
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

// Evict idle sessions, least recently used first, until `need` bytes fit.
static bool make_room(gcfg_server *s, size_t need) {
    while (s->used + need > s->budget) {
        Session *victim = nullptr;
        for (auto &kv : s->sessions) {
            Session *ss = kv.second;
            if (ss->busy) continue;
            bool pending = false;
            for (Request *q : s->queue)
                if ((q->kind == REQ_EDIT || q->kind == REQ_FRAME) && q->session == ss->id) pending = true;
            if (pending) continue;
            if (!victim || ss->last_used < victim->last_used) victim = ss;
        }
        if (!victim) return false;
        destroy_session(s, victim);
    }
    return true;
}

static int admit_image(gcfg_server *s, Request *r) {
    Session *ss = new (std::nothrow) Session();
    if (!ss) return GCFG_ERR_NOMEM;
    ss->id = r->session;
    ss->W = r->W;
    ss->H = r->H;
    ss->bytes = session_cost(s->eng, r->W, r->H);
    ss->dt = new (std::nothrow) DtContext();
    int rc = ss->dt ? ss->dt->init(r->W, r->H, s->eng.stream(), true) : GCFG_ERR_NOMEM;
    if (rc == GCFG_OK && hipMalloc((void **)&ss->d_label, (size_t)r->W * r->H * sizeof(int32_t)) != hipSuccess) {
        set_error("server: label buffer allocation failed");
        rc = GCFG_ERR_NOMEM;
    }
    if (rc == GCFG_OK) rc = ss->dt->upload_image(r->rgb);
    if (rc == GCFG_OK) rc = ss->dt->upload_seeds(r->seeds);
    if (rc == GCFG_OK) rc = ss->dt->compute_owned();
    if (rc == GCFG_OK) rc = ss->dt->freeze_model();  // later frames of the session reuse this colour model
    if (rc == GCFG_OK) {
        ss->slot = s->eng.add_instance(r->W, r->H, ss->dt->pos_, ss->dt->neg_, ss->dt->up_, ss->dt->down_,
                                       ss->dt->left_, ss->dt->right_, 0, ss->d_label);
        if (ss->slot < 0) rc = ss->slot;
    }
    if (rc == GCFG_OK) rc = s->eng.build({ss->slot});
    if (rc != GCFG_OK) {
        if (ss->slot >= 0) s->eng.remove_instance(ss->slot);
        delete ss->dt;
        if (ss->d_label) (void)hipFree(ss->d_label);
        delete ss;
        return rc;
    }
    s->used += ss->bytes;
    s->sessions[ss->id] = ss;
    s->by_slot[ss->slot] = ss;
    ss->busy = true;
    ss->last_used = now_ms();
    return GCFG_OK;
}

static int collect(gcfg_server *s) {
    if (s->running.empty()) return GCFG_OK;
    int rc = s->eng.fetch_states(s->states.data(), s->cap_slots);
    if (rc != GCFG_OK) return rc;
    std::vector<Request *> still;
    for (Request *r : s->running) {
        auto it = s->sessions.find(r->session);
        if (it == s->sessions.end()) {
            push_result(s, r, GCFG_ERR_STATE, 0);
            delete r;
            continue;
        }
        Session *ss = it->second;
        const InstState &st = s->states[ss->slot];
        if (st.phase != gcfg::PH_DONE) {
            still.push_back(r);
            continue;
        }
        int status = st.err ? GCFG_ERR_INVALID : GCFG_OK;
        if (status == GCFG_OK && r->h_label) {
            if (hipMemcpyAsync(r->h_label, ss->d_label, (size_t)ss->W * ss->H * sizeof(int32_t),
                               hipMemcpyDeviceToHost, s->eng.stream()) != hipSuccess ||
                hipStreamSynchronize(s->eng.stream()) != hipSuccess)
                status = GCFG_ERR_HIP;
        }
        ss->busy = false;
        ss->last_used = now_ms();
        push_result(s, r, status, st.flow);
        delete r;
    }
    s->running.swap(still);
    return GCFG_OK;
}

static int admit(gcfg_server *s) {
    std::deque<Request *> keep;
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
                DtContext *dt = ss->dt;
                rc = dt->upload_image(r->rgb);
                if (rc == GCFG_OK) rc = dt->compute_swap();
                if (rc == GCFG_OK)
                    rc = s->eng.reframe(ss->slot, dt->pos_, dt->neg_, dt->up_, dt->down_, dt->left_, dt->right_, 0);
            }
            if (rc == GCFG_OK) rc = s->eng.rearm({ss->slot});
            if (rc != GCFG_OK) {
                push_result(s, r, rc, 0);
                delete r;
                continue;
            }
            r->t_admit = now_ms();
            ss->busy = true;
            ss->last_used = r->t_admit;
            s->running.push_back(r);
            continue;
        }
        const size_t need = session_cost(s->eng, r->W, r->H);
        if (need > s->budget || s->eng.tiles_for(r->W, r->H) > s->eng.cap_tiles()) {
            set_error("server: request larger than the HBM budget");
            push_result(s, r, GCFG_ERR_BUDGET, 0);
            delete r;
            continue;
        }
        if (blocked || !make_room(s, need) || (size_t)s->eng.tiles_for(r->W, r->H) > s->eng.tiles_free()) {
            blocked = true;
            keep.push_back(r);
            continue;
        }
        r->t_admit = now_ms();
        int rc = admit_image(s, r);
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
        s->running.push_back(r);
    }
    s->queue.swap(keep);
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
    // Arena sized for the budget if every byte went to tiles and their inputs.
    const size_t per_tile = Engine::bytes_per_tile() + (size_t)gcfg::TN * 32;
    const int tiles = (int)std::max<size_t>(1, hbm_budget_bytes / per_tile);
    int rc = s->eng.init(tiles, s->cap_slots, s->opt);
    if (rc != GCFG_OK) {
        delete s;
        return rc;
    }
    s->used = 0;
    s->states.resize(s->cap_slots);
    *out = s;
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
    s->queue.push_back(r);
    if (request_id) *request_id = r->id;
    return GCFG_OK;
}

int gcfg_server_poll(gcfg_server *s, int32_t wait, int32_t max_results, gcfg_result *results,
                     int32_t *n_results) {
    if (!s || (max_results > 0 && !results)) return GCFG_ERR_INVALID;
    int rc = GCFG_OK;
    for (;;) {
        rc = collect(s);
        if (rc != GCFG_OK) break;
        rc = admit(s);
        if (rc != GCFG_OK) break;
        if (!s->running.empty()) {
            rc = s->eng.pump(s->opt.steps_per_poll);
            if (rc != GCFG_OK) break;
        }
        if (!wait || !s->ready.empty() || (s->running.empty() && s->queue.empty())) break;
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
    if (in_flight) *in_flight = (int32_t)s->running.size();
    if (queued) *queued = (int32_t)s->queue.size();
    if (bytes_in_use) *bytes_in_use = s->used;
    return GCFG_OK;
}

int gcfg_server_destroy(gcfg_server *s) {
    if (!s) return GCFG_OK;
    s->eng.sync();
    for (Request *r : s->queue) delete r;
    for (Request *r : s->running) delete r;
    std::vector<Session *> all;
    for (auto &kv : s->sessions) all.push_back(kv.second);
    for (Session *ss : all) destroy_session(s, ss);
    delete s;
    return GCFG_OK;
}

}  // extern "C"
