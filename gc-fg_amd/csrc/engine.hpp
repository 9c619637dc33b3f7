// Engine shared by the plain (GC_*), batched and serving paths.
//
// Device layout ("arena"): every instance (one image's graph) owns a
// contiguous range of TILES; a tile is TW x TH nodes stored contiguously
// (tile-major), row-major inside the tile. Per node: residual capacities
// int4 {left, right, up, down} (16 B), excess int32 (4 B), distance label
// int32 (4 B) and a residual mask byte (1 B: bits 0-3 residual arc l/r/u/d,
// bit 4 excess > 0, bit 5 excess < 0), kept current by the push-relabel
// passes so a global relabel starts from the mask alone.
// Excess < 0 means remaining sink capacity (t-links are pre-saturated,
// GraphCutKernels.cu:196-197). Pixels outside the interior are inert nodes
// with zero capacities (GraphCutKernels.cu:122-146). With the wavefront
// kernels the capacities of a large batch are materialised lazily: the build
// writes excess, distance and mask only, and a tile's capacities are read from
// the inputs until its first push pass writes them (InstDesc::lazy,
// KParams::mat, ::inflow).
//
// Control: a scheduler kernel advances one state machine per instance
// (START -> global relabel -> push/relabel colour passes -> ... -> FIN ->
// DONE) and emits compacted tile queues; persistent push/relabel and
// global-relabel kernels consume the queues. A "step" is the three launches
// sched -> pr -> gr; the host replays chunks of steps (optionally as a
// hipGraph) and polls a pinned live-instance counter.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

#include "gcfg.h"

namespace gcfg {

constexpr int TW = 64;                  // tile width (one wavefront row)
constexpr int TH = 16;                  // tile height
constexpr int TN = TW * TH;             // nodes per tile
constexpr int NT = 256;                 // threads per workgroup (4 wavefronts)
constexpr int RPT = TN / NT;            // rows per thread (vertical strip)
constexpr int DINF = 0x3fffffff;        // distance of nodes that cannot reach the sink
constexpr int NUM_XCD = 8;              // MI355X: 8 accelerator complex dies, one L2 each
constexpr int INFLOW = 2 * TH + 2 * TW; // per-tile boundary slots of pending reverse-arc capacity
constexpr int LW = TW + 2;              // LDS row pitch with halo
constexpr int LH = TH + 2;

enum Phase : int32_t { PH_IDLE = 0, PH_START = 1, PH_GR = 2, PH_PR = 3, PH_FIN = 4, PH_DONE = 5 };
// OP_GR_INIT (workgroup kernels) derives masks from capacities and excess;
// OP_GR_REINIT (wavefront kernels) starts the BFS from the kept masks, which
// the build, the push passes and edits keep current.
enum GrOp : int32_t { OP_GR_INIT = 1, OP_GR_RELAX = 2, OP_FIN = 3, OP_GR_REINIT = 4 };
constexpr int MB_EPOS = 16, MB_ENEG = 32;  // mask bits: excess > 0, excess < 0
// InstState::cap of a reachability round (wavefront kernels): a zero-weight
// flood from the sinks, distances 0 (can reach a node with excess < 0) or
// DINF, run instead of an exact round to confirm that a capped schedule is done.
constexpr int CAP_FLOOD = 0;

struct TileInfo {
    int32_t inst;
    int32_t tx, ty;
    int32_t nbr[4];  // left, right, up, down tile ids (-1: none)
    int32_t pad;
};

struct InstDesc {
    int32_t W, H, tiles_x, tiles_y;
    int32_t tile0, ntiles, n_interior, penalty;
    const int32_t *pos, *neg, *up, *down, *left, *right;
    int32_t *label_out;  // device W*H
    // 8-neighbourhood only: capacities to (x-1,y-1), (x+1,y-1), (x-1,y+1), (x+1,y+1)
    const int32_t *upleft, *upright, *downleft, *downright;
    // 1: built with lazy capacities (wavefront kernels): tiles get their
    // capacities before their first push pass, and flow pushed into a
    // neighbour tile adds the reverse arc to its inflow slots. 0: the build
    // wrote every capacity and pushes update the neighbour's directly.
    int32_t lazy;
    int32_t pad_;
    // Packed labels (device, may be null): row Y of the image is tiles_x
    // 64-bit words at label_bits + Y * tiles_x, bit x of word k = label of
    // pixel (64 k + x, Y); bits past W are 0. Written with, or instead of,
    // label_out.
    uint64_t *label_bits;
    // Pinned host copy of label_bits (device-visible pointer, may be null):
    // the scheduler copies the instance's packed words there when the
    // instance finishes, while the other instances are still being solved.
    // Both pointers 16-byte aligned for 16-byte copies (else 8-byte ones).
    uint64_t *label_host;
};

struct InstState {
    int32_t phase;
    int32_t pr_iter;
    int32_t gr_rounds;
    int32_t steps;
    int64_t sum_pos;       // change of the interior sum of pos by edits since the build
    int32_t wcap;          // cap of the capped round a reachability round follows (its witnesses)
    int32_t pad2_;
    int64_t flow;
    int64_t pr_tiles;      // tile push-relabel passes scheduled
    int64_t gr_tiles;      // tile BFS sweeps scheduled (initial and relaxation)
    int32_t err;
    int32_t sweeps;        // colour-pass pairs between global relabels
    int64_t gr_init_tiles; // of which initial sweeps (INIT / REINIT)
    int32_t cap;           // distance cap of the current global relabel (DINF: exact, CAP_FLOOD: reachability)
    int32_t partial;       // 1: the current exact round recomputes only the blocks near pushes
    int64_t gr_exact_tiles;  // relabel items of exact rounds (those write labels)
    int64_t fin_items;       // FIN items (blocks left unreachable)
};

// Queue counters of one step. Steps alternate between two sets: the
// scheduler of step s appends to set s % 2 and clears set (s + 1) % 2, which
// the consumers of step s - 1 have finished with, so no kernel needs a
// grid-wide "last workgroup" ticket to reset them.
struct QCtl {
    int32_t qlen_pr, head_pr, qlen_gr, head_gr;
};

struct Ctl {
    QCtl q[2];
    int32_t n_live, pad0, pad1, pad2;
    int64_t pr_iters, gr_iters;  // inner iterations / relaxation rounds (profile mode only)
};

struct KParams {
    // arena
    int4 *caps;
    int32_t *exc, *dist, *act_flag, *dirty;
    // Per global-relabel block (at its anchor tile), from its last BFS sweep:
    // 0 every node unreachable, 1 all at or beyond the cap, 2 some node below it.
    int32_t *live;
    // Per block: within cap distance of a tile changed by a push pass since the
    // last relabel round (capped rounds recompute only these).
    int32_t *stale;
    // Per tile: [2t] interior sum of pos (build), [2t + 1] interior sum of
    // max(excess, 0) (FIN); the scheduler adds them up per instance.
    int64_t *tsum;
    uint8_t *mask;
    TileInfo *tiles;
    InstDesc *desc;
    InstState *state;
    Ctl *ctl;
    int32_t *queue_pr, *queue_gr;
    int32_t *bmap;  // build list position -> tile id (written by k_arm for k_buildw)
    // Lazily materialised residual capacities (wavefront kernels). mat[t] = 0:
    // tile t's capacities have not been written since the build and are those
    // of the inputs (border rule applied); 1: they are in caps. Either way,
    // inflow[t * INFLOW + k] holds reverse-arc capacity still to be added to a
    // boundary node by flow that neighbouring tiles pushed into it: slots
    // [0, TH) column 0 left arcs, [TH, 2 TH) column TW-1 right arcs,
    // [2 TH, 2 TH + TW) row 0 up arcs, [2 TH + TW, INFLOW) row TH-1 down arcs.
    int32_t *mat;
    int32_t *inflow;
    // 1: some instance has more than 2048 tiles, so the scheduler runs
    // 1024-thread workgroups (host-side launch choice).
    int32_t sched_wide;
    int32_t n_slots;
    int32_t inner_iters;
    int32_t variant;  // 0: workgroup-per-tile kernels (LDS), 1: wavefront-per-tile kernels (registers)
    int32_t gr_rounds;  // max relaxation rounds per tile per global-relabel sweep
    int32_t profile;    // count inner iterations / relaxation rounds (extra atomics)
    int32_t gbx, gby;   // global-relabel block, in tiles (one wavefront per tile)
    int32_t gr_cap;     // distance cap of intermediate global relabels (0: always exact)
    int32_t gr_cap_late;  // cap of the capped rounds after the first (>= gr_cap; 0 when gr_cap is)
    // 8-neighbourhood (nbhd == 8; workgroup kernels only): per node the
    // diagonal residual capacities int4 {up-left, up-right, down-left,
    // down-right} and a second mask byte (bits 0-3: those arcs residual); per
    // tile the diagonal neighbour tile ids in the same order (-1: none).
    int32_t nbhd;
    int4 *caps8;
    uint8_t *mask8;
    int4 *nbr8;
};

// Kernel launchers (solver.hip). With events, a launch records the
// kernel's own start and stop times in them (profile mode).
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
// Arms instances d_slots[0, n) and clears the iteration counters; with
// clear_tiles > 0 also writes the tile records and build map of the first
// clear_tiles build-list positions (d_slots[n, 2n) = tile offsets).
void launch_arm(const KParams &p, const int32_t *d_slots, int n, int warm, int sweeps, hipStream_t s,
                int clear_tiles);
void launch_build(const KParams &p, const int32_t *d_slots, int n, int total_tiles, hipStream_t s,
                  const LaunchEvents &ev);
void launch_step(const KParams &p, int grid_pr, int grid_gr, int parity, hipStream_t s);
void launch_sched(const KParams &p, int parity, hipStream_t s, const LaunchEvents &ev);
void launch_pr(const KParams &p, int grid, int parity, hipStream_t s, const LaunchEvents &ev);
void launch_gr(const KParams &p, int grid, int parity, hipStream_t s, const LaunchEvents &ev);
bool gr_block_supported(int bx, int by);
// Wavefronts of the push kernel that the kernel parameters select (variant,
// neighbourhood, inner iterations) resident on one CU at once.
int pr_resident_waves_per_cu(const KParams &p);
bool options_valid(const gcfg_options &o);  // sets the error message when not
void launch_edit(const KParams &p, int slot, const int32_t *d_strokes, int n_strokes, int32_t delta,
                 hipStream_t s);
void launch_read_state(const KParams &p, int slot, int32_t *d_caps, int32_t *d_exc, int32_t *d_dist,
                       int32_t *d_caps8, hipStream_t s);
void launch_materialize(const KParams &p, int slot, int ntiles, hipStream_t s);
// in1: pos, neg, up, down, left, right of the new inputs (up == null: penalty pen1).
void launch_reframe(const KParams &p, int slot, int ntiles, const int32_t *const *in1, int32_t pen1,
                    hipStream_t s);

// Orders a handle's internal stream against a caller stream: `before` makes
// the internal stream wait for everything already queued on the caller's,
// `after` makes the caller's wait for everything queued internally so far.
// With no caller stream both are no-ops.
struct StreamJoin {
    hipStream_t user = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int set(hipStream_t s);
    int before(hipStream_t own);
    int after(hipStream_t own);
    ~StreamJoin();
};

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define GCFG_HIP(call)                                          \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return ::gcfg::hip_fail(_e, #call); \
    } while (0)

// One engine = one arena on one device + one stream.
class Engine {
   public:
    Engine() = default;
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Capacity in tiles and instance slots; nbhd 4 or 8 (neighbourhood of
    // every instance of this engine).
    int init(int cap_tiles, int cap_slots, const gcfg_options &opt, int nbhd = 4);
    // Reserve a tile range and a slot for a W x H instance; returns slot or < 0.
    int add_instance(int W, int H, const int32_t *pos, const int32_t *neg, const int32_t *up,
                     const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                     int32_t *label_out);
    void remove_instance(int slot);
    // Packed label destination of an instance (see InstDesc::label_bits); null: none.
    void set_label_bits(int slot, uint64_t *bits) { h_desc_[slot].label_bits = bits; }
    // Host destination the packed labels are copied to on completion (see
    // InstDesc::label_host); null: the caller copies them.
    void set_label_host(int slot, uint64_t *host) { h_desc_[slot].label_host = host; }
    // Words of an instance's packed label map.
    static size_t label_words(int W, int H) { return (size_t)((W + TW - 1) / TW) * (size_t)H; }
    // diag: the four diagonal planes (up-left, up-right, down-left,
    // down-right) of an 8-neighbourhood instance with per-arc capacities.
    int set_inputs(int slot, const int32_t *pos, const int32_t *neg, const int32_t *up, const int32_t *down,
                   const int32_t *left, const int32_t *right, int32_t penalty, int32_t *label_out,
                   const int32_t *const *diag = nullptr);
    int neighborhood() const { return nbhd_; }
    // Upload descriptors, re-initialise state and build the residual graphs.
    int build(const std::vector<int> &slots);
    // Upload the descriptors that changed (e.g. a new label destination),
    // queued after the work already on the stream.
    int upload_desc();
    int upload_desc_slot(int slot);
    // The changed descriptors of `slots`, one staged copy per run of
    // consecutive slots.
    int upload_desc_slots(const std::vector<int> &slots);
    // Re-arm instances from their current residual graph (warm start).
    int rearm(const std::vector<int> &slots);
    // Run steps until every armed instance is done. polled_ctl() then holds
    // the counters as of the final completion poll.
    int run(int64_t max_steps);
    const Ctl &polled_ctl() const { return *pin_ctl_; }
    // States of slots [first, first + n) into pinned memory (one copy, one
    // synchronisation); pinned_states() points at them.
    int fetch_states_pinned(int first, int n);
    const InstState *pinned_states() const { return pin_states_; }
    // Launch up to `steps` steps without waiting; returns live count seen at
    // the last completed poll via live_out (non-blocking).
    int pump(int steps);
    int live_count(int32_t *out);
    int edit(int slot, const std::vector<int32_t> &strokes, int32_t delta);
    // Warm start of a solved 4-neighbourhood instance onto new inputs (same
    // shape): reparametrise its residual graph, then make the new inputs the
    // instance's own; follow with rearm().
    int reframe(int slot, const int32_t *pos, const int32_t *neg, const int32_t *up, const int32_t *down,
                const int32_t *left, const int32_t *right, int32_t penalty);
    // Write every tile's residual capacities of the listed instances to caps
    // (needed by code that reads caps directly: the workgroup kernels, the
    // video warm start).
    int materialize(const std::vector<int> &slots);
    // Residual capacities (W*H*4), excess (W*H), distance labels (W*H) and,
    // for the 8-neighbourhood, diagonal residual capacities (W*H*4) in image
    // layout; any output may be null.
    int read_state(int slot, int32_t *h_caps, int32_t *h_exc, int32_t *h_dist = nullptr,
                   int32_t *h_caps8 = nullptr);
    // Queue a copy of the states of slots [0, n) into pinned host memory and
    // record ev after it on the engine stream (non-blocking).
    int snapshot_states(InstState *h_dst, int n, hipEvent_t ev);
    int fetch_state(int slot, InstState *out);
    int fetch_states(InstState *out, int n);  // slots [0, n)
    int fetch_ctl(Ctl *out);
    int sync();

    hipStream_t stream() const { return stream_; }
    // Host bytes copied to the device in stream order through the engine's
    // pinned ring (no wait for queued kernels; a wrap of the ring waits for
    // the stream).
    int stage(void *d_dst, const void *h_src, size_t bytes);
    const InstDesc &desc(int slot) const { return h_desc_[slot]; }
    size_t tiles_free() const;
    int cap_tiles() const { return cap_tiles_; }
    int tiles_for(int W, int H) const { return ((W + TW - 1) / TW) * ((H + TH - 1) / TH); }
    // Instances hold at most 2^30 pixels (int32 node counts and 32-bit tile
    // arithmetic throughout).
    static bool shape_supported(int W, int H) { return W > 0 && H > 0 && (int64_t)W * H <= (int64_t)1 << 30; }
    static size_t bytes_per_tile(int nbhd = 4) {
        return (size_t)TN * (16 + 4 + 4 + 1) + 100 + INFLOW * 4 + (nbhd == 8 ? (size_t)TN * (16 + 1) + 16 : 0);
    }
    gcfg_options opt;
    int steps_run = 0;
    // Steps the previous run needed (set by the caller from the instance
    // states); the next run launches that many before its first poll, so a
    // repeated solve of similar inputs neither polls early nor replays idle
    // steps.
    int steps_hint = 0;
    // Event timing of the launches since the last reset (profile mode).
    double prof_sched_ms = 0, prof_pr_ms = 0, prof_gr_ms = 0, prof_build_ms = 0;
    int64_t prof_steps = 0;
    // Keeps a timed build that has not been collected yet: the plain graph
    // builds in one call and solves (resetting the profile) in the next.
    void reset_profile() {
        prof_sched_ms = prof_pr_ms = prof_gr_ms = prof_build_ms = 0;
        prof_steps = 0;
    }

   private:
    int alloc_tiles(int n);
    void free_tiles(int start, int n);
    int launch_steps(int n);
    int collect_profile();
    void apply_options();

    int dev_ = 0;
    int cu_count_ = 0;
    int nbhd_ = 4;
    bool any_lazy() const;    // some live instance still has capacities only in its inputs
    void update_sched_width();  // kp_.sched_wide from the live instances (drops stale step graphs)
    hipStream_t stream_ = nullptr;
    KParams kp_{};
    int cap_tiles_ = 0, cap_slots_ = 0;
    int grid_pr_ = 0, grid_gr_ = 0;
    std::vector<InstDesc> h_desc_;
    std::vector<InstDesc> dev_desc_;  // what the device descriptor table holds
    std::vector<int32_t> dev_list_;   // what d_slots_ holds after a build (empty: unknown)
    std::vector<char> slot_used_;
    std::vector<std::pair<int, int>> free_ranges_;  // (start, count)
    int32_t *d_slots_ = nullptr;
    Ctl *pin_ctl_ = nullptr;
    InstState *pin_states_ = nullptr;  // cap_slots_ states (fetch_states_pinned)
    int32_t *d_strokes_ = nullptr;
    int strokes_cap_ = 0;
    // Pinned ring for small host-to-device uploads (descriptors, slot lists,
    // strokes): each upload takes the next bytes of the ring; a wrap waits for
    // the stream, after which every copy queued from the ring has read its
    // bytes. Uploads therefore never wait for queued kernels.
    uint8_t *ring_ = nullptr;
    size_t ring_cap_ = 0, ring_off_ = 0;
    // Captured step chunks, one per chunk length and starting queue parity
    // (a chunk of any length continues the parity sequence, so a solve
    // replays exactly the steps it needs); dropped when the options change.
    struct StepGraph {
        int steps;
        int parity;
        int wide;  // scheduler launch width the chunk was captured with
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs_;
    void drop_graphs();
    int step_graph(int steps, int parity, hipGraphExec_t *out);
    hipGraphExec_t cached_graph(int steps, int parity) const;
    int parity_ = 0;  // parity of the next step to launch
    std::vector<void *> allocs_;
    std::vector<hipEvent_t> events_;  // 6 per profiled step: start/stop of sched, pr, gr
    int events_pending_ = 0;
    QCtl *pin_qtrace_ = nullptr;      // queue lengths per profiled step (verbose >= 2)
    hipEvent_t build_ev_[2] = {nullptr, nullptr};  // the build kernel's start / stop (profile mode)
    bool build_timed_ = false;
    int qtrace_cap_ = 0;
};

double now_ms();
int coordinate_search(gcfg_options start, int trials, const std::function<double(const gcfg_options &)> &solve,
                      gcfg_options *best, double *best_ms);

}  // namespace gcfg
