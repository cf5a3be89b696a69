// Host-side solve engine: problem slots resident in HBM, launch tables and
// the chunked CG loop shared by single solves, batches and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/suli_ps.h"
#include "ps_internal.h"

namespace sps {

// Device memory handed out in whole blocks and kept for reuse by size.
class Pool {
public:
    ~Pool();
    void *get(size_t bytes, std::string &err);
    void put(void *p, size_t bytes);
    void release_all();
    size_t held() const { return held_; }
    void reserve(size_t blocks) { free_.reserve(blocks); }

private:
    std::vector<std::pair<size_t, void *>> free_;   // (bytes, block), unordered
    size_t held_ = 0;   // bytes allocated from the device (free + in use)
};

// What a problem needs besides its arrays.
struct SolveOpts {
    bool mixed = false;
    int stop_rule = SULI_PS_STOP_REFERENCE;
    double tol = 1e-8;
    int max_iter = 10000;
    bool keep_null = false;
    int kz = 0;          // planes per 3-D tile (0: tuned or default for the size)
    int vx = 0;          // lane code (x cells per lane, kPipelined), FP32 only (0: tuned or 1)
    bool store_r = false;   // mixed: the FP32 steps store the residual
};

// One problem resident on the device.
struct Slot {
    bool used = false;
    bool active = false;       // in the launch tables
    bool finished = false;     // final_ seen, finish/unscale enqueued
    Geo g{};
    SolveOpts o;
    ProbDev pd{};
    void *block = nullptr;     // the problem's device memory
    size_t bytes = 0;
    bool own_block = true;     // taken from the engine's pool (else the caller's)
    double *prd1_out = nullptr;  // device prd1 (halo box) written at the end
    const uint8_t *solid = nullptr;  // solid-cell mask (device), if any
    int n0 = 0;
    int user = -1;             // caller's index for this problem
    CgScalars status{};        // last status read on the host
};

// Bytes of device memory a problem of this size and precision needs (for
// any tile depth kz >= kMinKz).
size_t problem_bytes(int nx, int ny, int nz, bool mixed);

constexpr int kMinKz = 2;

// Allocation guard.  While a NoAllocScope lives on a thread, every path of
// the engine that would allocate device or pinned host memory, create an
// event or instantiate a graph fails instead (SULI_PS_ENOMEM, with a
// message naming the site), so an entry point that promises not to allocate
// cannot do so by accident.
bool no_alloc_active();
class NoAllocScope {
public:
    NoAllocScope();
    ~NoAllocScope();
    NoAllocScope(const NoAllocScope &) = delete;
    NoAllocScope &operator=(const NoAllocScope &) = delete;

private:
    bool prev_;
};

class Engine {
public:
    Engine() = default;
    ~Engine();
    int init(hipStream_t stream, int max_slots);
    const std::string &error() const { return err_; }
    // Sizes the launch tables, the finish tables, the staging ring and the
    // host-side lists for `entries` table entries, so that solves whose
    // tables fit need no allocation (see NoAllocScope).  Never shrinks.
    int reserve(int entries);
    int reserved_entries() const { return reserved_; }
    // Tile shape a size would get without tuning now: the tuned or fixed
    // shape if one is recorded, else the default rule (kz, vx = 0).
    void cached_shape(int nx, int ny, int nz, bool mixed, int *kz, int *vx) const;

    // Admits a problem: allocates its slot, enqueues the setup kernels on the
    // stream and schedules its iteration 1 at the next global step.
    // prd1_out receives the solution (device, reference layout) when done.
    // block (optional): problem_bytes() of device memory the caller owns, used
    // instead of the engine's pool; wait_ev (optional): the setup kernels
    // wait for it (inputs copied on another stream).
    int admit(const suli_ps_grid &gr, const SetupIn &in, double *prd1_out, const SolveOpts &o, int user,
              int *slot_out, void *block = nullptr, hipEvent_t wait_ev = nullptr);
    // Runs chunks until every admitted problem has finished; each finished
    // problem's prd1_out holds its solution and status its final state.
    int run_all(bool timing);
    // One scheduling round: keeps two chunks of the active problems in
    // flight and reads the older one's state.  Appends the slots that
    // finished; their finish and un-scaling kernels are queued on the stream
    // (prd1_out is final once the stream has passed them).
    int pump(bool timing, std::vector<int> &finished);
    // Copies the final scalars of these slots to their status (synchronises).
    int read_final(const std::vector<int> &slots);
    int pending_chunks() const { return npend_; }
    int chunks_enqueued() const { return chunk_idx_; }
    // Device address of a slot's scalars (valid while the slot is held).
    const CgScalars *device_status(int slot) const { return d_status_ + slot; }
    hipStream_t stream() const { return s_; }
    // Frees a finished slot (its memory returns to the pool).
    void release(int slot);
    int active_count() const;
    const Slot &slot(int i) const { return slots_[i]; }
    size_t pool_bytes() const { return pool_.held(); }

    // SpMV timing samples of the last run_all (problem slot 0).
    double spmv_ms_mean() const { return spmv_ms_; }
    int spmv_samples() const { return spmv_n_; }
    // Times the plain SpMV of slot 0's operator (FP64 or FP32).
    int time_spmv(bool fp32, int reps, double *ms);
    // Replay captured HIP graphs of kRefreshGrain steps (default) or launch
    // every kernel from the host.  Timed runs always launch directly.
    void set_graphs(bool on) { graphs_ = on; }
    // Held while a graph is captured: other threads of the process that
    // allocate or free device memory (which synchronises the device and
    // breaks a capture in progress) take it too.
    void set_capture_mutex(std::mutex *m) { capture_mutex_ = m; }
    // Tile depth per problem size: with autotune on, a size's first
    // admission times the per-step kernel over candidate depths on a
    // scratch problem and keeps the fastest for every later problem of that
    // size (results then depend on the size only, as without tuning).
    void set_autotune(bool on) { autotune_ = on; }
    int tune(int nx, int ny, int nz, bool mixed, int *kz_out, int *vx_out, double *us_out);
    // Fixes the tile shape of a size (as if tuned).
    void set_tiles(int nx, int ny, int nz, bool mixed, int kz, int vx);
    // Fixed chunk length (steps between host polls; rounded up to the
    // refresh grain), or automatic when <= 0.
    void set_chunk(int steps) { chunk_ = steps; }
    bool graphs() const { return graphs_; }
    int graph_captures() const { return captures_; }
    // captured graphs launched since the last run_all began
    int graph_replays() const { return replays_; }

private:
    int fail(int code, const std::string &msg);
    int rebuild_tables();
    int enqueue_chunk(int K, bool timing);
    int enqueue_steps_direct(int nsteps, bool timing);
    int enqueue_group_graph();
    int capture_graphs();
    void retire_graphs();
    void reap_graphs(bool wait);
    int finish(const std::vector<int> &done);
    int next_chunk_len() const;
    bool refresh_may_be_due(int step) const;

    hipStream_t s_ = nullptr;
    std::string err_;
    Pool pool_;
    std::vector<Slot> slots_;
    ProbDev *d_probs_ = nullptr;       // device descriptors [max_slots]
    CgScalars *d_status_ = nullptr;    // device scalars [max_slots]
    CgScalars *h_status_ = nullptr;    // pinned [2][max_slots]
    CgScalars *h_final_ = nullptr;     // pinned [max_slots]
    hipEvent_t ev_status_[2] = {nullptr, nullptr};
    int status_step_[2] = {0, 0};      // global step when each status copy was queued
    int pend_[2] = {0, 0};             // status slots of chunks not yet read (FIFO)
    int pend_head_ = 0, npend_ = 0;
    int max_slots_ = 0;
    int step_ = 0;                     // global step of the last enqueued iteration
    int chunk_idx_ = 0;
    bool tables_dirty_ = true;
    void *d_fin_ = nullptr;            // tables of the problems being finished
    size_t fin_cap_ = 0;
    // launch tables (device, with pinned staging ring)
    struct Tables {
        std::vector<TileRef> t3_64, t1_64, t3_32, t1_32;
        std::vector<int> mixed_probs;
    } tab_;
    TileRef *d_tab_ = nullptr;
    int *d_probs_list_ = nullptr;
    size_t tab_cap_ = 0, probs_cap_ = 0;
    void *h_stage_[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev_stage_[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t stage_cap_ = 0;
    int stage_next_ = 0;
    // offsets of the four tables and the mixed-problem list in d_tab_/d_probs_list_
    int off_t3_64 = 0, off_t1_64 = 0, off_t3_32 = 0, off_t1_32 = 0;
    // the FP32 3-D table holds the problems of one step kind after another
    // (lane_kind): where each kind's entries start and how many
    int sub_t3_32[kLaneKinds] = {}, cnt_t3_32[kLaneKinds] = {};
    // launches the mixed step over every kind's entries; the first launch
    // carries the step counter update `next`.  concurrent (direct launches
    // only, never under graph capture): the kinds run on the engine's
    // stream and up to kSideStreams more, forked from it and joined back by
    // events, so that the launches of a step overlap; the problems of
    // different kinds share no memory, and each launch is the same
    int launch_steps_m(StepRef sr, int *ran, bool concurrent = false);
    static constexpr int kSideStreams = 3;   // with the engine's own: HIP's default 4 hardware queues per process
    hipStream_t side_[kSideStreams] = {};
    hipEvent_t fork_ev_ = nullptr, join_ev_[kSideStreams] = {};
    // graph replay: step counter on the device, two captured groups of
    // kRefreshGrain steps (without / with the residual replacement)
    int *d_step_ = nullptr;            // two counters, alternating by group parity
    bool dstep_valid_ = true;          // the counter of the next group holds step_
    bool graphs_ = true;
    long long table_cells_ = 0;   // cells of the problems in the launch tables
    bool autotune_ = false;
    struct Tuned {
        int kz, vx;   // vx: lane code
        double us;
    };
    std::map<std::tuple<int, int, int, bool>, Tuned> tuned_;
    int chunk_ = 0;
    std::mutex *capture_mutex_ = nullptr;
    bool graph_valid_ = false;
    hipGraphExec_t gx_[4] = {nullptr, nullptr, nullptr, nullptr};   // [parity][replacement]
    int captures_ = 0;
    int replays_ = 0;
    // replaced graphs, destroyed once the stream has passed this event
    std::vector<std::pair<hipGraphExec_t, hipEvent_t>> retired_;
    // timing
    std::vector<hipEvent_t> ev_spmv_;
    int *d_ran_ = nullptr;
    int ntimed_ = 0;
    double spmv_ms_ = 0.0;
    int spmv_n_ = 0;
    // host-side lists of run_all / pump / finish, kept with their capacity
    std::vector<int> run_slots_, run_finished_, done_, ran_;
    std::vector<TileRef> fin_t3_;
    int reserved_ = 0;
};

}  // namespace sps
