// engine.hpp -- host side of the device step: packs systems into device arrays
// and drives the kernels.  A single-system context is a batch of one.
#pragma once

#include "../../include/ocm.h"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <tuple>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ocm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Owned copy of one ocm_force_desc.
struct HostForce {
    int kind = 0, group = 0, use_periodic = 0, n_terms = 0, n_categories = 0;
    std::vector<int> indices;
    std::vector<double> params, globals, particle_params;
    std::vector<int> category;       // empty: every bead is category 0
    std::vector<uint8_t> allow;      // empty: all category pairs allowed
    double cutoff = 0, switch_distance = 0;
    std::vector<std::pair<int, int>> exclusions;
};

struct HostSystem {
    int n = 0;
    std::vector<double> masses;
    bool periodic = false;
    double box[3] = {0, 0, 0};
    std::vector<HostForce> forces;
    ocm_integrator_desc integ{};
};

HostSystem copy_system(const ocm_system_desc& d, const ocm_integrator_desc& it);
int bonded_arity(int kind);
int bonded_nparams(int kind);
int pair_npp(int kind);
bool is_bonded(int kind);
bool is_pair(int kind);
void natural_spline_y2(const double* y, int n, double h, double* y2);

struct TimingSlot {
    double ms = 0;
    long long launches = 0;
};

class Engine {
public:
    Engine(std::vector<HostSystem> systems, const ocm_options& opt);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int n_systems() const { return static_cast<int>(sys_.size()); }
    int n_particles(int s) const { return sys_.at(s).n; }
    uint64_t seed(int s) const { return sys_.at(s).integ.seed; }

    void set_positions(int s, const double* xyz);
    void set_velocities(int s, const double* xyz);
    void set_velocities_to_temperature(int s, double T, uint64_t seed);
    // Advances every system by n steps.  Returns OCM_OK or OCM_ERR_NUMERIC
    // (some system produced NaN/overflow; see system_status).
    int step(long long n);
    // Captures the step graphs that step(n) will launch from the current
    // state (assuming no rollback), so that call spends no time on capture.
    void prepare(long long n);
    // Whether prepare(n) has work to do: the first force evaluation, or a
    // chunk of the coming n steps with no captured graph yet.  Host-side only.
    bool needs_prepare(long long n) const;
    // false: step runs a chunk that has no captured graph eagerly instead of
    // capturing one.  For callers whose other threads make HIP calls while
    // this engine steps: they capture through prepare, at a time they choose.
    void set_capture_in_step(bool on) { capture_in_step_ = on; }
    // Capped steepest descent until the RMS force component of every system
    // is <= tol or max_iter iterations ran.  Returns iterations done.
    int minimize(double tol, int max_iter);
    std::vector<double> potentials();   // per system, FP64 energy pass on the current list
    void get_state(int s, int what, uint32_t groups, ocm_state* out);
    long long neighbor_pairs(int s, int force_index, int32_t* out, long long cap);
    // The same pairs through the library's fixed device scratch: no device
    // allocation (rows are not widened; cap must hold every pair).  *count is
    // set before any size error is raised.
    void neighbor_pairs_noalloc(int s, int force_index, int32_t* out, long long cap, long long* count);
    int system_status(int s) const { return sys_status_.at(s); }
    // Continues system s's step count (and its noise stream) from steps_done.
    void set_steps_done(int s, long long steps_done);
    std::size_t device_bytes() const { return device_bytes_; }
    int n_beads_total() const { return d_.nb; }
    int bead_begin(int s) const { return bead_begin_.at(s); }
    hipStream_t stream() const { return stream_; }
    // Enqueues on the engine stream a copy of every bead position into dst
    // (device memory, 3*nb doubles: the x block, then y, then z).
    void enqueue_position_copy(double* dst);

    // 0 off; 1 HIP events around every launch (eager); 2 device-clock stamps
    // of the pair kernel inside the step graphs
    void set_timing(int mode);
    TimingSlot timing(int k) const { return timing_slots_.at(k); }
    void list_stats(long long* entries, long long* members, long long* rebuilds);
    void inner_stats(long long* entries, int* prune);
    void list_counts(long long out[4]);
    void launch_shape(int* team, int* u, int* grid, int* prune_team);
    int resident_threads() const { return d_.res_n > 0 ? d_.res_threads : 0; }

private:
    void pack();
    void pack_resident(const std::vector<BTerm>& terms, const std::vector<DevPairGroup>& groups,
                       const std::vector<float>& types_f);
    int step_resident(long long n);
    void alloc_device();
    void build_pair_blocks(int team);
    void ensure_forces();
    int pair_group_of(int s, int force_index) const;
    int rebuild_now();
    void carry_errors();
    void enqueue_forces(const ForceBuf& f, bool rebuild, uint32_t groups);
    // List schedule of one chunk: which steps rebuild the outer list and which
    // prune the inner one, from the list ages carried in from earlier chunks.
    struct ChunkPlan {
        int k = 0, interval = 0, prune = 0;
        int age0 = 0, page0 = 0, fresh = 0;   // ages at the chunk start; fresh: rebuild at its first step
        int age1 = 0, page1 = 0;              // ages after the chunk
        int rebuilds = 0;
        std::vector<uint8_t> rebuild, pruned;
        std::array<int, 7> key(int stamps) const { return {k, interval, prune, age0, page0, fresh, stamps}; }
    };
    ChunkPlan plan_chunk(int k) const;
    ChunkPlan plan_chunk_at(int k, int age, int page, int fresh) const;
    void enqueue_chunk(const ChunkPlan& p);
    hipGraphExec_t chunk_graph(const ChunkPlan& p, bool may_capture);
    int run_chunk(int k);
    void restore();
    void stale_lists() { fresh_ = 1; }   // the next chunk rebuilds and prunes at its first step
    void check_hip(hipError_t e, const char* what);
    void sync();
    void layout_rows();
    void grow_rows();
    void autotune_team();
    void drop_graphs();
    void release() noexcept;   // frees every device and pinned buffer, the graphs and the stream
    template <typename F> void timed(int slot, F&& f);
    void collect_timing();

    std::vector<HostSystem> sys_;
    ocm_options opt_;
    hipStream_t stream_ = nullptr;
    DevPtrs d_{};
    std::vector<void*> allocs_;
    template <typename T> T* dalloc(size_t n);
    template <typename T> void upload(T* dst, const std::vector<T>& src);

    // host mirrors
    std::vector<int> bead_begin_;
    std::vector<std::vector<int>> force_group_of_;   // [sys][force] -> pair group or -1
    std::vector<int> group_member_begin_, group_member_count_;
    std::vector<DevPairGroup> host_groups_;
    std::vector<long long> group_list_estimate_;
    std::vector<uint32_t> sys_group_mask_;
    std::vector<int> sys_status_;
    std::vector<int> err_carry_;     // flags of evaluations outside a chunk, reported with the next one
    std::vector<int64_t> sys_step_origin_;
    bool forces_dirty_ = true;
    bool capture_in_step_ = true;
    // resident steps leave the standard path's lists behind the positions:
    // the next evaluation that reads them rebuilds first
    bool lists_stale_ = false;
    static constexpr int kResChunk = 2000;   // resident steps per launch
    static constexpr long long kResListLdsMax = 8192;   // list entries per system kept in LDS
    static constexpr double kResSkin = 0.3;             // nm, at most, for resident lists
    // host copies of the members' beads and exclusion rows, held while packing
    std::vector<int> excl_bead_, excl_start_h_, excl_count_h_, excl_h_;
    int interval_ = 20;
    int chunk_ = 200;           // steps per chunk (chunk_steps, capped for large systems)
    int clean_chunks_ = 0;      // chunks since the interval was last shortened
    int recover_after_ = 4;     // clean chunks before the interval doubles back
    int prune_ = 0;             // steps between inner-list prunes (0: no inner list; 1: suspended)
    double prune_hm2_ = 1e300;  // (margin/2)^2 over all groups, the smallest
    int prune_clean_ = 0;       // the same backoff for the prune interval
    int prune_recover_after_ = 4;
    long long rebuilds_ = 0;
    bool timing_ = false;
    std::array<TimingSlot, 7> timing_slots_{};
    int timing_mode_ = 0;
    unsigned long long* kstamp_ = nullptr;   // 2 words per pair workgroup per step of a chunk
    long long kstamp_words_ = 0;
    int kstamp_grid_ = 0;
    void ensure_stamps();
    double wall_khz_ = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending_events_;
    std::vector<hipEvent_t> event_pool_;
    size_t event_next_ = 0;
    bool team_tuned_ = false;

    // snapshot buffers: the chunk-start state (positions, velocities, forces
    // and the two step slots), written by the chunk's first kernel
    double* bstate_ = nullptr;
    // pinned host words the chunk's last kernel writes: the batch flag word,
    // then one error word per system (cleared on the device as they are copied)
    int* status_host_ = nullptr;
    unsigned long long* stamp_host_ = nullptr;   // 2 words per step of a chunk (timing mode 2)

    // ages of the lists, carried from chunk to chunk: steps since the outer
    // list was built and since the inner list was pruned; fresh_ = 1 when the
    // lists no longer match the state (after a rollback or an evaluation
    // outside a chunk), so the next chunk rebuilds at its first step
    int age_ = 0, page_ = 0, fresh_ = 1;
    // captured chunks, least recently used first out when over kMaxGraphs
    struct GraphEntry {
        hipGraphExec_t exec;
        long long used;
    };
    std::map<std::array<int, 7>, GraphEntry> graphs_;
    long long graph_clock_ = 0;
    long long col_bytes_ = 0;
    std::size_t device_bytes_ = 0;
};

}  // namespace ocm
