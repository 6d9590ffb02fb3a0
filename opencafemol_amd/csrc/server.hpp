// server.hpp -- continuous-batching job server over the packed-batch engine.
//
// Jobs run in two lanes, each a batch with its own worker thread, engine and
// streams, so the two batches run side by side on the device: jobs small
// enough for resident stepping (up to kResMaxBeads beads; one workgroup each,
// whole chunks on chip) in one, the others in the other.  A small job then
// does not advance at the pace of a large one.  In each lane jobs wait in a
// FIFO queue until there is a free slot and their estimated device memory
// fits the budget (shared by the lanes); then the lane's batch is repacked
// with its active jobs and all of them advance together.  A worker steps its
// batch in stretches that end at the nearest frame or completion of any of
// its jobs, so a new arrival waits at most one stretch (capped at
// `quantum_steps`).  Frames are snapshotted on the device, copied to pinned
// host memory on a copy stream while the next stretch runs, and unpacked by a
// frame thread.
//
// A job's trajectory does not depend on what else is in the batch or where
// the stretches end: forces are sums in fixed point, the noise is keyed by the
// job's own seed, step and bead, and the integrator state after every step is
// the same bits however steps are grouped into calls.
#pragma once

#include "engine.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace ocm {

struct ServerJob {
    long long id = 0;
    int lane = 0;                      // 0: resident-sized, 1: the rest
    HostSystem sys;
    std::vector<double> x, v;          // state at `done` steps (host copy)
    bool v_set = false;                // false: Maxwell-Boltzmann at first admission
    long long nsteps = 0, frame_interval = 0, done = 0;
    std::size_t est_bytes = 0;
    int status = OCM_OK;
    bool finished = false;
    int frames_pending = 0;            // snapshots taken, not yet unpacked
    std::string message;
    double t_submit = 0, t_start = -1, t_end = -1;  // ms since server creation
};

struct ServerFrame {
    long long job = 0, step = 0;
    std::vector<double> x;             // n*3 nm
};

class Server {
public:
    explicit Server(const ocm_server_options& opt);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    long long submit(HostSystem sys, std::vector<double> x, std::vector<double> v, long long nsteps,
                     long long frame_interval);
    // 1: frame returned; 0: timed out.  Throws if cap is too small (frame kept).
    int next_frame(int timeout_ms, ServerFrame& out, long long cap_particles, long long* n_needed);
    // 1: job finished and removed (result moved out); 0: timed out.
    int wait(long long job, int timeout_ms, ServerJob& result);
    ocm_server_stats stats();

private:
    struct FrameSlot {
        double* dev = nullptr;         // 3*nb doubles: x block, y block, z block
        double* host = nullptr;        // pinned
        hipEvent_t copied = nullptr, staged = nullptr;
        bool busy = false;
        long long nb = 0;              // beads of the batch the slot serves
        long long cap = 0;             // beads the buffers hold
        std::vector<std::pair<long long, long long>> jobs;   // (job, step)
        std::vector<std::pair<int, int>> ranges;              // (bead begin, count)
    };

    // One batch and its worker.  Fields marked (mu_) are shared with the API
    // threads and the other lane under the server mutex; the rest belong to
    // the lane's worker (the frame thread reads `slots` entries it was handed).
    struct Lane {
        int id = 0;
        std::deque<std::shared_ptr<ServerJob>> queue;      // (mu_) waiting, FIFO
        std::vector<std::shared_ptr<ServerJob>> active;    // (mu_) admitted, in engine order once packed
        bool membership_changed = false;                    // (mu_)
        std::size_t engine_bytes = 0;                       // (mu_)
        std::unique_ptr<Engine> engine;
        std::vector<std::shared_ptr<ServerJob>> packed;     // engine system s -> job
        hipStream_t copy_stream = nullptr;
        std::vector<FrameSlot> slots;
        std::thread worker;
    };
    static constexpr int kLanes = 2;

    void loop(Lane& L);
    void frame_loop();
    bool admissible_locked(const Lane& L) const;
    bool admit_locked(Lane& L);
    void repack(Lane& L);
    void stage_frames(Lane& L, const std::vector<int>& systems);
    void release_slots(Lane& L);
    void drain_slots(Lane& L);
    void finish_job(const std::shared_ptr<ServerJob>& j, int status, const std::string& msg);
    void fail_lane(Lane& L, int code, const std::string& msg);
    double now_ms() const;
    std::size_t estimate_bytes(const HostSystem& s) const;
    int active_jobs_locked() const;
    // Device gate.  HIP work of the lanes and the frame thread runs inside it
    // (shared); graph capture, repacks and teardown hold it alone.  A HIP call
    // from another thread during a capture can invalidate the capture, and
    // hipFree synchronizes the whole device, so a repack freeing an engine
    // while the other lane kept submitting would wait for one of its
    // stretches per free.  Held alone, the gate waits for the other threads'
    // current HIP work once and keeps them out until it is released.
    void gate_enter(bool exclusive);
    void gate_exit(bool exclusive);

    ocm_server_options opt_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_, frame_cv_, slot_cv_;
    Lane lanes_[kLanes];
    std::map<long long, std::shared_ptr<ServerJob>> jobs_;     // all not yet collected
    std::deque<ServerFrame> frames_;
    std::vector<double> latencies_, waits_;
    bool stop_ = false;
    int workers_done_ = 0;
    long long next_id_ = 1;
    std::size_t active_bytes_ = 0;                             // over both lanes
    long long submitted_ = 0, completed_ = 0, failed_ = 0, repacks_ = 0, frames_made_ = 0;
    double simulated_ns_ = 0, busy_ms_ = 0, repack_ms_ = 0, first_submit_ = -1, last_end_ = -1;
    std::deque<std::pair<int, int>> staged_;                   // (lane, slot) awaiting the frame thread (mu_)
    int gate_users_ = 0, gate_waiting_ = 0;                    // (mu_)
    bool gate_held_ = false;                                   // (mu_)
    std::condition_variable gate_cv_;

    std::thread framer_;
};

}  // namespace ocm
