// server.hpp -- continuous-batching job server over the packed-batch engine.
//
// One worker thread owns the device.  Jobs wait in a FIFO queue until the
// batch has a free slot and their estimated device memory fits the budget;
// then the batch is repacked with every active job and all of them advance
// together, one launch per step stage.  The worker steps the batch in
// stretches that end at the nearest frame or completion of any job, so a new
// arrival waits at most one stretch (capped at `quantum_steps`).  Frames are
// snapshotted on the device, copied to pinned host memory on a second stream
// while the next stretch runs, and unpacked by a frame thread.
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

    void loop();
    void frame_loop();
    bool admit_locked();
    void repack();
    void stage_frames(const std::vector<int>& systems);
    void release_slots();
    void drain_slots();
    void finish_job(const std::shared_ptr<ServerJob>& j, int status, const std::string& msg);
    double now_ms() const;
    std::size_t estimate_bytes(const HostSystem& s) const;

    ocm_server_options opt_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_, frame_cv_, slot_cv_;
    std::deque<std::shared_ptr<ServerJob>> queue_;
    std::vector<std::shared_ptr<ServerJob>> active_;          // admitted, in engine order once packed
    std::map<long long, std::shared_ptr<ServerJob>> jobs_;     // all not yet collected
    std::deque<ServerFrame> frames_;
    std::vector<double> latencies_, waits_;
    bool stop_ = false, worker_done_ = false, membership_changed_ = false;
    long long next_id_ = 1;
    std::size_t active_bytes_ = 0;
    long long submitted_ = 0, completed_ = 0, failed_ = 0, repacks_ = 0, frames_made_ = 0;
    double simulated_ns_ = 0, busy_ms_ = 0, repack_ms_ = 0, first_submit_ = -1, last_end_ = -1;
    std::size_t engine_bytes_ = 0;

    // worker-thread state
    std::unique_ptr<Engine> engine_;
    std::vector<std::shared_ptr<ServerJob>> packed_;   // engine system s -> job
    hipStream_t copy_stream_ = nullptr;
    std::vector<FrameSlot> slots_;
    std::deque<int> staged_;                           // slot indices awaiting the frame thread (under mu_)
    int team_ = 0;

    std::thread worker_, framer_;
};

}  // namespace ocm
