// server.cpp -- continuous-batching job server (the request loop of a served
// job stream).  The reference runs one simulation per process
// (src/OpenAICG2Plus.cpp:12-93); this loop keeps many resident on one device.
#include "server.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ocm {

namespace {

constexpr int kSlots = 4;   // frame snapshots in flight

void hip_ok(hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory) throw Error(OCM_ERR_OOM, std::string(what) + ": out of device memory");
    if (e != hipSuccess) throw Error(OCM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Nearest-rank percentile.
double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const std::size_t rank = static_cast<std::size_t>(std::ceil(q * static_cast<double>(v.size())));
    return v[std::min(v.size() - 1, rank == 0 ? 0 : rank - 1)];
}

}  // namespace

Server::Server(const ocm_server_options& opt) : opt_(opt), t0_(std::chrono::steady_clock::now()) {
    if (opt_.max_active_jobs <= 0) opt_.max_active_jobs = 256;
    if (opt_.max_queued_frames <= 0) opt_.max_queued_frames = 4096;
    if (opt_.quantum_steps <= 0) opt_.quantum_steps = 2000;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Error(OCM_ERR_NO_DEVICE, "no HIP device found");
    if (opt_.engine.device < 0 || opt_.engine.device >= ndev) throw Error(OCM_ERR_NO_DEVICE, "device ordinal out of range");
    hip_ok(hipSetDevice(opt_.engine.device), "hipSetDevice");
    hipDeviceProp_t prop;
    hip_ok(hipGetDeviceProperties(&prop, opt_.engine.device), "hipGetDeviceProperties");
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
        throw Error(OCM_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
    if (opt_.hbm_budget_bytes <= 0) {
        std::size_t free_b = 0, total_b = 0;
        hip_ok(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        opt_.hbm_budget_bytes = static_cast<int64_t>(0.8 * static_cast<double>(free_b));
    }
    hip_ok(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking), "hipStreamCreate");
    worker_ = std::thread([this] { loop(); });
    framer_ = std::thread([this] { frame_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    slot_cv_.notify_all();
    frame_cv_.notify_all();
    done_cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    if (framer_.joinable()) framer_.join();
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
}

double Server::now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
}

// Device bytes a system takes when packed: per-bead state, snapshots and
// records, bonded terms, and a neighbor list sized as the engine sizes it.
std::size_t Server::estimate_bytes(const HostSystem& s) const {
    double bytes = 512.0 * s.n;
    const double skin = opt_.engine.skin;
    for (const HostForce& f : s.forces) {
        if (is_bonded(f.kind)) {
            bytes += 72.0 * f.n_terms;
            continue;
        }
        double per_bead;
        if (s.periodic) {
            const double rl = f.cutoff + skin;
            const double rho = s.n / (s.box[0] * s.box[1] * s.box[2]);
            per_bead = rho * 4.18879 * rl * rl * rl * 3.0 + 8.0;
            bytes += 16.0 * (s.box[0] / rl) * (s.box[1] / rl) * (s.box[2] / rl);
        } else {
            per_bead = std::min(s.n, 256);
        }
        const double list_bytes = opt_.engine.prune_interval > 1 ? 8.0 : 4.0;   // outer (+ inner) list
        bytes += s.n * (list_bytes * per_bead + 48.0) + 8.0 * f.exclusions.size();
    }
    return static_cast<std::size_t>(bytes) + (1u << 20);
}

long long Server::submit(HostSystem sys, std::vector<double> x, std::vector<double> v, long long nsteps,
                         long long frame_interval) {
    if (nsteps < 1) throw Error(OCM_ERR_INVALID, "nsteps must be >= 1");
    if (frame_interval < 0) throw Error(OCM_ERR_INVALID, "frame_interval must be >= 0");
    if (static_cast<long long>(x.size()) != 3LL * sys.n) throw Error(OCM_ERR_INVALID, "positions length mismatch");
    for (double c : x)
        if (!std::isfinite(c)) throw Error(OCM_ERR_INVALID, "position is not finite");
    auto j = std::make_shared<ServerJob>();
    j->est_bytes = estimate_bytes(sys);
    j->sys = std::move(sys);
    j->x = std::move(x);
    j->v_set = !v.empty();
    j->v = v.empty() ? std::vector<double>(j->x.size(), 0.0) : std::move(v);
    j->nsteps = nsteps;
    j->frame_interval = frame_interval;
    long long id;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) throw Error(OCM_ERR_STATE, "server is shutting down");
        if (static_cast<long long>(j->est_bytes) > opt_.hbm_budget_bytes)
            throw Error(OCM_ERR_OOM, "job needs about " + std::to_string(j->est_bytes) +
                                         " device bytes, more than the server budget");
        id = next_id_++;
        j->id = id;
        j->t_submit = now_ms();
        if (first_submit_ < 0) first_submit_ = j->t_submit;
        jobs_[id] = j;
        queue_.push_back(j);
        ++submitted_;
    }
    work_cv_.notify_all();
    return id;
}

// FIFO admission: the head of the queue enters when a slot is free and its
// estimate fits beside the active jobs; nothing overtakes it.
bool Server::admit_locked() {
    bool any = false;
    while (!queue_.empty() && static_cast<int>(active_.size()) < opt_.max_active_jobs) {
        auto& j = queue_.front();
        if (!active_.empty() && static_cast<long long>(active_bytes_ + j->est_bytes) > opt_.hbm_budget_bytes) break;
        active_bytes_ += j->est_bytes;
        if (j->t_start < 0) {   // a job handed back after a failed repack keeps its first admission
            j->t_start = now_ms();
            waits_.push_back(j->t_start - j->t_submit);
        }
        active_.push_back(j);
        queue_.pop_front();
        any = true;
    }
    return any;
}

void Server::finish_job(const std::shared_ptr<ServerJob>& j, int status, const std::string& msg) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        j->status = status;
        j->message = msg;
        j->finished = true;
        j->t_end = now_ms();
        last_end_ = j->t_end;
        if (status == OCM_OK) {
            ++completed_;
            latencies_.push_back(j->t_end - j->t_submit);
        } else {
            ++failed_;
        }
        auto it = std::find(active_.begin(), active_.end(), j);
        if (it != active_.end()) {
            active_bytes_ -= j->est_bytes;
            active_.erase(it);
            membership_changed_ = true;
        }
    }
    done_cv_.notify_all();
}

// Waits until every frame snapshot has been unpacked.
void Server::drain_slots() {
    std::unique_lock<std::mutex> lk(mu_);
    slot_cv_.wait(lk, [&] {
        for (const FrameSlot& s : slots_)
            if (s.busy) return false;
        return true;
    });
}

void Server::release_slots() {
    drain_slots();
    for (FrameSlot& s : slots_) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.staged) (void)hipEventDestroy(s.staged);
    }
    slots_.clear();
}

void Server::repack() {
    const double t = now_ms();
    // carry the state of jobs that stay over from the old batch
    if (engine_) {
        for (std::size_t s = 0; s < packed_.size(); ++s) {
            auto& j = packed_[s];
            if (j->finished) continue;
            ocm_state st;
            std::memset(&st, 0, sizeof(st));
            st.positions = j->x.data();
            st.velocities = j->v.data();
            engine_->get_state(static_cast<int>(s), OCM_STATE_POSITIONS | OCM_STATE_VELOCITIES, 0u, &st);
            j->v_set = true;
        }
        drain_slots();   // snapshots in flight index the old batch's beads
        engine_.reset();
        engine_bytes_ = 0;
    }
    packed_.clear();
    std::vector<std::shared_ptr<ServerJob>> members;
    {
        std::lock_guard<std::mutex> lk(mu_);
        members = active_;
    }
    if (members.empty()) return;
    std::vector<HostSystem> systems;
    systems.reserve(members.size());
    for (auto& j : members) systems.push_back(j->sys);
    try {
        engine_ = std::make_unique<Engine>(std::move(systems), opt_.engine);
        for (std::size_t s = 0; s < members.size(); ++s) {
            auto& j = members[s];
            engine_->set_positions(static_cast<int>(s), j->x.data());
            if (j->v_set) {
                engine_->set_velocities(static_cast<int>(s), j->v.data());
            } else {
                engine_->set_velocities_to_temperature(static_cast<int>(s), j->sys.integ.temperature, j->sys.integ.seed);
            }
            engine_->set_steps_done(static_cast<int>(s), j->done);
        }
        const long long nb = engine_->n_beads_total();
        if (slots_.empty() || slots_[0].cap < nb) {
            // frame slots outlive repacks; they grow (by half again) only when a batch outgrows them
            release_slots();
            const long long cap = std::max(nb, nb + nb / 2);
            slots_.resize(kSlots);
            for (FrameSlot& sl : slots_) {
                sl.cap = cap;
                hip_ok(hipMalloc(&sl.dev, 3 * cap * sizeof(double)), "hipMalloc");
                hip_ok(hipHostMalloc(&sl.host, 3 * cap * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
                hip_ok(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming), "event");
                hip_ok(hipEventCreateWithFlags(&sl.staged, hipEventDisableTiming), "event");
            }
        }
        for (FrameSlot& sl : slots_) sl.nb = nb;
    } catch (const Error& e) {
        engine_.reset();
        release_slots();
        if (e.code == OCM_ERR_OOM && members.size() > 1) {
            // the estimate was short: hand the newest job back to the queue
            std::lock_guard<std::mutex> lk(mu_);
            auto j = active_.back();
            active_.pop_back();
            active_bytes_ -= j->est_bytes;
            queue_.push_front(j);
            opt_.hbm_budget_bytes = static_cast<int64_t>(active_bytes_);
            membership_changed_ = true;
            return;
        }
        for (auto& j : members) finish_job(j, e.code, e.what());
        return;
    }
    packed_ = std::move(members);
    engine_bytes_ = engine_->device_bytes();
    std::lock_guard<std::mutex> lk(mu_);
    ++repacks_;
    repack_ms_ += now_ms() - t;
}

// Snapshots every bead position on the engine stream (ordered before the next
// stretch), then copies the snapshot to pinned memory on the copy stream.
void Server::stage_frames(const std::vector<int>& systems) {
    int slot = -1;
    {
        std::unique_lock<std::mutex> lk(mu_);
        slot_cv_.wait(lk, [&] {
            if (stop_) return true;
            if (static_cast<int>(frames_.size()) >= opt_.max_queued_frames) return false;
            for (int i = 0; i < kSlots; ++i)
                if (!slots_[i].busy) { slot = i; return true; }
            return false;
        });
        if (slot < 0) return;   // shutting down
        slots_[slot].busy = true;
    }
    FrameSlot& sl = slots_[slot];
    sl.jobs.clear();
    sl.ranges.clear();
    for (int s : systems) {
        sl.jobs.emplace_back(packed_[s]->id, packed_[s]->done);
        sl.ranges.emplace_back(engine_->bead_begin(s), engine_->n_particles(s));
    }
    const std::size_t bytes = 3 * sl.nb * sizeof(double);
    engine_->enqueue_position_copy(sl.dev);
    hip_ok(hipEventRecord(sl.copied, engine_->stream()), "event");
    hip_ok(hipStreamWaitEvent(copy_stream_, sl.copied, 0), "wait");
    hip_ok(hipMemcpyAsync(sl.host, sl.dev, bytes, hipMemcpyDeviceToHost, copy_stream_), "frame copy");
    hip_ok(hipEventRecord(sl.staged, copy_stream_), "event");
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (int s : systems) ++packed_[s]->frames_pending;
        staged_.push_back(slot);
        frames_made_ += static_cast<long long>(systems.size());
    }
    frame_cv_.notify_all();
}

void Server::frame_loop() {
    (void)hipSetDevice(opt_.engine.device);
    while (true) {
        int slot;
        {
            std::unique_lock<std::mutex> lk(mu_);
            frame_cv_.wait(lk, [&] { return !staged_.empty() || (stop_ && worker_done_); });
            if (staged_.empty()) return;
            slot = staged_.front();
            staged_.pop_front();
        }
        FrameSlot& sl = slots_[slot];
        (void)hipEventSynchronize(sl.staged);
        std::vector<ServerFrame> out(sl.jobs.size());
        const double* hx = sl.host;
        const double* hy = sl.host + sl.nb;
        const double* hz = sl.host + 2 * sl.nb;
        for (std::size_t k = 0; k < sl.jobs.size(); ++k) {
            out[k].job = sl.jobs[k].first;
            out[k].step = sl.jobs[k].second;
            const int b0 = sl.ranges[k].first, n = sl.ranges[k].second;
            out[k].x.resize(3 * static_cast<std::size_t>(n));
            for (int i = 0; i < n; ++i) {
                out[k].x[3 * i] = hx[b0 + i];
                out[k].x[3 * i + 1] = hy[b0 + i];
                out[k].x[3 * i + 2] = hz[b0 + i];
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto& f : out) {
                auto it = jobs_.find(f.job);
                if (it != jobs_.end()) --it->second->frames_pending;
                frames_.push_back(std::move(f));
            }
            sl.busy = false;
        }
        slot_cv_.notify_all();
        done_cv_.notify_all();
    }
}

void Server::loop() {
    (void)hipSetDevice(opt_.engine.device);
    while (true) {
        bool need_repack;
        {
            std::unique_lock<std::mutex> lk(mu_);
            work_cv_.wait(lk, [&] { return stop_ || !queue_.empty() || !active_.empty() || membership_changed_; });
            if (stop_) break;
            if (admit_locked()) membership_changed_ = true;
            need_repack = membership_changed_;
            membership_changed_ = false;
        }
        const double t = now_ms();
        if (need_repack) {
            try {
                repack();
            } catch (const std::exception& e) {
                release_slots();
                engine_.reset();
                std::vector<std::shared_ptr<ServerJob>> members;
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    members = active_;
                }
                for (auto& j : members) finish_job(j, OCM_ERR_HIP, e.what());
                packed_.clear();
                continue;
            }
        }
        if (!engine_) {
            std::lock_guard<std::mutex> lk(mu_);
            busy_ms_ += now_ms() - t;
            continue;
        }
        // stretch: up to the nearest frame or completion of any job
        long long n = opt_.quantum_steps;
        for (auto& j : packed_) {
            long long next = j->nsteps;
            if (j->frame_interval > 0) next = std::min(next, (j->done / j->frame_interval + 1) * j->frame_interval);
            n = std::min(n, next - j->done);
        }
        std::vector<int> frame_sys, end_sys;
        try {
            engine_->step(n);
            for (std::size_t s = 0; s < packed_.size(); ++s) {
                auto& j = packed_[s];
                j->done += n;
                const int st = engine_->system_status(static_cast<int>(s));
                if (st != OCM_OK) { end_sys.push_back(static_cast<int>(s)); continue; }
                if (j->frame_interval > 0 && j->done % j->frame_interval == 0) frame_sys.push_back(static_cast<int>(s));
                if (j->done >= j->nsteps) end_sys.push_back(static_cast<int>(s));
            }
            if (!frame_sys.empty()) stage_frames(frame_sys);
            for (int s : end_sys) {
                auto& j = packed_[s];
                ocm_state st;
                std::memset(&st, 0, sizeof(st));
                st.positions = j->x.data();
                st.velocities = j->v.data();
                engine_->get_state(s, OCM_STATE_POSITIONS | OCM_STATE_VELOCITIES, 0u, &st);
                const int status = engine_->system_status(s);
                double dt_ps = j->sys.integ.dt;
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    simulated_ns_ += j->done * dt_ps * 1e-3;
                }
                finish_job(j, status, status == OCM_OK ? "" : "NaN or fixed-point force overflow detected");
            }
        } catch (const std::exception& e) {
            const int code = dynamic_cast<const Error*>(&e) ? static_cast<const Error&>(e).code : OCM_ERR_HIP;
            for (auto& j : packed_)
                if (!j->finished) finish_job(j, code, e.what());
            release_slots();
            engine_.reset();
            packed_.clear();
        }
        std::lock_guard<std::mutex> lk(mu_);
        busy_ms_ += now_ms() - t;
    }
    // shutdown: let the frame thread finish the snapshots in flight, then free
    release_slots();
    engine_.reset();
    packed_.clear();
    {
        std::lock_guard<std::mutex> lk(mu_);
        worker_done_ = true;
    }
    frame_cv_.notify_all();
}

int Server::next_frame(int timeout_ms, ServerFrame& out, long long cap_particles, long long* n_needed) {
    std::unique_lock<std::mutex> lk(mu_);
    auto ready = [&] { return !frames_.empty(); };
    if (timeout_ms < 0) done_cv_.wait(lk, ready);
    else if (!done_cv_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return 0;
    const long long n = static_cast<long long>(frames_.front().x.size() / 3);
    if (n_needed) *n_needed = n;
    if (n > cap_particles) throw Error(OCM_ERR_INVALID, "frame buffer too small; the frame stays queued");
    out = std::move(frames_.front());
    frames_.pop_front();
    lk.unlock();
    slot_cv_.notify_all();
    return 1;
}

int Server::wait(long long job, int timeout_ms, ServerJob& result) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw Error(OCM_ERR_INVALID, "unknown job id (never submitted or already collected)");
    std::shared_ptr<ServerJob> j = it->second;
    auto ready = [&] { return j->finished && j->frames_pending == 0; };
    if (timeout_ms < 0) done_cv_.wait(lk, ready);
    else if (!done_cv_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return 0;
    jobs_.erase(job);
    result = std::move(*j);
    return 1;
}

ocm_server_stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ocm_server_stats st;
    std::memset(&st, 0, sizeof(st));
    st.submitted = submitted_;
    st.completed = completed_;
    st.failed = failed_;
    st.active = static_cast<int64_t>(active_.size());
    st.queued = static_cast<int64_t>(queue_.size());
    st.repacks = repacks_;
    st.frames = frames_made_;
    st.latency_p50_ms = percentile(latencies_, 0.50);
    st.latency_p99_ms = percentile(latencies_, 0.99);
    st.wait_p50_ms = percentile(waits_, 0.50);
    st.wait_p99_ms = percentile(waits_, 0.99);
    st.simulated_ns = simulated_ns_;
    st.busy_ms = busy_ms_;
    st.repack_ms = repack_ms_;
    st.makespan_ms = (first_submit_ >= 0 && last_end_ >= 0) ? last_end_ - first_submit_ : 0.0;
    st.device_bytes = static_cast<int64_t>(engine_bytes_);
    st.budget_bytes = opt_.hbm_budget_bytes;
    return st;
}

}  // namespace ocm
