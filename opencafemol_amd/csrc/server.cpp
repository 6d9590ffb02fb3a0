// server.cpp -- continuous-batching job server (the request loop of a served
// job stream).  The reference runs one simulation per process
// (src/OpenAICG2Plus.cpp:12-93); this loop keeps many resident on one device.
#include "server.hpp"
#include "trace.hpp"

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

// Whether a job can step resident: at most kResMaxBeads beads, at most
// kResMaxGroups nonbonded forces and, when periodic, a box of at most kResWide
// nm.  The engine checks the rest (type rows) when it packs the lane; a batch
// with a member that fails there steps the standard path, with the same bits.
bool resident_sized(const HostSystem& s) {
    if (s.n > kResMaxBeads) return false;
    int pair_forces = 0;
    for (const HostForce& f : s.forces) pair_forces += is_pair(f.kind) ? 1 : 0;
    if (pair_forces > kResMaxGroups) return false;
    return !s.periodic || std::max({s.box[0], s.box[1], s.box[2]}) <= kResWide;
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
    for (int l = 0; l < kLanes; ++l) {
        lanes_[l].id = l;
        hip_ok(hipStreamCreateWithFlags(&lanes_[l].copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    }
    for (int l = 0; l < kLanes; ++l) {
        Lane* L = &lanes_[l];
        L->worker = std::thread([this, L] { loop(*L); });
    }
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
    for (Lane& L : lanes_)
        if (L.worker.joinable()) L.worker.join();
    if (framer_.joinable()) framer_.join();
    for (Lane& L : lanes_)
        if (L.copy_stream) (void)hipStreamDestroy(L.copy_stream);
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
    // a resident-sized system's own list holds every member pair (4 bytes each)
    if (s.n <= kResMaxBeads) bytes += 2.0 * s.n * s.n;
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
    // resident-sized jobs get the lane of their own (when resident stepping is on)
    j->lane = opt_.engine.resident != 0 && resident_sized(sys) ? 0 : 1;
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
        lanes_[j->lane].queue.push_back(j);
        ++submitted_;
    }
    work_cv_.notify_all();
    return id;
}

void Server::gate_enter(bool exclusive) {
    std::unique_lock<std::mutex> lk(mu_);
    if (exclusive) {
        ++gate_waiting_;
        gate_cv_.wait(lk, [&] { return gate_users_ == 0 && !gate_held_; });
        --gate_waiting_;
        gate_held_ = true;
    } else {
        // a waiting repack goes first
        gate_cv_.wait(lk, [&] { return !gate_held_ && gate_waiting_ == 0; });
        ++gate_users_;
    }
}

void Server::gate_exit(bool exclusive) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (exclusive) gate_held_ = false;
        else --gate_users_;
    }
    gate_cv_.notify_all();
}

int Server::active_jobs_locked() const {
    int n = 0;
    for (const Lane& L : lanes_) n += static_cast<int>(L.active.size());
    return n;
}

// FIFO admission within a lane: the head of its queue enters when a job slot
// is free and its estimate fits beside the active jobs of both lanes (or
// nothing is active); nothing overtakes it.
bool Server::admissible_locked(const Lane& L) const {
    if (L.queue.empty() || active_jobs_locked() >= opt_.max_active_jobs) return false;
    if (active_jobs_locked() == 0) return true;
    return static_cast<long long>(active_bytes_ + L.queue.front()->est_bytes) <= opt_.hbm_budget_bytes;
}

bool Server::admit_locked(Lane& L) {
    bool any = false;
    while (admissible_locked(L)) {
        auto j = L.queue.front();
        active_bytes_ += j->est_bytes;
        if (j->t_start < 0) {   // a job handed back after a failed repack keeps its first admission
            j->t_start = now_ms();
            waits_.push_back(j->t_start - j->t_submit);
        }
        L.active.push_back(j);
        L.queue.pop_front();
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
        Lane& L = lanes_[j->lane];
        auto it = std::find(L.active.begin(), L.active.end(), j);
        if (it != L.active.end()) {
            active_bytes_ -= j->est_bytes;
            L.active.erase(it);
            L.membership_changed = true;
        }
    }
    done_cv_.notify_all();
    work_cv_.notify_all();   // freed memory may admit the other lane's head
}

// Waits until every frame snapshot of the lane has been unpacked.
void Server::drain_slots(Lane& L) {
    std::unique_lock<std::mutex> lk(mu_);
    slot_cv_.wait(lk, [&] {
        for (const FrameSlot& s : L.slots)
            if (s.busy) return false;
        return true;
    });
}

void Server::release_slots(Lane& L) {
    drain_slots(L);
    for (FrameSlot& s : L.slots) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.staged) (void)hipEventDestroy(s.staged);
    }
    L.slots.clear();
}

void Server::repack(Lane& L) {
    const TraceRange trace("ocm.server.repack");
    const double t = now_ms();
    // carry the state of jobs that stay over from the old batch
    if (L.engine) {
        for (std::size_t s = 0; s < L.packed.size(); ++s) {
            auto& j = L.packed[s];
            if (j->finished) continue;
            ocm_state st;
            std::memset(&st, 0, sizeof(st));
            st.positions = j->x.data();
            st.velocities = j->v.data();
            L.engine->get_state(static_cast<int>(s), OCM_STATE_POSITIONS | OCM_STATE_VELOCITIES, 0u, &st);
            j->v_set = true;
        }
        drain_slots(L);   // snapshots in flight index the old batch's beads
        L.engine.reset();
        std::lock_guard<std::mutex> lk(mu_);
        L.engine_bytes = 0;
    }
    L.packed.clear();
    std::vector<std::shared_ptr<ServerJob>> members;
    {
        std::lock_guard<std::mutex> lk(mu_);
        members = L.active;
    }
    if (members.empty()) return;
    std::vector<HostSystem> systems;
    systems.reserve(members.size());
    for (auto& j : members) systems.push_back(j->sys);
    try {
        L.engine = std::make_unique<Engine>(std::move(systems), opt_.engine);
        L.engine->set_capture_in_step(false);
        for (std::size_t s = 0; s < members.size(); ++s) {
            auto& j = members[s];
            L.engine->set_positions(static_cast<int>(s), j->x.data());
            if (j->v_set) {
                L.engine->set_velocities(static_cast<int>(s), j->v.data());
            } else {
                L.engine->set_velocities_to_temperature(static_cast<int>(s), j->sys.integ.temperature,
                                                        j->sys.integ.seed);
            }
            L.engine->set_steps_done(static_cast<int>(s), j->done);
        }
        const long long nb = L.engine->n_beads_total();
        if (L.slots.empty() || L.slots[0].cap < nb) {
            // frame slots outlive repacks; they grow (by half again) only when a batch outgrows them
            release_slots(L);
            const long long cap = std::max(nb, nb + nb / 2);
            L.slots.resize(kSlots);
            for (FrameSlot& sl : L.slots) {
                sl.cap = cap;
                hip_ok(hipMalloc(&sl.dev, 3 * cap * sizeof(double)), "hipMalloc");
                hip_ok(hipHostMalloc(&sl.host, 3 * cap * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
                hip_ok(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming), "event");
                hip_ok(hipEventCreateWithFlags(&sl.staged, hipEventDisableTiming), "event");
            }
        }
        for (FrameSlot& sl : L.slots) sl.nb = nb;
    } catch (const Error& e) {
        L.engine.reset();
        release_slots(L);
        if (e.code == OCM_ERR_OOM && members.size() > 1) {
            // the estimate was short: hand the newest job back to the queue
            std::lock_guard<std::mutex> lk(mu_);
            auto j = L.active.back();
            L.active.pop_back();
            active_bytes_ -= j->est_bytes;
            L.queue.push_front(j);
            opt_.hbm_budget_bytes = static_cast<int64_t>(active_bytes_);
            L.membership_changed = true;
            return;
        }
        for (auto& j : members) finish_job(j, e.code, e.what());
        return;
    }
    L.packed = std::move(members);
    std::lock_guard<std::mutex> lk(mu_);
    L.engine_bytes = L.engine->device_bytes();
    ++repacks_;
    repack_ms_ += now_ms() - t;
}

// Snapshots every bead position on the engine stream (ordered before the next
// stretch), then copies the snapshot to pinned memory on the lane's copy stream.
void Server::stage_frames(Lane& L, const std::vector<int>& systems) {
    const TraceRange trace("ocm.server.frames");
    int slot = -1;
    {
        std::unique_lock<std::mutex> lk(mu_);
        slot_cv_.wait(lk, [&] {
            if (stop_) return true;
            if (static_cast<int>(frames_.size()) >= opt_.max_queued_frames) return false;
            for (int i = 0; i < kSlots; ++i)
                if (!L.slots[i].busy) { slot = i; return true; }
            return false;
        });
        if (slot < 0) return;   // shutting down
        L.slots[slot].busy = true;
    }
    FrameSlot& sl = L.slots[slot];
    sl.jobs.clear();
    sl.ranges.clear();
    for (int s : systems) {
        sl.jobs.emplace_back(L.packed[s]->id, L.packed[s]->done);
        sl.ranges.emplace_back(L.engine->bead_begin(s), L.engine->n_particles(s));
    }
    const std::size_t bytes = 3 * sl.nb * sizeof(double);
    gate_enter(false);
    try {
        L.engine->enqueue_position_copy(sl.dev);
        hip_ok(hipEventRecord(sl.copied, L.engine->stream()), "event");
        hip_ok(hipStreamWaitEvent(L.copy_stream, sl.copied, 0), "wait");
        hip_ok(hipMemcpyAsync(sl.host, sl.dev, bytes, hipMemcpyDeviceToHost, L.copy_stream), "frame copy");
        hip_ok(hipEventRecord(sl.staged, L.copy_stream), "event");
    } catch (...) {
        gate_exit(false);
        {
            std::lock_guard<std::mutex> lk(mu_);
            sl.busy = false;
        }
        slot_cv_.notify_all();
        throw;
    }
    gate_exit(false);
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (int s : systems) ++L.packed[s]->frames_pending;
        staged_.emplace_back(L.id, slot);
        frames_made_ += static_cast<long long>(systems.size());
    }
    frame_cv_.notify_all();
}

void Server::frame_loop() {
    (void)hipSetDevice(opt_.engine.device);
    while (true) {
        int lane, slot;
        {
            std::unique_lock<std::mutex> lk(mu_);
            frame_cv_.wait(lk, [&] { return !staged_.empty() || (stop_ && workers_done_ == kLanes); });
            if (staged_.empty()) return;
            lane = staged_.front().first;
            slot = staged_.front().second;
            staged_.pop_front();
        }
        FrameSlot& sl = lanes_[lane].slots[slot];
        gate_enter(false);
        (void)hipEventSynchronize(sl.staged);
        gate_exit(false);
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

// Ends every job of the lane's batch with an error and drops the batch.
void Server::fail_lane(Lane& L, int code, const std::string& msg) {
    std::vector<std::shared_ptr<ServerJob>> members;
    {
        std::lock_guard<std::mutex> lk(mu_);
        members = L.active;
    }
    for (auto& j : members)
        if (!j->finished) finish_job(j, code, msg);
    release_slots(L);
    L.engine.reset();
    L.packed.clear();
}

void Server::loop(Lane& L) {
    (void)hipSetDevice(opt_.engine.device);
    // Gate discipline: no thread holds the gate while it waits for another
    // thread (a free frame slot, the frame thread, the caller), so every
    // holder leaves in finite time.  HIP work runs in the shared gate; graph
    // capture, repacks and teardown in the exclusive one, after this lane's
    // snapshots have been unpacked (the frame thread needs the shared gate).
    auto fail = [&](int code, const std::string& msg) {
        drain_slots(L);
        gate_enter(true);
        fail_lane(L, code, msg);
        gate_exit(true);
    };
    auto code_of = [](const std::exception& e) {
        return dynamic_cast<const Error*>(&e) ? static_cast<const Error&>(e).code : OCM_ERR_HIP;
    };
    while (true) {
        bool need_repack;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // a queued head that cannot be admitted yet (the other lane holds
            // the memory) does not wake the worker: finish_job notifies when
            // memory frees
            work_cv_.wait(lk, [&] {
                return stop_ || admissible_locked(L) || !L.active.empty() || L.membership_changed;
            });
            if (stop_) break;
            if (admit_locked(L)) L.membership_changed = true;
            need_repack = L.membership_changed;
            L.membership_changed = false;
        }
        const double t = now_ms();
        if (need_repack) {
            drain_slots(L);   // snapshots in flight index the old batch's beads
            bool ok = true;
            std::string msg;
            gate_enter(true);
            try {
                repack(L);
            } catch (const std::exception& e) {
                ok = false;
                msg = e.what();
            }
            gate_exit(true);
            if (!ok) {
                fail(OCM_ERR_HIP, msg);
                continue;
            }
        }
        if (!L.engine) {
            std::lock_guard<std::mutex> lk(mu_);
            busy_ms_ += now_ms() - t;
            continue;
        }
        // stretch: up to the nearest frame or completion of any job
        long long n = opt_.quantum_steps;
        for (auto& j : L.packed) {
            long long next = j->nsteps;
            if (j->frame_interval > 0) next = std::min(next, (j->done / j->frame_interval + 1) * j->frame_interval);
            n = std::min(n, next - j->done);
        }
        std::vector<int> frame_sys, end_sys;
        bool failed = false;
        int fail_code = OCM_OK;
        std::string fail_msg;
        // graphs the stretch needs are captured with the other threads held
        // off (their HIP calls would invalidate a capture in progress); the
        // engine never captures inside step
        if (L.engine->needs_prepare(n)) {
            gate_enter(true);
            try {
                L.engine->prepare(n);
            } catch (const std::exception& e) {
                failed = true;
                fail_code = code_of(e);
                fail_msg = e.what();
            }
            gate_exit(true);
        }
        if (!failed) {
            gate_enter(false);
            try {
                const TraceRange trace(L.id == 0 ? "ocm.server.stretch.lane0" : "ocm.server.stretch.lane1");
                L.engine->step(n);
                for (std::size_t s = 0; s < L.packed.size(); ++s) {
                    auto& j = L.packed[s];
                    j->done += n;
                    const int st = L.engine->system_status(static_cast<int>(s));
                    if (st != OCM_OK) { end_sys.push_back(static_cast<int>(s)); continue; }
                    if (j->frame_interval > 0 && j->done % j->frame_interval == 0)
                        frame_sys.push_back(static_cast<int>(s));
                    if (j->done >= j->nsteps) end_sys.push_back(static_cast<int>(s));
                }
            } catch (const std::exception& e) {
                failed = true;
                fail_code = code_of(e);
                fail_msg = e.what();
            }
            gate_exit(false);
        }
        if (!failed && !frame_sys.empty()) {
            try {
                stage_frames(L, frame_sys);   // waits for a slot outside the gate
            } catch (const std::exception& e) {
                failed = true;
                fail_code = code_of(e);
                fail_msg = e.what();
            }
        }
        if (!failed && !end_sys.empty()) {
            gate_enter(false);
            try {
                for (int s : end_sys) {
                    auto& j = L.packed[s];
                    ocm_state st;
                    std::memset(&st, 0, sizeof(st));
                    st.positions = j->x.data();
                    st.velocities = j->v.data();
                    L.engine->get_state(s, OCM_STATE_POSITIONS | OCM_STATE_VELOCITIES, 0u, &st);
                    const int status = L.engine->system_status(s);
                    double dt_ps = j->sys.integ.dt;
                    {
                        std::lock_guard<std::mutex> lk(mu_);
                        simulated_ns_ += j->done * dt_ps * 1e-3;
                    }
                    finish_job(j, status, status == OCM_OK ? "" : "NaN or fixed-point force overflow detected");
                }
            } catch (const std::exception& e) {
                failed = true;
                fail_code = code_of(e);
                fail_msg = e.what();
            }
            gate_exit(false);
        }
        if (failed) fail(fail_code, fail_msg);
        std::lock_guard<std::mutex> lk(mu_);
        busy_ms_ += now_ms() - t;
    }
    // shutdown: let the frame thread finish the snapshots in flight, then free
    drain_slots(L);
    gate_enter(true);
    release_slots(L);
    L.engine.reset();
    L.packed.clear();
    gate_exit(true);
    {
        std::lock_guard<std::mutex> lk(mu_);
        ++workers_done_;
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
    st.active = active_jobs_locked();
    st.queued = static_cast<int64_t>(lanes_[0].queue.size() + lanes_[1].queue.size());
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
    st.device_bytes = static_cast<int64_t>(lanes_[0].engine_bytes + lanes_[1].engine_bytes);
    st.budget_bytes = opt_.hbm_budget_bytes;
    return st;
}

}  // namespace ocm
