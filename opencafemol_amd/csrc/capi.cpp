// capi.cpp -- extern "C" entry points declared in include/ocm.h.
// Every call catches C++ exceptions and turns them into a status code plus a
// thread-local message (ocm_last_error), as the reference's main() turns its
// std::runtime_error into an exit status (src/OpenAICG2Plus.cpp:84-93).
#include "../../include/ocm.h"
#include "engine.hpp"
#include "server.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

struct ocm_context {
    ocm::Engine* eng;
};

struct ocm_server {
    ocm::Server* srv;
};

namespace {
thread_local std::string g_last_error;

template <typename F> int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const ocm::Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return OCM_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return OCM_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return OCM_ERR_INVALID;
    }
}

ocm::Engine& eng(ocm_context* c) {
    if (!c || !c->eng) throw ocm::Error(OCM_ERR_INVALID, "NULL context");
    return *c->eng;
}
void need(const void* p, const char* what) {
    if (!p) throw ocm::Error(OCM_ERR_INVALID, std::string(what) + " is NULL");
}
void check_sys(ocm::Engine& e, int s) {
    if (s < 0 || s >= e.n_systems()) throw ocm::Error(OCM_ERR_INVALID, "system index out of range");
}
ocm::Server& server(ocm_server* s) {
    if (!s || !s->srv) throw ocm::Error(OCM_ERR_INVALID, "NULL server");
    return *s->srv;
}
void check_n(ocm::Engine& e, int s, int32_t n) {
    if (n != e.n_particles(s)) throw ocm::Error(OCM_ERR_INVALID, "array length does not match the particle count");
}
}  // namespace

extern "C" {

void ocm_options_default(ocm_options* o) {
    if (!o) return;
    o->device = 0;
    o->rebuild_interval = 40;
    o->skin = 0.7;
    o->team_size = 0;
    o->use_graph = 1;
    o->chunk_steps = 200;
    o->prune_interval = 5;
    o->resident = -1;
    o->entries_per_lane = 0;
}

int ocm_batch_create(const ocm_system_desc* systems, const ocm_integrator_desc* integrators, int32_t n,
                     const ocm_options* options, ocm_context** out) {
    return guarded([&] {
        need(systems, "systems");
        need(integrators, "integrators");
        need(out, "out");
        *out = nullptr;
        if (n <= 0) throw ocm::Error(OCM_ERR_INVALID, "n_systems must be > 0");
        ocm_options opt;
        ocm_options_default(&opt);
        if (options) opt = *options;
        std::vector<ocm::HostSystem> hs;
        hs.reserve(n);
        for (int s = 0; s < n; ++s) hs.push_back(ocm::copy_system(systems[s], integrators[s]));
        auto* e = new ocm::Engine(std::move(hs), opt);
        *out = new ocm_context{e};
        return OCM_OK;
    });
}

int ocm_context_create(const ocm_system_desc* system, const ocm_integrator_desc* integrator,
                       const ocm_options* options, ocm_context** out) {
    return ocm_batch_create(system, integrator, 1, options, out);
}

int ocm_context_destroy(ocm_context* ctx) {
    return guarded([&] {
        if (!ctx) return OCM_OK;
        delete ctx->eng;
        delete ctx;
        return OCM_OK;
    });
}

int ocm_batch_set_positions(ocm_context* ctx, int32_t s, const double* xyz, int32_t n) {
    return guarded([&] {
        auto& e = eng(ctx);
        check_sys(e, s);
        need(xyz, "xyz");
        check_n(e, s, n);
        e.set_positions(s, xyz);
        return OCM_OK;
    });
}
int ocm_set_positions(ocm_context* ctx, const double* xyz, int32_t n) {
    return ocm_batch_set_positions(ctx, 0, xyz, n);
}

int ocm_batch_set_velocities(ocm_context* ctx, int32_t s, const double* xyz, int32_t n) {
    return guarded([&] {
        auto& e = eng(ctx);
        check_sys(e, s);
        need(xyz, "xyz");
        check_n(e, s, n);
        e.set_velocities(s, xyz);
        return OCM_OK;
    });
}
int ocm_set_velocities(ocm_context* ctx, const double* xyz, int32_t n) {
    return ocm_batch_set_velocities(ctx, 0, xyz, n);
}

int ocm_set_velocities_to_temperature(ocm_context* ctx, double T, uint64_t seed) {
    return guarded([&] {
        auto& e = eng(ctx);
        e.set_velocities_to_temperature(0, T, seed);
        return OCM_OK;
    });
}

int ocm_prepare_steps(ocm_context* ctx, int64_t nsteps) {
    return guarded([&] {
        eng(ctx).prepare(nsteps);
        return OCM_OK;
    });
}

int ocm_step(ocm_context* ctx, int64_t nsteps) {
    return guarded([&] {
        auto& e = eng(ctx);
        const int st = e.step(nsteps);
        if (st == OCM_ERR_NUMERIC) g_last_error = "NaN or fixed-point force overflow detected";
        return st;
    });
}

int ocm_minimize(ocm_context* ctx, double tolerance, int32_t max_iterations) {
    return guarded([&] {
        eng(ctx).minimize(tolerance, max_iterations);
        return OCM_OK;
    });
}

int ocm_batch_get_state(ocm_context* ctx, int32_t s, int32_t what, uint32_t groups, ocm_state* out) {
    return guarded([&] {
        auto& e = eng(ctx);
        check_sys(e, s);
        need(out, "out");
        e.get_state(s, what, groups, out);
        return OCM_OK;
    });
}
int ocm_get_state(ocm_context* ctx, int32_t what, uint32_t groups, ocm_state* out) {
    return ocm_batch_get_state(ctx, 0, what, groups, out);
}

int ocm_get_neighbor_pairs(ocm_context* ctx, int32_t force_index, int32_t* out_pairs, int64_t cap,
                           int64_t* count) {
    return guarded([&] {
        auto& e = eng(ctx);
        need(count, "count");
        if (cap > 0) need(out_pairs, "out_pairs");
        *count = e.neighbor_pairs(0, force_index, out_pairs, cap);
        return OCM_OK;
    });
}

int ocm_get_neighbor_pairs_noalloc(ocm_context* ctx, int32_t force_index, int32_t* out_pairs, int64_t cap,
                                   int64_t* count) {
    return guarded([&] {
        auto& e = eng(ctx);
        need(count, "count");
        if (cap < 0) throw ocm::Error(OCM_ERR_INVALID, "cap must not be negative");
        if (cap > 0) need(out_pairs, "out_pairs");
        long long n = 0;
        try {
            e.neighbor_pairs_noalloc(0, force_index, out_pairs, cap, &n);
        } catch (...) {
            *count = n;
            throw;
        }
        *count = n;
        return OCM_OK;
    });
}

int ocm_get_launch_shape(ocm_context* ctx, int32_t* team, int32_t* entries_per_lane, int32_t* grid,
                         int32_t* prune_team) {
    return guarded([&] {
        need(team, "team");
        need(entries_per_lane, "entries_per_lane");
        need(grid, "grid");
        need(prune_team, "prune_team");
        int t = 0, u = 0, g = 0, p = 0;
        eng(ctx).launch_shape(&t, &u, &g, &p);
        *team = t;
        *entries_per_lane = u;
        *grid = g;
        *prune_team = p;
        return OCM_OK;
    });
}

int ocm_get_resident(ocm_context* ctx, int32_t* threads) {
    return guarded([&] {
        need(threads, "threads");
        *threads = eng(ctx).resident_threads();
        return OCM_OK;
    });
}

int ocm_set_timing(ocm_context* ctx, int32_t enable) {
    return guarded([&] {
        eng(ctx).set_timing(enable);
        return OCM_OK;
    });
}

int ocm_get_timing(ocm_context* ctx, int32_t kernel, double* total_ms, int64_t* launches) {
    return guarded([&] {
        if (kernel < 0 || kernel > 6) throw ocm::Error(OCM_ERR_INVALID, "kernel id out of range");
        const auto t = eng(ctx).timing(kernel);
        if (total_ms) *total_ms = t.ms;
        if (launches) *launches = t.launches;
        return OCM_OK;
    });
}

int ocm_get_list_stats(ocm_context* ctx, int64_t* entries, int64_t* members, int64_t* rebuilds) {
    return guarded([&] {
        long long a = 0, b = 0, c = 0;
        eng(ctx).list_stats(&a, &b, &c);
        if (entries) *entries = a;
        if (members) *members = b;
        if (rebuilds) *rebuilds = c;
        return OCM_OK;
    });
}

int ocm_get_list_counts(ocm_context* ctx, int64_t* counts) {
    return guarded([&] {
        need(counts, "counts");
        long long c[4];
        eng(ctx).list_counts(c);
        for (int k = 0; k < 4; ++k) counts[k] = c[k];
        return OCM_OK;
    });
}

int ocm_get_inner_list_stats(ocm_context* ctx, int64_t* entries, int32_t* prune_interval) {
    return guarded([&] {
        long long a = 0;
        int p = 0;
        eng(ctx).inner_stats(&a, &p);
        if (entries) *entries = a;
        if (prune_interval) *prune_interval = p;
        return OCM_OK;
    });
}

int ocm_batch_system_status(ocm_context* ctx, int32_t s) {
    try {
        auto& e = eng(ctx);
        check_sys(e, s);
        return e.system_status(s);
    } catch (const ocm::Error& err) {
        g_last_error = err.what();
        return err.code;
    }
}

// ------------------------------------------------------------ checkpoints
// File layout (native byte order): "OCMCKPT" + version byte 1, int32 bead
// count, int32 0, int64 steps taken, uint64 integrator seed, then positions
// and velocities as n*3 doubles each.
namespace {
constexpr char kCkptMagic[8] = {'O', 'C', 'M', 'C', 'K', 'P', 'T', 1};

struct FileCloser {
    void operator()(std::FILE* f) const { if (f) std::fclose(f); }
};
}  // namespace

int ocm_batch_write_checkpoint(ocm_context* ctx, int32_t s, const char* path) {
    return guarded([&] {
        auto& e = eng(ctx);
        check_sys(e, s);
        need(path, "path");
        const int n = e.n_particles(s);
        std::vector<double> x(3 * static_cast<size_t>(n)), v(3 * static_cast<size_t>(n));
        ocm_state st{};
        st.positions = x.data();
        st.velocities = v.data();
        e.get_state(s, OCM_STATE_POSITIONS | OCM_STATE_VELOCITIES, 0xFFFFFFFFu, &st);
        std::unique_ptr<std::FILE, FileCloser> f(std::fopen(path, "wb"));
        if (!f) throw ocm::Error(OCM_ERR_INVALID, std::string("cannot open checkpoint file for writing: ") + path);
        const int32_t head[2] = {n, 0};
        const int64_t step = st.step;
        const uint64_t seed = e.seed(s);
        bool ok = std::fwrite(kCkptMagic, 1, 8, f.get()) == 8 && std::fwrite(head, 4, 2, f.get()) == 2 &&
                  std::fwrite(&step, 8, 1, f.get()) == 1 && std::fwrite(&seed, 8, 1, f.get()) == 1 &&
                  std::fwrite(x.data(), 8, x.size(), f.get()) == x.size() &&
                  std::fwrite(v.data(), 8, v.size(), f.get()) == v.size();
        ok = std::fflush(f.get()) == 0 && ok;
        if (!ok) throw ocm::Error(OCM_ERR_INVALID, std::string("writing checkpoint file failed: ") + path);
        return OCM_OK;
    });
}

int ocm_batch_read_checkpoint(ocm_context* ctx, int32_t s, const char* path) {
    return guarded([&] {
        auto& e = eng(ctx);
        check_sys(e, s);
        need(path, "path");
        std::unique_ptr<std::FILE, FileCloser> f(std::fopen(path, "rb"));
        if (!f) throw ocm::Error(OCM_ERR_INVALID, std::string("cannot open checkpoint file: ") + path);
        char magic[8];
        int32_t head[2];
        int64_t step = 0;
        uint64_t seed = 0;
        if (std::fread(magic, 1, 8, f.get()) != 8 || std::memcmp(magic, kCkptMagic, 8) != 0)
            throw ocm::Error(OCM_ERR_INVALID, std::string("not a checkpoint file of this version: ") + path);
        if (std::fread(head, 4, 2, f.get()) != 2 || std::fread(&step, 8, 1, f.get()) != 1 ||
            std::fread(&seed, 8, 1, f.get()) != 1)
            throw ocm::Error(OCM_ERR_INVALID, std::string("truncated checkpoint file: ") + path);
        const int n = e.n_particles(s);
        if (head[0] != n)
            throw ocm::Error(OCM_ERR_INVALID, "checkpoint holds " + std::to_string(head[0]) + " beads, the system " +
                                                  std::to_string(n));
        if (seed != e.seed(s))
            throw ocm::Error(OCM_ERR_INVALID, "checkpoint seed " + std::to_string(seed) +
                                                  " differs from the integrator's " + std::to_string(e.seed(s)) +
                                                  ": the noise stream would not continue");
        if (step < 0) throw ocm::Error(OCM_ERR_INVALID, "checkpoint step count is negative");
        std::vector<double> x(3 * static_cast<size_t>(n)), v(3 * static_cast<size_t>(n));
        if (std::fread(x.data(), 8, x.size(), f.get()) != x.size() ||
            std::fread(v.data(), 8, v.size(), f.get()) != v.size())
            throw ocm::Error(OCM_ERR_INVALID, std::string("truncated checkpoint file: ") + path);
        e.set_positions(s, x.data());
        e.set_velocities(s, v.data());
        e.set_steps_done(s, step);
        return OCM_OK;
    });
}

int ocm_write_checkpoint(ocm_context* ctx, const char* path) { return ocm_batch_write_checkpoint(ctx, 0, path); }
int ocm_read_checkpoint(ocm_context* ctx, const char* path) { return ocm_batch_read_checkpoint(ctx, 0, path); }

int32_t ocm_batch_size(ocm_context* ctx) {
    if (!ctx || !ctx->eng) return 0;
    return ctx->eng->n_systems();
}

void ocm_server_options_default(ocm_server_options* o) {
    if (!o) return;
    ocm_options_default(&o->engine);
    o->hbm_budget_bytes = 0;
    o->max_active_jobs = 256;
    o->max_queued_frames = 4096;
    o->quantum_steps = 2000;
}

int ocm_server_create(const ocm_server_options* options, ocm_server** out) {
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        ocm_server_options opt;
        ocm_server_options_default(&opt);
        if (options) opt = *options;
        auto* s = new ocm::Server(opt);
        *out = new ocm_server{s};
        return OCM_OK;
    });
}

int ocm_server_destroy(ocm_server* srv) {
    return guarded([&] {
        if (!srv) return OCM_OK;
        delete srv->srv;
        delete srv;
        return OCM_OK;
    });
}

int ocm_server_submit(ocm_server* srv, const ocm_system_desc* system, const ocm_integrator_desc* integrator,
                      const double* positions, const double* velocities, int32_t n, int64_t nsteps,
                      int64_t frame_interval, int64_t* job_id) {
    return guarded([&] {
        auto& S = server(srv);
        need(system, "system");
        need(integrator, "integrator");
        need(positions, "positions");
        need(job_id, "job_id");
        ocm::HostSystem hs = ocm::copy_system(*system, *integrator);
        if (n != hs.n) throw ocm::Error(OCM_ERR_INVALID, "array length does not match the particle count");
        std::vector<double> x(positions, positions + 3 * static_cast<size_t>(n));
        std::vector<double> v;
        if (velocities) v.assign(velocities, velocities + 3 * static_cast<size_t>(n));
        *job_id = S.submit(std::move(hs), std::move(x), std::move(v), nsteps, frame_interval);
        return OCM_OK;
    });
}

int ocm_server_next_frame(ocm_server* srv, int32_t timeout_ms, int64_t* job, int64_t* step, double* xyz,
                          int32_t cap, int32_t* n) {
    return guarded([&] {
        auto& S = server(srv);
        need(n, "n");
        if (cap > 0) need(xyz, "xyz");
        ocm::ServerFrame f;
        long long needed = 0;
        *n = 0;
        try {
            if (!S.next_frame(timeout_ms, f, cap, &needed)) {
                g_last_error = "no frame within the timeout";
                return OCM_ERR_TIMEOUT;
            }
        } catch (...) {
            *n = static_cast<int32_t>(needed);
            throw;
        }
        *n = static_cast<int32_t>(needed);
        if (job) *job = f.job;
        if (step) *step = f.step;
        std::copy(f.x.begin(), f.x.end(), xyz);
        return OCM_OK;
    });
}

int ocm_server_wait(ocm_server* srv, int64_t job, int32_t timeout_ms, double* positions, double* velocities,
                    int32_t cap, ocm_job_info* info) {
    return guarded([&] {
        auto& S = server(srv);
        ocm::ServerJob r;
        if (!S.wait(job, timeout_ms, r)) {
            g_last_error = "job still running";
            return OCM_ERR_TIMEOUT;
        }
        if (info) {
            info->status = r.status;
            info->n_particles = r.sys.n;
            info->steps_done = r.done;
            info->submit_ms = r.t_submit;
            info->start_ms = r.t_start;
            info->end_ms = r.t_end;
        }
        if ((positions || velocities) && cap < r.sys.n)
            throw ocm::Error(OCM_ERR_INVALID, "result buffer too small (the job is collected)");
        if (positions) std::copy(r.x.begin(), r.x.end(), positions);
        if (velocities) std::copy(r.v.begin(), r.v.end(), velocities);
        if (r.status != OCM_OK) g_last_error = r.message;
        return OCM_OK;
    });
}

int ocm_server_get_stats(ocm_server* srv, ocm_server_stats* out) {
    return guarded([&] {
        need(out, "out");
        *out = server(srv).stats();
        return OCM_OK;
    });
}

const char* ocm_last_error(void) { return g_last_error.c_str(); }

const char* ocm_version(void) { return "opencafemol_amd 0.1 (gfx950)"; }

}  // extern "C"
