// capi.cpp -- extern "C" entry points declared in include/ocm.h.
// Every call catches C++ exceptions and turns them into a status code plus a
// thread-local message (ocm_last_error), as the reference's main() turns its
// std::runtime_error into an exit status (src/OpenAICG2Plus.cpp:84-93).
#include "../../include/ocm.h"
#include "engine.hpp"

#include <new>
#include <string>

struct ocm_context {
    ocm::Engine* eng;
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
void check_n(ocm::Engine& e, int s, int32_t n) {
    if (n != e.n_particles(s)) throw ocm::Error(OCM_ERR_INVALID, "array length does not match the particle count");
}
}  // namespace

extern "C" {

void ocm_options_default(ocm_options* o) {
    if (!o) return;
    o->device = 0;
    o->rebuild_interval = 20;
    o->skin = 0.3;
    o->team_size = 0;
    o->use_graph = 1;
    o->chunk_steps = 200;
    o->reserved = 0;
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

int ocm_set_timing(ocm_context* ctx, int32_t enable) {
    return guarded([&] {
        eng(ctx).set_timing(enable != 0);
        return OCM_OK;
    });
}

int ocm_get_timing(ocm_context* ctx, int32_t kernel, double* total_ms, int64_t* launches) {
    return guarded([&] {
        if (kernel < 0 || kernel > 4) throw ocm::Error(OCM_ERR_INVALID, "kernel id out of range");
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

int32_t ocm_batch_size(ocm_context* ctx) {
    if (!ctx || !ctx->eng) return 0;
    return ctx->eng->n_systems();
}

const char* ocm_last_error(void) { return g_last_error.c_str(); }

const char* ocm_version(void) { return "opencafemol_amd 0.1 (gfx950)"; }

}  // extern "C"
