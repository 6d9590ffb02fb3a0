// engine.cpp -- packing of systems into device arrays and the chunked step loop.
#include "engine.hpp"
#include "trace.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <set>
#include <sstream>

namespace ocm {

// Pair-group members above which a system counts as large: 64-row list-build
// work items, and chunks capped at two list lifetimes.
constexpr int kLargeMembers = 1 << 17;
// Members from which an open-boundary pair group gets a cell grid; below it
// every member pair is a candidate.
constexpr int kOpenGridMembers = 2048;
// Estimated list entries (whole batch) below which no inner list is kept.
constexpr long long kPruneMinEntries = 1 << 14;
// Systems in a batch from which the pair kernel's table is split into XCD slices.
constexpr int kXcdSliceSystems = 16;

// ------------------------------------------------------------ kind tables
bool is_bonded(int k) { return k >= OCM_HARMONIC_BOND && k <= OCM_COSINE_DIHEDRAL; }
bool is_pair(int k) { return k >= OCM_EXCLUDED_VOLUME && k <= OCM_ISOLF_ATTRACTIVE; }
int bonded_arity(int k) {
    switch (k) {
    case OCM_HARMONIC_BOND: case OCM_GAUSSIAN_BOND: case OCM_GO_CONTACT: case OCM_CAPPED_GO_CONTACT: return 2;
    case OCM_HARMONIC_ANGLE: case OCM_FLEXIBLE_LOCAL_ANGLE: return 3;
    case OCM_GAUSSIAN_DIHEDRAL: case OCM_FLEXIBLE_LOCAL_DIHEDRAL: case OCM_COSINE_DIHEDRAL: return 4;
    default: return 0;
    }
}
int bonded_nparams(int k) {
    switch (k) {
    case OCM_HARMONIC_BOND: return 2;
    case OCM_GAUSSIAN_BOND: return 3;
    case OCM_GO_CONTACT: return 2;
    case OCM_HARMONIC_ANGLE: return 2;
    case OCM_FLEXIBLE_LOCAL_ANGLE: return 1;
    case OCM_GAUSSIAN_DIHEDRAL: return 3;
    case OCM_FLEXIBLE_LOCAL_DIHEDRAL: return 1;
    case OCM_CAPPED_GO_CONTACT: return 3;
    case OCM_COSINE_DIHEDRAL: return 3;
    default: return 0;
    }
}
static int bonded_nglobals(int k) {
    if (k == OCM_FLEXIBLE_LOCAL_ANGLE) return 13;
    if (k == OCM_FLEXIBLE_LOCAL_DIHEDRAL) return 8;
    return 0;
}
int pair_npp(int k) {
    switch (k) {
    case OCM_EXCLUDED_VOLUME: return 1;
    case OCM_DEBYE_HUCKEL: return 1;
    case OCM_LJ_REPULSIVE: case OCM_LJ_ATTRACTIVE: case OCM_WCA: return 2;
    case OCM_ISOLF_ATTRACTIVE: return 3;
    default: return 0;   // the uniform kinds carry their parameters as globals
    }
}
static int pair_nglobals(int k) {
    if (k == OCM_EXCLUDED_VOLUME) return 1;
    if (k == OCM_DEBYE_HUCKEL) return 3;
    if (k == OCM_UNIFORM_WCA || k == OCM_UNIFORM_LJ_ATTRACTIVE) return 2;
    return 0;
}
// The four kinds of round shape (EXV, DH, LJ repulsive and attractive) merge
// into one pair group when their lists agree; the others each get a group of
// their own (their per-type parameters reuse the first type-row slots).
static bool solo_pair_kind(int k) { return k >= OCM_WCA; }
static int pair_bit(int k) { return 1 << (k - OCM_EXCLUDED_VOLUME); }
static_assert(OCM_ISOLF_ATTRACTIVE - OCM_EXCLUDED_VOLUME < kPairKinds, "pair kinds fit the kind bits");

// Natural cubic spline second derivatives on uniform knots (the fit OpenMM's
// Continuous1DFunction gives the FLA y-table).
void natural_spline_y2(const double* y, int n, double h, double* y2) {
    for (int i = 0; i < n; ++i) y2[i] = 0.0;
    if (n < 3) return;
    std::vector<double> diag(n), rhs(n);
    for (int i = 1; i < n - 1; ++i) {
        diag[i] = 4.0 * h;
        rhs[i] = 6.0 * (y[i + 1] - 2.0 * y[i] + y[i - 1]) / h;
    }
    for (int i = 2; i < n - 1; ++i) {
        const double m = h / diag[i - 1];
        diag[i] -= m * h;
        rhs[i] -= m * rhs[i - 1];
    }
    y2[n - 2] = rhs[n - 2] / diag[n - 2];
    for (int i = n - 3; i >= 1; --i) y2[i] = (rhs[i] - h * y2[i + 1]) / diag[i];
}

static void invalid(const std::string& m) { throw Error(OCM_ERR_INVALID, m); }

HostSystem copy_system(const ocm_system_desc& d, const ocm_integrator_desc& it) {
    HostSystem s;
    if (d.n_particles <= 0) invalid("system has no particles");
    if (!d.masses) invalid("masses pointer is NULL");
    if (d.n_forces < 0 || (d.n_forces > 0 && !d.forces)) invalid("forces pointer is NULL");
    s.n = d.n_particles;
    s.masses.assign(d.masses, d.masses + s.n);
    for (double m : s.masses)
        if (!(m >= 0.0) || !std::isfinite(m)) invalid("mass must be finite and >= 0");
    s.periodic = d.use_periodic_box != 0;
    for (int c = 0; c < 3; ++c) {
        s.box[c] = d.box[c];
        if (s.periodic && !(d.box[c] > 0.0)) invalid("periodic box edges must be > 0");
    }
    if (!(it.dt > 0.0) || !(it.temperature >= 0.0) || !(it.friction >= 0.0))
        invalid("integrator needs dt > 0, temperature >= 0, friction >= 0");
    s.integ = it;
    for (int fi = 0; fi < d.n_forces; ++fi) {
        const ocm_force_desc& f = d.forces[fi];
        HostForce h;
        h.kind = f.kind;
        h.group = f.group;
        h.use_periodic = f.use_periodic;
        if (f.group < 0 || f.group > 31) invalid("force group must be in 0..31");
        if (h.use_periodic && !s.periodic) invalid("force uses periodic boundaries but the system has no box");
        if (is_bonded(f.kind)) {
            const int ar = bonded_arity(f.kind), np = bonded_nparams(f.kind), ng = bonded_nglobals(f.kind);
            if (f.n_terms < 0 || (f.n_terms > 0 && (!f.indices || !f.params))) invalid("bonded force arrays missing");
            if (f.n_globals != ng || (ng > 0 && !f.globals)) {
                std::ostringstream o;
                o << "force kind " << f.kind << " needs " << ng << " globals, got " << f.n_globals;
                invalid(o.str());
            }
            h.n_terms = f.n_terms;
            h.indices.assign(f.indices, f.indices + static_cast<size_t>(ar) * f.n_terms);
            h.params.assign(f.params, f.params + static_cast<size_t>(np) * f.n_terms);
            if (ng) h.globals.assign(f.globals, f.globals + ng);
            for (int v : h.indices)
                if (v < 0 || v >= s.n) invalid("bonded term index out of range");
            for (double p : h.params)
                if (!std::isfinite(p)) invalid("bonded parameter is not finite");
            for (double g : h.globals)
                if (!std::isfinite(g)) invalid("force global is not finite");
        } else if (is_pair(f.kind)) {
            const int npp = pair_npp(f.kind), ng = pair_nglobals(f.kind);
            if (npp > 0 && !f.particle_params) invalid("nonbonded force needs particle_params");
            if (f.n_globals != ng || (ng > 0 && !f.globals)) {
                std::ostringstream o;
                o << "force kind " << f.kind << " needs " << ng << " globals, got " << f.n_globals;
                invalid(o.str());
            }
            if (!(f.cutoff > 0.0)) invalid("nonbonded cutoff must be > 0");
            h.cutoff = f.cutoff;
            h.switch_distance = f.switch_distance > 0 ? f.switch_distance : 0.0;
            if (h.switch_distance > 0 && f.kind != OCM_LJ_ATTRACTIVE && f.kind != OCM_UNIFORM_LJ_ATTRACTIVE)
                invalid("switching is supported for the Lennard-Jones attractive kinds only");
            if (h.switch_distance >= h.cutoff) invalid("switch distance must be below the cutoff");
            if (npp > 0) h.particle_params.assign(f.particle_params, f.particle_params + static_cast<size_t>(npp) * s.n);
            if (ng) h.globals.assign(f.globals, f.globals + ng);
            if (f.category) {
                if (f.n_categories < 1 || f.n_categories > 32) invalid("n_categories must be in 1..32");
                h.n_categories = f.n_categories;
                h.category.assign(f.category, f.category + s.n);
                for (int c : h.category)
                    if (c < -1 || c >= f.n_categories) invalid("category out of range");
                if (f.category_allow) {
                    h.allow.assign(f.category_allow, f.category_allow + f.n_categories * f.n_categories);
                    for (int a = 0; a < f.n_categories; ++a)
                        for (int b = 0; b < f.n_categories; ++b)
                            if ((h.allow[a * f.n_categories + b] != 0) != (h.allow[b * f.n_categories + a] != 0))
                                invalid("category_allow must be symmetric");
                }
            } else {
                h.n_categories = 1;
            }
            for (int i = 0; i < s.n; ++i) {
                const bool part = h.category.empty() || h.category[i] >= 0;
                if (!part) continue;
                for (int k = 0; k < npp; ++k)
                    if (!std::isfinite(h.particle_params[static_cast<size_t>(npp) * i + k]))
                        invalid("participating bead has a non-finite parameter");
            }
            if (f.n_exclusions < 0 || (f.n_exclusions > 0 && !f.exclusions)) invalid("exclusions pointer is NULL");
            h.exclusions.reserve(f.n_exclusions);
            for (int64_t e = 0; e < f.n_exclusions; ++e) {
                int a = f.exclusions[2 * e], b = f.exclusions[2 * e + 1];
                if (a < 0 || b < 0 || a >= s.n || b >= s.n) invalid("exclusion index out of range");
                if (a == b) continue;
                if (a > b) std::swap(a, b);
                h.exclusions.emplace_back(a, b);
            }
            std::sort(h.exclusions.begin(), h.exclusions.end());
            h.exclusions.erase(std::unique(h.exclusions.begin(), h.exclusions.end()), h.exclusions.end());
        } else {
            std::ostringstream o;
            o << "unknown force kind " << f.kind;
            invalid(o.str());
        }
        s.forces.push_back(std::move(h));
    }
    return s;
}

// --------------------------------------------------------------- engine
void Engine::check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw Error(OCM_ERR_OOM, std::string(what) + ": out of device memory");
        throw Error(OCM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

template <typename T> T* Engine::dalloc(size_t n) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    check_hip(hipMalloc(&p, bytes), "hipMalloc");
    allocs_.push_back(p);
    device_bytes_ += bytes;
    return static_cast<T*>(p);
}

template <typename T> void Engine::upload(T* dst, const std::vector<T>& src) {
    if (src.empty()) return;
    check_hip(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
}

Engine::Engine(std::vector<HostSystem> systems, const ocm_options& opt) : sys_(std::move(systems)), opt_(opt) {
    if (sys_.empty()) invalid("no systems");
    if (opt_.rebuild_interval < 1) opt_.rebuild_interval = 1;
    if (opt_.chunk_steps < 1) opt_.chunk_steps = 200;
    if (!(opt_.skin >= 0.0)) invalid("skin must be >= 0");
    if (opt_.team_size != 0) {
        const int u = opt_.entries_per_lane, t = opt_.team_size;
        if (u != 0 && u != 4 && !(u == 2 && (t == 8 || t == 16 || t == 32)))
            invalid("entries_per_lane must be 4, or 2 with a team of 8, 16 or 32");
    }
    interval_ = opt_.rebuild_interval;
    prune_ = opt_.prune_interval > 1 ? opt_.prune_interval : 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Error(OCM_ERR_NO_DEVICE, "no HIP device found");
    if (opt_.device < 0 || opt_.device >= ndev) throw Error(OCM_ERR_NO_DEVICE, "device ordinal out of range");
    check_hip(hipSetDevice(opt_.device), "hipSetDevice");
    hipDeviceProp_t prop;
    check_hip(hipGetDeviceProperties(&prop, opt_.device), "hipGetDeviceProperties");
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
        throw Error(OCM_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    try {
        pack();
    } catch (...) {
        // no destructor runs for a constructor that throws: free what the
        // packing allocated (a server retrying after running out of memory
        // would otherwise leak it)
        release();
        throw;
    }
    // A stale list rolls the whole chunk back.  For large systems (a step of
    // milliseconds) chunks are capped at two list lifetimes, so a rollback
    // costs at most that; the host check per chunk stays negligible beside it.
    chunk_ = opt_.chunk_steps;
    if (d_.nm > kLargeMembers) chunk_ = std::min(chunk_, std::max(2 * opt_.rebuild_interval, 32));
}

Engine::~Engine() { release(); }

void Engine::release() noexcept {
    // nothing may still run on the stream when its graphs and buffers go
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    event_pool_.clear();
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    if (status_host_) (void)hipHostFree(status_host_);
    status_host_ = nullptr;
    if (stamp_host_) (void)hipHostFree(stamp_host_);
    stamp_host_ = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
}

// Callers guarantee the stream is idle (every chunk ends with a sync).
void Engine::drop_graphs() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second.exec);
    graphs_.clear();
}

// Merge key of a nonbonded force: forces with equal keys share one neighbor
// list and are evaluated in one pass over it.
struct PairKey {
    int periodic;
    double cutoff;
    const HostForce* f;
};
static bool same_list(const HostForce& a, const HostForce& b) {
    return a.use_periodic == b.use_periodic && a.cutoff == b.cutoff && a.category == b.category &&
           a.n_categories == b.n_categories && a.allow == b.allow && a.exclusions == b.exclusions;
}

void Engine::pack() {
    const int nsys = n_systems();
    // ---- beads
    int nb = 0;
    bead_begin_.resize(nsys);
    for (int s = 0; s < nsys; ++s) { bead_begin_[s] = nb; nb += sys_[s].n; }
    d_.nb = nb;
    d_.nsys = nsys;
    std::vector<int> sys_of(nb), local(nb);
    std::vector<double> inv_m(nb), mass(nb);
    std::vector<DevSys> dsys(nsys);
    sys_status_.assign(nsys, OCM_OK);
    err_carry_.assign(nsys, 0);
    sys_step_origin_.assign(nsys, 0);
    sys_group_mask_.assign(nsys, 0);
    const double kB = 1.380649e-23 * 6.02214076e23 / 1000.0;  // kJ/mol/K
    for (int s = 0; s < nsys; ++s) {
        const HostSystem& H = sys_[s];
        for (int i = 0; i < H.n; ++i) {
            const int b = bead_begin_[s] + i;
            sys_of[b] = s;
            local[b] = i;
            mass[b] = H.masses[i];
            inv_m[b] = H.masses[i] > 0 ? 1.0 / H.masses[i] : 0.0;
        }
        DevSys& D = dsys[s];
        std::memset(&D, 0, sizeof(D));
        D.periodic = H.periodic;
        for (int c = 0; c < 3; ++c) {
            D.box[c] = H.box[c];
            D.inv_box[c] = H.periodic ? 1.0 / H.box[c] : 0.0;
        }
        D.bead_begin = bead_begin_[s];
        D.bead_count = H.n;
        D.half_dt = 0.5 * H.integ.dt;
        D.c1 = std::exp(-H.integ.friction * H.integ.dt);
        D.noise_kT = kB * H.integ.temperature * (1.0 - D.c1 * D.c1);
        D.seed = H.integ.seed;
        D.step_origin = 0;
        for (const HostForce& f : H.forces) sys_group_mask_[s] |= 1u << f.group;
    }

    // ---- bonded terms and tables
    std::vector<BTerm> terms;
    std::vector<FlaTable> fla;
    std::vector<FldTable> fld;
    for (int s = 0; s < nsys; ++s) {
        const HostSystem& H = sys_[s];
        for (const HostForce& f : H.forces) {
            if (!is_bonded(f.kind)) continue;
            int table = 0;
            if (f.kind == OCM_FLEXIBLE_LOCAL_ANGLE) {
                FlaTable T{};
                for (int k = 0; k < 10; ++k) T.y[k] = f.globals[k];
                T.tmin = f.globals[10];
                T.tmax = f.globals[11];
                T.emin = f.globals[12];
                if (!(T.tmax > T.tmin)) invalid("FlexibleLocalAngle needs max_theta > min_theta");
                T.h = (T.tmax - T.tmin) / 9.0;
                natural_spline_y2(T.y, 10, T.h, T.y2);
                table = static_cast<int>(fla.size());
                fla.push_back(T);
            } else if (f.kind == OCM_FLEXIBLE_LOCAL_DIHEDRAL) {
                FldTable T{};
                for (int k = 0; k < 8; ++k) T.c[k] = f.globals[k];
                table = static_cast<int>(fld.size());
                fld.push_back(T);
            }
            const int ar = bonded_arity(f.kind), np = bonded_nparams(f.kind);
            for (int t = 0; t < f.n_terms; ++t) {
                BTerm T{};
                for (int k = 0; k < 4; ++k) T.idx[k] = k < ar ? bead_begin_[s] + f.indices[ar * t + k] : -1;
                T.kind = f.kind;
                T.meta = (f.group & 31) | ((f.use_periodic ? 1 : 0) << 5) | (table << 8);
                for (int k = 0; k < 3; ++k) T.p[k] = k < np ? f.params[np * t + k] : 0.0;
                terms.push_back(T);
            }
        }
    }
    std::stable_sort(terms.begin(), terms.end(), [](const BTerm& a, const BTerm& b) { return a.kind < b.kind; });
    d_.nt = static_cast<int>(terms.size());

    // ---- pair groups
    std::vector<DevPairGroup> groups;
    std::vector<int> member_group, member_bead, member_type, member_cat, excl_start, excl_count, excl;
    std::vector<int> cell_group;
    std::vector<float> types_f;
    std::vector<double> types_d;
    int ncell_total = 0;
    int open_grids = 0;
    int max_types = 0;
    force_group_of_.assign(nsys, {});
    for (int s = 0; s < nsys; ++s) {
        const HostSystem& H = sys_[s];
        force_group_of_[s].assign(H.forces.size(), -1);
        std::vector<std::vector<int>> merged;   // lists of force indices
        for (int fi = 0; fi < static_cast<int>(H.forces.size()); ++fi) {
            const HostForce& f = H.forces[fi];
            if (!is_pair(f.kind)) continue;
            bool placed = false;
            for (auto& g : merged) {
                const HostForce& h = H.forces[g[0]];
                if (solo_pair_kind(f.kind) || solo_pair_kind(h.kind)) continue;
                bool kind_taken = false;
                for (int o : g) kind_taken |= H.forces[o].kind == f.kind;
                if (!kind_taken && same_list(h, f)) { g.push_back(fi); placed = true; break; }
            }
            if (!placed) merged.push_back({fi});
        }
        for (const auto& g : merged) {
            const HostForce& f0 = H.forces[g[0]];
            const int gid = static_cast<int>(groups.size());
            DevPairGroup G;
            std::memset(&G, 0, sizeof(G));
            G.sys = s;
            G.periodic = f0.use_periodic ? 1 : 0;
            G.rc = f0.cutoff;
            G.rc2 = f0.cutoff * f0.cutoff;
            double skin = opt_.skin;
            if (G.periodic) {
                const double half = 0.5 * std::min({H.box[0], H.box[1], H.box[2]});
                if (f0.cutoff > half) invalid("nonbonded cutoff exceeds half the periodic box");
                skin = std::min(skin, half - f0.cutoff);
            }
            const double rl = f0.cutoff + skin;
            G.rl2 = rl * rl;
            G.half_skin2 = 0.25 * skin * skin;
            const double margin = std::min(kPruneMargin, skin);
            G.rin2 = (f0.cutoff + margin) * (f0.cutoff + margin);
            G.rin2_wide = G.rin2;
            if (G.periodic) {
                // the FP32 prune's wrapped coordinates lie in [0, L): each axis of
                // a difference is off by at most 2.5 ulp(L); 4 ulp per axis bounds it
                const double L = std::max({H.box[0], H.box[1], H.box[2]});
                const double delta = 4.0 * std::ldexp(1.0, std::ilogb(L) - 23);
                const double rw = f0.cutoff + margin + 2.0 * delta;
                G.rin2_wide = rw * rw;
            }
            prune_hm2_ = std::min(prune_hm2_, 0.25 * margin * margin * (1.0 - 1e-6));
            for (int c = 0; c < 3; ++c) {
                G.box[c] = H.box[c];
                G.inv_box[c] = G.periodic ? 1.0 / H.box[c] : 0.0;
            }
            for (int k = 0; k < kPairKinds; ++k) G.egroup[k] = 0;
            const HostForce* byk[kPairKinds] = {};
            for (int fi : g) {
                const HostForce& f = H.forces[fi];
                const int k = f.kind - OCM_EXCLUDED_VOLUME;
                byk[k] = &f;
                G.kinds |= pair_bit(f.kind);
                G.egroup[k] = f.group;
                force_group_of_[s][fi] = gid;
            }
            if (byk[0]) {
                G.exv_eps = byk[0]->globals[0];
                G.exv_corr = std::pow(1.0 / f0.cutoff, 12);
            }
            if (byk[1]) {
                G.dh_K = byk[1]->globals[0];
                if (!(byk[1]->globals[1] > 0)) invalid("Debye length must be > 0");
                G.dh_inv_lambda = 1.0 / byk[1]->globals[1];
                G.dh_corr = byk[1]->globals[2];
            }
            // the attractive LJ kinds' switch (a group holds at most one of them)
            for (const HostForce* lja : {byk[3], byk[OCM_UNIFORM_LJ_ATTRACTIVE - OCM_EXCLUDED_VOLUME]}) {
                if (!lja) continue;
                G.lja_rsw = lja->switch_distance;
                if (G.lja_rsw > 0) G.lja_inv_sw = 1.0 / (f0.cutoff - G.lja_rsw);
            }
            for (int k : {OCM_UNIFORM_WCA, OCM_UNIFORM_LJ_ATTRACTIVE}) {
                const HostForce* u = byk[k - OCM_EXCLUDED_VOLUME];
                if (!u) continue;
                G.u_sigma = u->globals[0];
                G.u_eps = u->globals[1];
                if (!(G.u_sigma > 0)) invalid("uniform sigma must be > 0");
            }
            const bool solo = solo_pair_kind(f0.kind);
            const int npp0 = pair_npp(f0.kind);
            // categories
            G.n_cat = f0.n_categories;
            for (int a = 0; a < G.n_cat; ++a) {
                uint32_t m = 0;
                for (int b = 0; b < G.n_cat; ++b)
                    if (f0.allow.empty() || f0.allow[a * G.n_cat + b]) m |= 1u << b;
                G.allow[a] = m;
            }
            // members and types
            G.member_begin = static_cast<int>(member_bead.size());
            G.type_begin = static_cast<int>(types_d.size() / kTypeStride);
            std::map<std::array<double, 6>, int> tmap;
            std::vector<int> bead_to_member(H.n, -1);
            for (int i = 0; i < H.n; ++i) {
                const int cat = f0.category.empty() ? 0 : f0.category[i];
                if (cat < 0) continue;
                std::array<double, 6> tup{0, 0, 0, 0, 0, 0};
                if (solo) {
                    // WCA {sigma, epsilon}, iSoLF {sigma, epsilon, omega}; the
                    // uniform kinds have one type
                    for (int k = 0; k < npp0; ++k) tup[k] = f0.particle_params[static_cast<size_t>(npp0) * i + k];
                } else {
                    if (byk[0]) tup[0] = byk[0]->particle_params[i];
                    if (byk[1]) tup[1] = byk[1]->particle_params[i];
                    if (byk[2]) { tup[2] = byk[2]->particle_params[2 * i]; tup[3] = byk[2]->particle_params[2 * i + 1]; }
                    if (byk[3]) { tup[4] = byk[3]->particle_params[2 * i]; tup[5] = byk[3]->particle_params[2 * i + 1]; }
                }
                auto it = tmap.find(tup);
                int ty;
                if (it == tmap.end()) {
                    ty = static_cast<int>(tmap.size());
                    tmap.emplace(tup, ty);
                    if (solo) {
                        // FP64 rows {sigma, epsilon, omega}; the force pass's FP32
                        // rows {sigma/2, sqrt(epsilon), omega/2}: the arithmetic
                        // means are then single adds and the geometric mean a product
                        const double row[kTypeStride] = {tup[0], tup[1], tup[2], 0, 0, 0, 0, 0};
                        for (int k = 0; k < kTypeStride; ++k) types_d.push_back(row[k]);
                        if (tup[1] < 0) invalid("WCA / iSoLF epsilon must be >= 0");
                        const float rowf[kTypeStride] = {0.5f * static_cast<float>(tup[0]),
                                                         static_cast<float>(std::sqrt(tup[1])),
                                                         0.5f * static_cast<float>(tup[2]), 0.f, 0.f, 0.f, 0.f, 0.f};
                        for (int k = 0; k < kTypeStride; ++k) types_f.push_back(rowf[k]);
                    } else {
                        const double row[kTypeStride] = {tup[0], tup[1], 0.0, 0.0, tup[2], tup[3], tup[4], tup[5]};
                        for (int k = 0; k < kTypeStride; ++k) {
                            types_d.push_back(row[k]);
                            // the force pass's FP32 rows hold half of each LJ epsilon and
                            // sigma (its mixing rules are then single adds)
                            const float v = static_cast<float>(row[k]);
                            types_f.push_back(k >= kTyLjrEps ? 0.5f * v : v);
                        }
                    }
                } else {
                    ty = it->second;
                }
                bead_to_member[i] = static_cast<int>(member_bead.size());
                member_group.push_back(gid);
                member_bead.push_back(bead_begin_[s] + i);
                member_type.push_back(ty | (cat << 16));  // category rides in bits 16-20
                member_cat.push_back(cat);
            }
            G.member_count = static_cast<int>(member_bead.size()) - G.member_begin;
            G.n_types = static_cast<int>(tmap.size());
            G.lj_shared = byk[2] && byk[3] ? 1 : 0;
            for (const auto& kv : tmap)
                if (kv.first[3] != kv.first[5]) G.lj_shared = 0;
            if (G.n_types * kTypeStride * sizeof(float) > 64 * 1024)
                invalid("more than 2048 distinct per-particle parameter sets in one nonbonded force");
            max_types = std::max(max_types, G.n_types);
            // exclusions in member space, symmetric CSR sorted by partner
            std::vector<std::vector<int>> adj(G.member_count);
            for (const auto& e : f0.exclusions) {  // partners stored as global bead ids
                const int a = bead_to_member[e.first], b = bead_to_member[e.second];
                if (a < 0 || b < 0) continue;
                adj[a - G.member_begin].push_back(bead_begin_[s] + e.second);
                adj[b - G.member_begin].push_back(bead_begin_[s] + e.first);
            }
            for (int m = 0; m < G.member_count; ++m) {
                auto& v = adj[m];
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
                excl_start.push_back(static_cast<int>(excl.size()));
                excl_count.push_back(static_cast<int>(v.size()));
                excl.insert(excl.end(), v.begin(), v.end());
            }
            // cell grid
            long long est_neighbors;
            if (G.periodic) {
                long long total = 1;
                for (int c = 0; c < 3; ++c) {
                    int n = static_cast<int>(std::floor(H.box[c] / rl));
                    if (n < 3) { n = 1; G.reach[c] = 0; } else { G.reach[c] = 1; }
                    G.ncell[c] = n;
                    total *= n;
                }
                while (total > (1LL << 24)) {  // cap the grid; coarser cells keep the list exact
                    total = 1;
                    for (int c = 0; c < 3; ++c) {
                        if (G.ncell[c] >= 6) G.ncell[c] /= 2;
                        if (G.ncell[c] < 3) { G.ncell[c] = 1; G.reach[c] = 0; }
                        total *= G.ncell[c];
                    }
                }
                for (int c = 0; c < 3; ++c) G.cell_scale[c] = G.ncell[c] / H.box[c];
                const double vol = H.box[0] * H.box[1] * H.box[2];
                const double rho = G.member_count / vol;
                est_neighbors = static_cast<long long>(G.member_count * (rho * 4.18879 * rl * rl * rl * 3.0 + 8.0));
            } else {
                for (int c = 0; c < 3; ++c) { G.ncell[c] = 1; G.reach[c] = 0; G.cell_scale[c] = 0; G.cell_lo[c] = 0; }
                // a large open group gets a grid over its bounding box, laid
                // out at every build (k_open_grid); small ones test all pairs
                if (G.member_count >= kOpenGridMembers) {
                    G.open_grid = 1;
                    G.grid_cap = std::min(std::max(G.member_count / 4, 27), 1 << 22);
                    ++open_grids;
                }
                est_neighbors = static_cast<long long>(G.member_count) * std::min(G.member_count, 256);
            }
            G.cell_begin = ncell_total;
            ncell_total += G.open_grid ? G.grid_cap : G.ncell[0] * G.ncell[1] * G.ncell[2];
            cell_group.resize(ncell_total, gid);
            group_list_estimate_.push_back(est_neighbors);
            group_member_begin_.push_back(G.member_begin);
            group_member_count_.push_back(G.member_count);
            groups.push_back(G);
        }
    }
    if (member_bead.size() >= (1u << 26))   // record offsets are 32-bit byte offsets
        throw Error(OCM_ERR_OOM, "more than 2^26 pair-group members");
    // a small list is read in the force pass's launch latency whatever its
    // length, so a prune launch cannot pay (62-bead Go chain: 8.95 -> 8.34
    // us/step without it); such batches read the full list
    long long est_entries = 0;
    for (long long e : group_list_estimate_) est_entries += e;
    if (est_entries < kPruneMinEntries) prune_ = 0;
    d_.nm = static_cast<int>(member_bead.size());
    // large lists: 64-row work items stage each stencil once for a whole cell
    // (there are still far more items than workgroups); small ones keep 16-row
    // items so the build spreads over the device
    d_.tile_chunk = d_.nm > kLargeMembers ? 64 : 16;
    d_.ncell = std::max(ncell_total, 1);
    d_.open_grids = open_grids;
    d_.ngroups = static_cast<int>(groups.size());
    d_.max_types = std::max(max_types, 1);


    // ---- device allocation and upload
    const int n3 = 3 * nb;
    double* pos = dalloc<double>(n3);
    double* vel = dalloc<double>(n3);
    long long* frc = dalloc<long long>(n3);
    long long* frc2 = dalloc<long long>(n3);
    d_.px = pos; d_.py = pos + nb; d_.pz = pos + 2 * nb;
    d_.vx = vel; d_.vy = vel + nb; d_.vz = vel + 2 * nb;
    d_.f = {frc, frc + nb, frc + 2 * nb};
    d_.f2 = {frc2, frc2 + nb, frc2 + 2 * nb};
    check_hip(hipMemset(pos, 0, sizeof(double) * n3), "memset");
    check_hip(hipMemset(vel, 0, sizeof(double) * n3), "memset");
    check_hip(hipMemset(frc, 0, sizeof(long long) * n3), "memset");
    bstate_ = dalloc<double>(3 * static_cast<size_t>(n3) + 2);
    {
        // pinned, device-mapped and coherent: the chunk's last kernel writes
        // the status words straight into it
        void* h = nullptr;
        check_hip(hipHostMalloc(&h, sizeof(int) * (1 + static_cast<size_t>(nsys)),
                                hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
        status_host_ = static_cast<int*>(h);
        std::memset(status_host_, 0, sizeof(int) * (1 + static_cast<size_t>(nsys)));
    }
    d_.sys = dalloc<DevSys>(nsys);
    upload(d_.sys, dsys);
    d_.sys0 = dsys[0];
    d_.sys_of = dalloc<int>(nb); upload(d_.sys_of, sys_of);
    d_.local_id = dalloc<int>(nb); upload(d_.local_id, local);
    d_.inv_m = dalloc<double>(nb); upload(d_.inv_m, inv_m);
    d_.mass = dalloc<double>(nb); upload(d_.mass, mass);
    d_.gstep = dalloc<long long>(2);
    check_hip(hipMemset(d_.gstep, 0, 2 * sizeof(long long)), "memset");
    d_.err = dalloc<int>(nsys);
    check_hip(hipMemset(d_.err, 0, nsys * sizeof(int)), "memset");
    d_.flags = dalloc<int>(1);
    check_hip(hipMemset(d_.flags, 0, sizeof(int)), "memset");
    d_.terms = dalloc<BTerm>(terms.size()); upload(d_.terms, terms);
    d_.fla = dalloc<FlaTable>(fla.size()); upload(d_.fla, fla);
    d_.fld = dalloc<FldTable>(fld.size()); upload(d_.fld, fld);
    // rows of a group get a fixed number of list slots: the estimate per row,
    // widened on overflow (grow_rows)
    for (int g = 0; g < static_cast<int>(groups.size()); ++g) {
        const long long per_row = group_list_estimate_[g] / std::max(groups[g].member_count, 1) + 8;
        groups[g].stride = static_cast<int>(std::min<long long>((per_row + 3) & ~3LL, 1 << 20));
    }
    host_groups_ = groups;
    d_.groups = dalloc<DevPairGroup>(std::max<size_t>(groups.size(), 1)); upload(d_.groups, groups);
    d_.group_need = dalloc<int>(std::max<size_t>(groups.size(), 1));
    check_hip(hipMemset(d_.group_need, 0, sizeof(int) * std::max<size_t>(groups.size(), 1)), "memset");
    const int nm = d_.nm;
    d_.member_group = dalloc<int>(nm); upload(d_.member_group, member_group);
    d_.member_bead = dalloc<int>(nm); upload(d_.member_bead, member_bead);
    d_.member_type = dalloc<int>(nm); upload(d_.member_type, member_type);
    d_.member_cat = dalloc<int>(nm); upload(d_.member_cat, member_cat);
    d_.member_slot = dalloc<int>(nm);
    d_.member_cell = dalloc<int>(nm);
    d_.member_rank = dalloc<int>(nm);
    {
        // fixed-stride per-bead member tables, so the integrator reaches a
        // bead's records with one dependent load
        std::vector<int> cnt(nb, 0), kpos(std::max(nm, 1), 0);
        for (int m = 0; m < nm; ++m) ++cnt[member_bead[m]];
        int maxm = 1;
        for (int b = 0; b < nb; ++b) maxm = std::max(maxm, cnt[b]);
        d_.maxm = maxm;
        std::vector<int> fillc(nb, 0);
        std::vector<double> hs2(nb, std::numeric_limits<double>::max());   // per bead: smallest (skin/2)^2
        for (int m = 0; m < nm; ++m) {
            const int b = member_bead[m];
            const int k = b * maxm + fillc[b]++;
            kpos[m] = k;
            hs2[b] = std::min(hs2[b], groups[member_group[m]].half_skin2);
        }
        d_.bead_slot = dalloc<int>(static_cast<size_t>(nb) * maxm);
        // slots of absent memberships stay -1; the build fills the others
        check_hip(hipMemset(d_.bead_slot, 0xFF, sizeof(int) * static_cast<size_t>(nb) * maxm), "memset");
        d_.bead_hs2 = dalloc<double>(hs2.size()); upload(d_.bead_hs2, hs2);
        d_.member_kpos = dalloc<int>(kpos.size()); upload(d_.member_kpos, kpos);
    }
    excl_bead_ = member_bead;
    excl_start_h_ = excl_start;
    excl_count_h_ = excl_count;
    excl_h_ = excl;
    d_.excl_start = dalloc<int>(nm); upload(d_.excl_start, excl_start);
    d_.excl_count = dalloc<int>(nm); upload(d_.excl_count, excl_count);
    d_.excl = dalloc<int>(excl.size()); upload(d_.excl, excl);
    d_.refx = dalloc<double>(nb); d_.refy = dalloc<double>(nb); d_.refz = dalloc<double>(nb);
    d_.rec = dalloc<PairRec>(nm);
    d_.slot_member = dalloc<int>(nm);
    d_.cell_count = dalloc<int>(d_.ncell);
    check_hip(hipMemset(d_.cell_count, 0, sizeof(int) * d_.ncell), "memset");   // each build leaves it zero
    d_.cell_start = dalloc<int>(d_.ncell + 1);   // exclusive scan; [ncell] holds the total
    d_.cell_group = dalloc<int>(d_.ncell);
    d_.tile_items = dalloc<int2>(static_cast<size_t>(nm) / 16 + std::min(d_.ncell, std::max(nm, 1)) + 16);
    d_.tile_nitems = dalloc<int>(1);
    if (cell_group.empty()) cell_group.assign(d_.ncell, 0);
    upload(d_.cell_group, cell_group);
    d_.row_count = dalloc<int>(nm);
    d_.row_start = dalloc<int>(nm);
    check_hip(hipMemset(d_.row_count, 0, sizeof(int) * std::max(nm, 1)), "memset");
    if (prune_ && nm > 0) {
        d_.row_count_in = dalloc<int>(nm);
        check_hip(hipMemset(d_.row_count_in, 0, sizeof(int) * nm), "memset");
        d_.prx = dalloc<double>(nb); d_.pry = dalloc<double>(nb); d_.prz = dalloc<double>(nb);
        d_.rec32 = dalloc<float4>(nm);
        d_.prune_hm2 = prune_hm2_;
    } else {
        prune_ = 0;
    }
    layout_rows();
    const int scan_n = std::max(d_.ncell, nm);
    d_.scan_partial = dalloc<int>(scan_n / 2048 + 2);
    d_.scan_total = dalloc<int>(2);
    check_hip(hipMemset(d_.scan_total, 0, 2 * sizeof(int)), "memset");
    d_.types_f = dalloc<float>(types_f.size()); upload(d_.types_f, types_f);
    d_.types_d = dalloc<double>(types_d.size()); upload(d_.types_d, types_d);
    d_.energy = dalloc<long long>(static_cast<size_t>(nsys) * 32);
    d_.ke = dalloc<long long>(nsys);
    d_.dump_count = dalloc<int>(1);
    int team = opt_.team_size;
    if (team != 4 && team != 8 && team != 16 && team != 32 && team != 64) team = 16;
    build_pair_blocks(team);
    team_tuned_ = opt_.team_size != 0;
    if (team_tuned_ && team == opt_.team_size && opt_.entries_per_lane != 0) d_.pair_u = opt_.entries_per_lane;
    if (opt_.resident != 0) pack_resident(terms, groups, types_f);
    excl_bead_.clear(); excl_start_h_.clear(); excl_count_h_.clear(); excl_h_.clear();
    // resident batches step outside the pair kernel: no launch shape to tune
    if (d_.res_n > 0) team_tuned_ = true;
    check_hip(hipDeviceSynchronize(), "upload sync");
}

// Resident stepping, when every system fits one workgroup: each system's
// bonded terms with local bead indices, its groups' FP32 type rows staged
// together, and a list region of nm (nm - 1) / 2 entries per group (every
// member pair fits, so the in-kernel build never overflows).  Leaves
// d_.res_n at 0 when some system does not qualify.
void Engine::pack_resident(const std::vector<BTerm>& terms, const std::vector<DevPairGroup>& groups,
                           const std::vector<float>& types_f) {
    const int nsys = n_systems();
    int max_n = 1;
    std::vector<std::vector<int>> sys_groups(nsys);
    for (int g = 0; g < static_cast<int>(groups.size()); ++g) sys_groups[groups[g].sys].push_back(g);
    for (int s = 0; s < nsys; ++s) {
        const HostSystem& H = sys_[s];
        if (H.n > kResMaxBeads || static_cast<int>(sys_groups[s].size()) > kResMaxGroups) return;
        int rows = 0;
        for (int g : sys_groups[s]) {
            const DevPairGroup& G = groups[g];
            // the build's FP32 coordinates: wrapped into a box of at most
            // kResWide nm, or open positions relative to bead 0
            if ((G.periodic != 0) != H.periodic) return;
            if (G.periodic && std::max({H.box[0], H.box[1], H.box[2]}) > kResWide) return;
            rows += G.n_types;
        }
        if (rows > kResMaxTypeRows) return;
        max_n = std::max(max_n, H.n);
    }
    std::vector<ResSys> tab(nsys);
    std::vector<ResGroup> grp;
    std::vector<BTerm> rterms;
    std::vector<float> rows;
    std::vector<int> rexcl;
    // per bead: the smallest (skin/2)^2 of its groups' resident lists
    std::vector<double> res_hs2(static_cast<size_t>(d_.nb), std::numeric_limits<double>::max());
    std::vector<std::vector<int>> sys_terms(nsys);
    for (int t = 0; t < static_cast<int>(terms.size()); ++t) {
        const int b = terms[t].idx[0];
        const int s = static_cast<int>(std::upper_bound(bead_begin_.begin(), bead_begin_.end(), b) -
                                       bead_begin_.begin()) - 1;
        sys_terms[s].push_back(t);
    }
    long long l0 = 0;
    for (int s = 0; s < nsys; ++s) {
        ResSys& R = tab[s];
        R.sys = s;
        R.nb = sys_[s].n;
        R.b0 = bead_begin_[s];
        R.t0 = static_cast<int>(rterms.size());
        R.nt = static_cast<int>(sys_terms[s].size());
        R.nseg = 0;
        for (int t : sys_terms[s]) {   // kind order is kept within the system
            BTerm T = terms[t];
            for (int k = 0; k < 4; ++k)
                if (T.idx[k] >= 0) T.idx[k] -= R.b0;
            const int local = static_cast<int>(rterms.size()) - R.t0;
            if (R.nseg == 0 || rterms.back().kind != T.kind) {
                if (R.nseg == kResMaxSeg) throw Error(OCM_ERR_INVALID, "resident terms are not in kind order");
                R.seg_t[R.nseg] = local;
                R.seg_n[R.nseg] = 0;
                ++R.nseg;
            }
            ++R.seg_n[R.nseg - 1];
            rterms.push_back(T);
        }
        R.g0 = static_cast<int>(grp.size());
        R.ng = static_cast<int>(sys_groups[s].size());
        R.row0 = static_cast<int>(rows.size() / kTypeStride);
        R.ex0 = static_cast<int>(rexcl.size());
        int trow = 0;
        for (int g : sys_groups[s]) {
            const DevPairGroup& G = groups[g];
            ResGroup Q{};
            Q.gid = g;
            Q.m0 = G.member_begin;
            Q.nm = G.member_count;
            Q.trow = trow;
            Q.l0 = l0;
            // exclusions by member index (the members of a group are in bead order)
            Q.exo = static_cast<int>(rexcl.size()) - R.ex0;
            {
                std::vector<int> member_of(sys_[s].n, -1);
                for (int i = 0; i < Q.nm; ++i) member_of[excl_bead_[Q.m0 + i] - R.b0] = i;
                std::vector<uint32_t> mask(Q.nm, 0u);
                std::vector<std::vector<int>> far(Q.nm);
                for (int i = 0; i < Q.nm; ++i) {
                    const int m = Q.m0 + i;
                    for (int e = excl_start_h_[m]; e < excl_start_h_[m] + excl_count_h_[m]; ++e) {
                        const int j = member_of[excl_h_[e] - R.b0];
                        if (j <= i) continue;   // the pair is tested from its lower member
                        if (j - i <= kResNearExcl) mask[i] |= 1u << (j - i - 1);
                        else far[i].push_back(j);
                    }
                    std::sort(far[i].begin(), far[i].end());
                }
                for (int i = 0; i < Q.nm; ++i) rexcl.push_back(static_cast<int>(mask[i]));
                int at = 0;
                for (int i = 0; i < Q.nm; ++i) { rexcl.push_back(at); at += static_cast<int>(far[i].size()); }
                rexcl.push_back(at);
                for (int i = 0; i < Q.nm; ++i) rexcl.insert(rexcl.end(), far[i].begin(), far[i].end());
            }
            // the resident list's skin: at most kResSkin (a list read every
            // step by one workgroup pays for its width at every step, a
            // rebuild is cheap); the FP32 test's threshold adds 1e-3 nm, well
            // above the 2.2e-4 nm error of coordinates below kResWide
            const double skin_r = std::min(2.0 * std::sqrt(G.half_skin2), kResSkin);
            const double rl = G.rc + skin_r;
            Q.rl2 = rl * rl;
            Q.rl2w = static_cast<float>((rl + 1e-3) * (rl + 1e-3));
            for (int i = 0; i < G.member_count; ++i) {
                double& h = res_hs2[excl_bead_[G.member_begin + i]];
                h = std::min(h, 0.25 * skin_r * skin_r);
            }
            l0 += std::max<long long>(1, static_cast<long long>(Q.nm) * (Q.nm - 1) / 2);
            for (int k = 0; k < G.n_types * kTypeStride; ++k)
                rows.push_back(types_f[static_cast<size_t>(G.type_begin) * kTypeStride + k]);
            trow += G.n_types;
            grp.push_back(Q);
        }
        R.nrow = trow;
        R.exn = static_cast<int>(rexcl.size()) - R.ex0;
    }
    // a step's work (terms, list entries, noise) is dealt over all threads:
    // 1024 of them keep every thread's share short
    d_.res_threads = 1024;
    d_.res_nbs = (max_n + 63) & ~63;
    // LDS: what one system's step reads, as far as it fits.  Up to one
    // workgroup per CU the whole 160 KB may go to one; beyond that half, so
    // that two workgroups share a CU and hide each other's latency.
    int max_nt = 0, max_ex = 0;
    long long max_pairs = 0;
    for (int s = 0; s < nsys; ++s) {
        max_nt = std::max(max_nt, tab[s].nt);
        max_ex = std::max(max_ex, tab[s].exn);
        long long p = 0;
        for (int g = tab[s].g0; g < tab[s].g0 + tab[s].ng; ++g)
            p += static_cast<long long>(grp[g].nm) * (grp[g].nm - 1) / 2;
        max_pairs = std::max(max_pairs, p);
    }
    int cus = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, opt_.device) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    const size_t budget = (nsys <= cus ? kResLdsMax : kResLdsMax / 2) - 1024;
    // exclusion tables are staged whole; they are small (chains exclude near
    // neighbours, which are mask bits)
    d_.res_ex_cap = max_ex;
    const size_t fixed = resident_lds_bytes(d_.res_nbs, 0, 0, d_.res_ex_cap);
    if (fixed > budget) { d_.res_n = 0; return; }
    const size_t room = budget - fixed;
    d_.res_list_cap = static_cast<int>(std::min<long long>({max_pairs, kResListLdsMax,
                                                            static_cast<long long>(room / sizeof(uint32_t))}));
    d_.res_term_cap = static_cast<int>(std::min<size_t>(static_cast<size_t>(max_nt),
                                                        (room - d_.res_list_cap * sizeof(uint32_t)) / sizeof(BTerm)));
    const size_t lds = resident_lds_bytes(d_.res_nbs, d_.res_term_cap, d_.res_list_cap, d_.res_ex_cap);
    check_hip(resident_prepare(d_.res_threads, lds), "hipFuncSetAttribute (resident LDS)");
    d_.res_tab = dalloc<ResSys>(tab.size()); upload(d_.res_tab, tab);
    d_.res_grp = dalloc<ResGroup>(std::max<size_t>(grp.size(), 1));
    if (!grp.empty()) upload(d_.res_grp, grp);
    d_.res_terms = dalloc<BTerm>(std::max<size_t>(rterms.size(), 1));
    if (!rterms.empty()) upload(d_.res_terms, rterms);
    d_.res_rows = dalloc<float>(std::max<size_t>(rows.size(), 1));
    if (!rows.empty()) upload(d_.res_rows, rows);
    d_.res_list = dalloc<uint32_t>(static_cast<size_t>(std::max<long long>(l0, 1)));
    d_.res_hs2 = dalloc<double>(std::max<size_t>(res_hs2.size(), 1));
    upload(d_.res_hs2, res_hs2);
    d_.res_excl = dalloc<int>(std::max<size_t>(rexcl.size(), 1));
    if (!rexcl.empty()) upload(d_.res_excl, rexcl);
    d_.res_n = nsys;
}

void Engine::build_pair_blocks(int team) {
    d_.team = team;
    const int rpb = kPairBlock / team;
    std::vector<int2> blocks;
    // Batches of many systems: the pair kernel's workgroups walk the table in
    // 8 contiguous slices, one per XCD (k_pair's SLICED walk), and the table
    // is laid out as 8 slices with the same mix of work: each holds an eighth
    // of the bonded-term chunks (leading: latency-bound, they start while the
    // pair rows run) and an eighth of every group's rows.  Measured on 256 Go
    // replicas: +5-8% over the plain walk, from L1 locality: the workgroups on
    // one CU take adjacent entries, and L2 requests fall by 18% with misses
    // unchanged (slices of whole systems, balanced by estimated work, gained
    // only 2% there and lost 6% on the served mix).  A single slab, whose
    // long rows sit in the middle of the cell order, lost 6% when split, so
    // single systems and small batches keep one slice and the plain walk.
    // Cuts fall on item boundaries.
    d_.xcd_sliced = d_.nsys >= kXcdSliceSystems ? 1 : 0;
    const int slices = d_.xcd_sliced ? 8 : 1;
    const auto cut = [&](long long n_items, int x) { return static_cast<int>(n_items * x / slices); };
    const long long n_bonded = (d_.nt + kPairBlock - 1) / kPairBlock;
    for (int x = 0; x < slices; ++x) {
        for (int c = cut(n_bonded, x); c < cut(n_bonded, x + 1); ++c) blocks.push_back(make_int2(-1, c * kPairBlock));
        for (int g = 0; g < d_.ngroups; ++g) {
            const long long n_items = (group_member_count_[g] + rpb - 1) / rpb;
            for (int c = cut(n_items, x); c < cut(n_items, x + 1); ++c)
                blocks.push_back(make_int2(g, group_member_begin_[g] + c * rpb));
        }
    }
    // the block table buffer is reallocated only when it grows
    static_assert(sizeof(int2) == 8, "int2 layout");
    if (static_cast<int>(blocks.size()) > d_.n_pair_blocks || !d_.pair_blocks) d_.pair_blocks = dalloc<int2>(blocks.size());
    upload(d_.pair_blocks, blocks);
    d_.n_pair_blocks = static_cast<int>(blocks.size());
    drop_graphs();
}

// ----------------------------------------------------------- timing
template <typename F> void Engine::timed(int slot, F&& f) {
    if (!timing_) { f(); return; }
    if (event_next_ + 2 > event_pool_.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            check_hip(hipEventCreate(&e), "hipEventCreate");
            event_pool_.push_back(e);
        }
    }
    hipEvent_t a = event_pool_[event_next_++], b = event_pool_[event_next_++];
    check_hip(hipEventRecord(a, stream_), "hipEventRecord");
    f();
    check_hip(hipEventRecord(b, stream_), "hipEventRecord");
    pending_events_.push_back({slot, {a, b}});
}

void Engine::set_timing(int mode) {
    if (mode < 0 || mode > 2) invalid("timing mode must be 0, 1 or 2");
    sync();
    timing_mode_ = mode;
    timing_ = mode == 1;
    if (mode == 2 && wall_khz_ == 0) {
        void* h = nullptr;
        check_hip(hipHostMalloc(&h, 2 * sizeof(unsigned long long) * static_cast<size_t>(opt_.chunk_steps),
                                hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
        stamp_host_ = static_cast<unsigned long long*>(h);
        int khz = 0;
        check_hip(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, opt_.device), "wall clock rate");
        if (khz <= 0) throw Error(OCM_ERR_HIP, "device wall clock rate unknown");
        wall_khz_ = khz;
    }
}

// Stamp buffer for the current pair grid; a new buffer means new kernel
// arguments, so captured graphs are dropped.
void Engine::ensure_stamps() {
    const int grid = pair_grid_size(d_);
    if (grid != kstamp_grid_) {
        kstamp_grid_ = grid;
        drop_graphs();
    }
    const long long words = 2LL * opt_.chunk_steps * std::max(grid, 1);
    if (words <= kstamp_words_) return;
    if (kstamp_) {
        auto it = std::find(allocs_.begin(), allocs_.end(), static_cast<void*>(kstamp_));
        if (it != allocs_.end()) { (void)hipFree(*it); allocs_.erase(it); }
    }
    kstamp_ = dalloc<unsigned long long>(static_cast<size_t>(words));
    kstamp_words_ = words;
    drop_graphs();
}

void Engine::collect_timing() {
    for (auto& pe : pending_events_) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            timing_slots_[pe.first].ms += ms;
            timing_slots_[pe.first].launches += 1;
        }
    }
    pending_events_.clear();
    event_next_ = 0;
}

void Engine::sync() {
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    if (timing_) collect_timing();
}

// ------------------------------------------------------------- stepping
void Engine::enqueue_forces(const ForceBuf& f, bool rebuild, uint32_t groups) {
    rebuild = rebuild || lists_stale_;
    lists_stale_ = false;
    if (d_.nm > 0) {
        if (rebuild) { timed(3, [&] { launch_rebuild(d_, stream_); }); ++rebuilds_; }
        else timed(4, [&] { launch_pack(d_, stream_); });
    }
    // bonded terms included; a full evaluation also prunes the inner list
    timed(0, [&] { launch_pair(d_, stream_, false, f, groups, kListPrune); });
}

// The list schedule of a chunk of k steps.  Ages count integration steps:
// after step s the outer list is rebuilt when it is due (age >= interval, or
// the lists are stale) and the inner list pruned at every rebuild and when
// its own age reaches the prune interval; other steps read the inner list.
// Ages carry over from chunk to chunk, so the list lifetime does not depend
// on how steps are grouped into calls.
Engine::ChunkPlan Engine::plan_chunk(int k) const { return plan_chunk_at(k, age_, page_, fresh_); }

Engine::ChunkPlan Engine::plan_chunk_at(int k, int age_in, int page_in, int fresh_in) const {
    ChunkPlan p;
    p.k = k;
    p.interval = interval_;
    p.prune = prune_ > 1 ? prune_ : 0;
    p.fresh = fresh_in;
    p.age0 = fresh_in ? 0 : age_in;
    p.page0 = fresh_in || !p.prune ? 0 : page_in;
    p.rebuild.assign(k, 0);
    p.pruned.assign(k, 0);
    int age = p.age0, page = p.page0;
    for (int s = 0; s < k; ++s) {
        ++age;
        ++page;
        const bool rb = (s == 0 && p.fresh) || age >= p.interval;
        if (rb) { age = 0; ++p.rebuilds; }
        const bool pr = p.prune > 1 && (rb || page >= p.prune);
        if (pr) page = 0;
        p.rebuild[s] = rb;
        p.pruned[s] = pr;
    }
    p.age1 = age;
    p.page1 = page;
    return p;
}

// One chunk: the snapshot a rollback restores, the steps, the closing kick,
// and the status words into pinned host memory.
void Engine::enqueue_chunk(const ChunkPlan& p) {
    launch_snapshot(d_, stream_, bstate_);
    for (int s = 0; s < p.k; ++s) {
        const bool rebuild = p.rebuild[s] != 0;
        const bool pruned = p.pruned[s] != 0;
        const int pr = p.prune > 1 ? (pruned ? kPruneWrite : kPruneCheck) : kPruneNone;
        // the integrator refreshes the pair records itself on non-rebuild steps
        timed(2, [&] { launch_integrate(d_, stream_, s & 1, (s + 1) & 1, s > 0 ? 1 : 0, rebuild ? 0 : 1, 1, pr); });
        if (d_.nm > 0 && rebuild) timed(3, [&] { launch_rebuild(d_, stream_); });
        unsigned long long* stamp =
            timing_mode_ == 2 ? kstamp_ + 2 * static_cast<size_t>(kstamp_grid_) * s : nullptr;
        // bonded terms run inside the pair kernel's force pass
        if (pruned) timed(6, [&] { launch_prune(d_, stream_); });
        const int list = p.prune > 1 ? kListInner : kListOuter;
        timed(0, [&] { launch_pair(d_, stream_, false, d_.f, 0xFFFFFFFFu, list, stamp); });
    }
    timed(2, [&] { launch_half_kick(d_, stream_, p.k & 1); });
    if (timing_mode_ == 2) launch_reduce_stamps(stream_, kstamp_, p.k, kstamp_grid_, stamp_host_);
    launch_status_out(d_, stream_, status_host_);
}

void Engine::restore() {
    const size_t n3 = 3 * static_cast<size_t>(d_.nb);
    const long long* b = reinterpret_cast<const long long*>(bstate_);
    check_hip(hipMemcpyAsync(d_.px, bstate_, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_), "restore");
    check_hip(hipMemcpyAsync(d_.vx, bstate_ + n3, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_), "restore");
    check_hip(hipMemcpyAsync(d_.f.fx, b + 2 * n3, n3 * sizeof(long long), hipMemcpyDeviceToDevice, stream_), "restore");
    check_hip(hipMemcpyAsync(d_.gstep, b + 3 * n3, 2 * sizeof(long long), hipMemcpyDeviceToDevice, stream_), "restore");
    // the lists were built for states past the restored one
    stale_lists();
}

// Row r of group g owns list slots [row_start[r], row_start[r] + stride_g);
// groups follow each other in col.  (Re)allocates col and uploads row_start.
void Engine::layout_rows() {
    std::vector<int> row_start(std::max(d_.nm, 1), 0);
    long long total = 0;
    for (const DevPairGroup& G : host_groups_) {
        for (int r = 0; r < G.member_count; ++r) {
            if (total > std::numeric_limits<int>::max() - G.stride)
                throw Error(OCM_ERR_OOM, "neighbor list exceeds 2^31 entries");
            row_start[G.member_begin + r] = static_cast<int>(total);
            total += G.stride;
        }
    }
    upload(d_.row_start, row_start);
    if (d_.col) {
        auto it = std::find(allocs_.begin(), allocs_.end(), static_cast<void*>(d_.col));
        if (it != allocs_.end()) { (void)hipFree(*it); allocs_.erase(it); }
        d_.col = nullptr;
    }
    d_.list_capacity = std::max<long long>(total, 4);
    d_.col = dalloc<int>(d_.list_capacity);
    col_bytes_ = d_.list_capacity * static_cast<long long>(sizeof(int));
    if (d_.row_count_in) {
        if (d_.col_in) {
            auto it = std::find(allocs_.begin(), allocs_.end(), static_cast<void*>(d_.col_in));
            if (it != allocs_.end()) { (void)hipFree(*it); allocs_.erase(it); }
            d_.col_in = nullptr;
        }
        d_.col_in = dalloc<int>(d_.list_capacity);
        col_bytes_ *= 2;
        // a widened layout leaves no valid inner rows until the next prune
        check_hip(hipMemset(d_.row_count_in, 0, sizeof(int) * std::max(d_.nm, 1)), "memset");
    }
}

// After a build reported rows longer than their slots: widen those groups to
// 1.25x the longest row seen (plus 8) and lay the list out again.
void Engine::grow_rows() {
    sync();
    const int ng = static_cast<int>(host_groups_.size());
    std::vector<int> need(std::max(ng, 1), 0);
    check_hip(hipMemcpy(need.data(), d_.group_need, sizeof(int) * need.size(), hipMemcpyDeviceToHost), "need");
    bool grew = false;
    for (int g = 0; g < ng; ++g) {
        if (need[g] <= host_groups_[g].stride) continue;
        const long long w = (static_cast<long long>(need[g] * 1.25) + 8 + 3) & ~3LL;
        if (w > (1 << 24)) throw Error(OCM_ERR_OOM, "a neighbor-list row exceeds 2^24 entries");
        host_groups_[g].stride = static_cast<int>(w);
        grew = true;
    }
    if (!grew) throw Error(OCM_ERR_HIP, "list overflow reported without a longer row");
    upload(d_.groups, host_groups_);
    check_hip(hipMemset(d_.group_need, 0, sizeof(int) * std::max(ng, 1)), "memset");
    layout_rows();
    drop_graphs();
}

// Device error words are cleared after every read.  Flags raised by a force
// evaluation outside a chunk (the one that starts stepping) are carried on the
// host and reported with the next accepted chunk.
void Engine::carry_errors() {
    sync();
    std::vector<int> err(d_.nsys);
    check_hip(hipMemcpy(err.data(), d_.err, sizeof(int) * d_.nsys, hipMemcpyDeviceToHost), "err");
    for (int s = 0; s < d_.nsys; ++s) err_carry_[s] |= err[s] & kErrNumeric;
    check_hip(hipMemsetAsync(d_.err, 0, sizeof(int) * d_.nsys, stream_), "memset");
}

void Engine::ensure_forces() {
    const TraceRange trace("ocm.forces");
    if (!forces_dirty_) return;
    for (int attempt = 0; attempt < 8; ++attempt) {
        check_hip(hipMemsetAsync(d_.flags, 0, sizeof(int), stream_), "memset");
        check_hip(hipMemsetAsync(d_.err, 0, sizeof(int) * d_.nsys, stream_), "memset");
        launch_zero_forces(d_, d_.f, stream_);
        enqueue_forces(d_.f, true, 0xFFFFFFFFu);
        sync();
        int flags = 0;
        check_hip(hipMemcpy(&flags, d_.flags, sizeof(int), hipMemcpyDeviceToHost), "flags");
        if (flags & kFlagListOverflow) {
            grow_rows();
            continue;
        }
        forces_dirty_ = false;
        // built and pruned outside a chunk (the prune positions are not
        // recorded there): the next chunk starts with a rebuild
        stale_lists();
        if (!team_tuned_) {
            autotune_team();   // clears the flags its scratch launches raised
            launch_zero_forces(d_, d_.f, stream_);
            enqueue_forces(d_.f, false, 0xFFFFFFFFu);   // again, with the tuned launch shape
        }
        carry_errors();
        return;
    }
    throw Error(OCM_ERR_OOM, "neighbor list did not fit after repeated growth");
}

// Project autotuner for the pair kernel's launch shape: lanes per row (team)
// and the number of workgroups of the fixed grid.  Forces are fixed-point
// sums, so every shape gives identical bits; the fastest on the current list
// wins.
void Engine::autotune_team() {
    team_tuned_ = true;
    if (d_.nm == 0) return;
    hipEvent_t a, b;
    check_hip(hipEventCreate(&a), "event");
    check_hip(hipEventCreate(&b), "event");
    float best = 1e30f;
    int best_team = d_.team, best_grid = d_.pair_grid, best_u = d_.pair_u;
    // (lanes per row, entries per lane per round)
    const int shapes[7][2] = {{8, 4}, {16, 4}, {32, 4}, {64, 4}, {8, 2}, {16, 2}, {32, 2}};
    const int grids[4] = {1024, 2048, 4096, 1 << 30};
    for (const auto& sh : shapes) {
        const int t = sh[0];
        build_pair_blocks(t);
        d_.pair_u = sh[1];
        for (int g : grids) {
            d_.pair_grid = g;
            // the inner list (pruned by the evaluation before) is what most steps read
            const int list = prune_ ? kListInner : kListOuter;
            launch_pair(d_, stream_, false, d_.f2, 0xFFFFFFFFu, list);  // warm-up into the scratch buffer
            check_hip(hipEventRecord(a, stream_), "event");
            for (int r = 0; r < 3; ++r) launch_pair(d_, stream_, false, d_.f2, 0xFFFFFFFFu, list);
            check_hip(hipEventRecord(b, stream_), "event");
            check_hip(hipEventSynchronize(b), "event");
            float ms = 0;
            check_hip(hipEventElapsedTime(&ms, a, b), "event");
            if (ms < best) { best = ms; best_team = t; best_grid = g; best_u = sh[1]; }
            if (g >= d_.n_pair_blocks) break;
        }
    }
    // the prune's team, on the same list (it rewrites the inner rows with the
    // same content each time)
    if (d_.col_in) {
        float best_p = 1e30f;
        int best_pt = d_.prune_team;
        for (int t : {16, 32, 64}) {
            d_.prune_team = t;
            launch_prune(d_, stream_);
            check_hip(hipEventRecord(a, stream_), "event");
            for (int r = 0; r < 3; ++r) launch_prune(d_, stream_);
            check_hip(hipEventRecord(b, stream_), "event");
            check_hip(hipEventSynchronize(b), "event");
            float ms = 0;
            check_hip(hipEventElapsedTime(&ms, a, b), "event");
            if (ms < best_p) { best_p = ms; best_pt = t; }
        }
        d_.prune_team = best_pt;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    build_pair_blocks(best_team);
    d_.pair_grid = best_grid;
    d_.pair_u = best_u;
    // the tuning launches may have flagged numeric trouble in the scratch pass only
    check_hip(hipMemsetAsync(d_.err, 0, sizeof(int) * d_.nsys, stream_), "memset");
    sync();
}

// The captured graph of a chunk plan, or nullptr when it runs eagerly.
// Graphs are kept per plan key (length, intervals, starting ages, stamping),
// up to kMaxGraphs; past that the least recently used one is dropped.
// Timing mode 1 (events around every launch) runs eagerly.
hipGraphExec_t Engine::chunk_graph(const ChunkPlan& p, bool may_capture) {
    constexpr std::size_t kMaxGraphs = 48;
    if (!opt_.use_graph || timing_) return nullptr;
    const auto key = p.key(timing_mode_ == 2 ? 1 : 0);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        if (!may_capture) return nullptr;
        if (graphs_.size() >= kMaxGraphs) {
            auto old = graphs_.begin();
            for (auto g = graphs_.begin(); g != graphs_.end(); ++g)
                if (g->second.used < old->second.used) old = g;
            (void)hipGraphExecDestroy(old->second.exec);   // the stream is idle between chunks
            graphs_.erase(old);
        }
        const TraceRange trace("ocm.capture");
        hipGraph_t g;
        check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
        enqueue_chunk(p);
        check_hip(hipStreamEndCapture(stream_, &g), "capture");
        hipGraphExec_t ge;
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        check_hip(e, "instantiate");
        it = graphs_.emplace(key, GraphEntry{ge, 0}).first;
    }
    it->second.used = ++graph_clock_;
    return it->second.exec;
}

bool Engine::needs_prepare(long long n) const {
    if (n <= 0) return false;
    if (forces_dirty_) return true;
    if (d_.res_n > 0 || !opt_.use_graph || timing_) return false;
    if (timing_mode_ == 2) return true;   // the stamp buffers may need sizing
    int age = age_, page = page_, fresh = fresh_;
    long long remaining = n;
    while (remaining > 0) {
        const int k = static_cast<int>(std::min<long long>(remaining, chunk_));
        const ChunkPlan p = plan_chunk_at(k, age, page, fresh);
        if (graphs_.find(p.key(0)) == graphs_.end()) return true;
        age = p.age1;
        page = p.page1;
        fresh = 0;
        remaining -= k;
    }
    return false;
}

void Engine::prepare(long long n) {
    const TraceRange trace("ocm.prepare");
    if (n < 0) invalid("negative step count");
    if (n == 0) return;
    ensure_forces();
    if (d_.res_n > 0) return;   // resident chunks capture nothing
    if (timing_mode_ == 2) ensure_stamps();
    // walk the chunks step(n) would run, with the ages they would leave
    const int age = age_, page = page_, fresh = fresh_;
    long long remaining = n;
    while (remaining > 0) {
        const int k = static_cast<int>(std::min<long long>(remaining, chunk_));
        const ChunkPlan p = plan_chunk(k);
        (void)chunk_graph(p, true);
        age_ = p.age1;
        page_ = p.page1;
        fresh_ = 0;
        remaining -= k;
    }
    age_ = age;
    page_ = page;
    fresh_ = fresh;
}

int Engine::run_chunk(int k) {
    const TraceRange trace("ocm.chunk");
    for (int attempt = 0; attempt < 16; ++attempt) {
        if (timing_mode_ == 2) ensure_stamps();
        const ChunkPlan p = plan_chunk(k);
        hipGraphExec_t ge = chunk_graph(p, capture_in_step_);
        if (ge) check_hip(hipGraphLaunch(ge, stream_), "hipGraphLaunch");
        else enqueue_chunk(p);
        rebuilds_ += p.rebuilds;
        sync();
        // the chunk's last kernel left the flag word and the error words here
        const int flags = status_host_[0];
        const int* err = status_host_ + 1;
        if (flags & kFlagListOverflow) {
            restore();
            grow_rows();
            continue;
        }
        bool stale = false, numeric = false, pruned_stale = false;
        for (int s = 0; s < d_.nsys; ++s) {
            if ((err[s] & kErrNumeric) || sys_status_[s] != OCM_OK) continue;
            if (err[s] & kErrStale) stale = true;
            if (err[s] & kErrPrune) pruned_stale = true;
        }
        if (stale && p.interval > 1) {
            // a bead moved more than skin/2 since the list was built: redo the
            // chunk with half the interval, and wait twice as long as last time
            // before trying a longer one again
            restore();
            interval_ = std::max(1, p.interval / 2);
            clean_chunks_ = 0;
            recover_after_ = std::min(recover_after_ * 2, 64);
            continue;
        }
        if (pruned_stale && p.prune > 1) {
            // the same for the inner list: a bead moved more than margin/2
            // since the prune; below 2 the inner list is not used
            restore();
            prune_ = p.prune / 2 > 1 ? p.prune / 2 : 1;
            prune_clean_ = 0;
            prune_recover_after_ = std::min(prune_recover_after_ * 2, 64);
            continue;
        }
        // accepted: report this chunk's flags and those carried from before it
        age_ = p.age1;
        page_ = p.page1;
        fresh_ = 0;
        for (int s = 0; s < d_.nsys; ++s) {
            if ((err[s] | err_carry_[s]) & kErrNumeric) { sys_status_[s] = OCM_ERR_NUMERIC; numeric = true; }
            err_carry_[s] = 0;
        }
        if (timing_mode_ == 2) {
            for (int q = 0; q < k; ++q) {
                const unsigned long long a = stamp_host_[2 * q], b = stamp_host_[2 * q + 1];
                if (a == ~0ull || b < a) continue;  // no block ran
                timing_slots_[5].ms += static_cast<double>(b - a) / wall_khz_;
                timing_slots_[5].launches += 1;
            }
        }
        if (interval_ < opt_.rebuild_interval && ++clean_chunks_ >= recover_after_) {
            interval_ = std::min(opt_.rebuild_interval, interval_ * 2);
            clean_chunks_ = 0;
            recover_after_ = std::max(4, recover_after_ / 2);
        }
        if (d_.col_in && prune_ < opt_.prune_interval && ++prune_clean_ >= prune_recover_after_) {
            prune_ = std::min(opt_.prune_interval, prune_ * 2);
            prune_clean_ = 0;
            prune_recover_after_ = std::max(4, prune_recover_after_ / 2);
        }
        return numeric ? OCM_ERR_NUMERIC : OCM_OK;
    }
    throw Error(OCM_ERR_HIP, "step chunk could not complete");
}

// Resident stepping: one launch per chunk runs every system's steps on chip;
// nothing rolls back (each workgroup rebuilds its own list when a bead has
// moved more than skin/2), so a chunk only reports numeric errors.
int Engine::step_resident(long long n) {
    const TraceRange trace("ocm.resident_steps");
    int status = OCM_OK;
    long long remaining = n;
    while (remaining > 0) {
        const int k = static_cast<int>(std::min<long long>(remaining, kResChunk));
        launch_resident(d_, stream_, k);
        launch_status_out(d_, stream_, status_host_);
        sync();
        const int* err = status_host_ + 1;
        for (int s = 0; s < d_.nsys; ++s) {
            if ((err[s] | err_carry_[s]) & kErrNumeric) {
                sys_status_[s] = OCM_ERR_NUMERIC;
                status = OCM_ERR_NUMERIC;
            }
            err_carry_[s] = 0;
        }
        remaining -= k;
    }
    // the standard path's lists and records no longer match the positions
    stale_lists();
    lists_stale_ = true;
    return status;
}

int Engine::step(long long n) {
    const TraceRange trace("ocm.step");
    if (n < 0) invalid("negative step count");
    if (n == 0) return OCM_OK;
    ensure_forces();
    if (d_.res_n > 0) return step_resident(n);
    int status = OCM_OK;
    long long remaining = n;
    while (remaining > 0) {
        const int k = static_cast<int>(std::min<long long>(remaining, chunk_));
        const int st = run_chunk(k);
        if (st != OCM_OK) status = st;
        remaining -= k;
    }
    return status;
}

// Potential energy of every system at the current positions, by the FP64
// energy pass over the list the last force evaluation built.
std::vector<double> Engine::potentials() {
    if (lists_stale_ && d_.nm > 0) {
        launch_rebuild(d_, stream_);
        ++rebuilds_;
        stale_lists();
    }
    lists_stale_ = false;
    launch_zero_energy(d_, stream_);
    launch_bonded(d_, stream_, true, d_.f2, 0xFFFFFFFFu);
    launch_pair(d_, stream_, true, d_.f2, 0xFFFFFFFFu);
    sync();
    check_hip(hipMemsetAsync(d_.err, 0, sizeof(int) * d_.nsys, stream_), "memset");
    std::vector<long long> e(static_cast<size_t>(d_.nsys) * 32);
    check_hip(hipMemcpy(e.data(), d_.energy, e.size() * sizeof(long long), hipMemcpyDeviceToHost), "energy");
    std::vector<double> out(d_.nsys, 0.0);
    for (int s = 0; s < d_.nsys; ++s)
        for (int g = 0; g < 32; ++g) out[s] += e[static_cast<size_t>(s) * 32 + g] * kFixInv;
    return out;
}

// LocalEnergyMinimizer::minimize (src/Simulator.cpp:42-46) is OpenMM's
// L-BFGS; this is an L-BFGS of its own over the coordinates of all mobile
// beads of the batch (the systems are independent, so their summed energy
// has the same minima), with device forces and FP64 device energies:
//   - direction from the last kHistory (s, y) pairs (two-loop recursion),
//     steepest descent when it is not a descent direction;
//   - the step capped so that no bead moves more than kMaxMove per
//     iteration (starts with overlapping beads have enormous gradients);
//   - backtracking line search on the sufficient-decrease condition;
//   - stop when the RMS force component of every system is <= tol (the
//     reference's criterion) or after max_iter iterations.
int Engine::minimize(double tol, int max_iter) {
    const TraceRange trace("ocm.minimize");
    if (!(tol >= 0.0)) invalid("tolerance must be >= 0");
    if (max_iter < 0) invalid("max_iterations must be >= 0");
    if (max_iter == 0) max_iter = 100000;
    constexpr int kHistory = 10;
    constexpr double kMaxMove = 0.05;   // nm per bead per iteration
    const size_t nb = static_cast<size_t>(d_.nb), n3 = 3 * nb;
    std::vector<char> mobile(nb, 0);
    for (int s = 0; s < d_.nsys; ++s)
        for (int i = 0; i < sys_[s].n; ++i) mobile[bead_begin_[s] + i] = sys_[s].masses[i] != 0.0;
    std::vector<long long> fbuf(n3);

    // energy and gradient (-force; 0 for fixed beads) at x, device left at x
    auto evaluate = [&](const std::vector<double>& x, std::vector<double>& g) {
        check_hip(hipMemcpy(d_.px, x.data(), n3 * sizeof(double), hipMemcpyHostToDevice), "positions");
        forces_dirty_ = true;
        ensure_forces();
        check_hip(hipMemcpy(fbuf.data(), d_.f.fx, n3 * sizeof(long long), hipMemcpyDeviceToHost), "forces");
        const std::vector<double> pe = potentials();
        for (size_t k = 0; k < n3; ++k) g[k] = mobile[k % nb] ? -(fbuf[k] * kFixInv) : 0.0;
        double e = 0;
        for (double v : pe) e += v;
        return e;
    };
    auto converged = [&](const std::vector<double>& g) {
        for (int s = 0; s < d_.nsys; ++s) {
            const HostSystem& H = sys_[s];
            double ss = 0.0;
            for (int i = 0; i < H.n; ++i) {
                const size_t b = bead_begin_[s] + i;
                for (int c = 0; c < 3; ++c) ss += g[c * nb + b] * g[c * nb + b];
            }
            if (H.n > 0 && std::sqrt(ss / (3.0 * H.n)) > tol) return false;
        }
        return true;
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double r = 0;
        for (size_t k = 0; k < n3; ++k) r += a[k] * b[k];
        return r;
    };

    std::vector<double> x(n3), g(n3), d(n3), xn(n3), gn(n3), q(n3);
    sync();
    check_hip(hipMemcpy(x.data(), d_.px, n3 * sizeof(double), hipMemcpyDeviceToHost), "positions");
    double e = evaluate(x, g);
    std::vector<std::vector<double>> S, Y;
    std::vector<double> rho;
    int it = 0;
    while (!converged(g) && it < max_iter) {
        // two-loop recursion: d = -H g
        q = g;
        std::vector<double> alpha(S.size());
        for (int k = static_cast<int>(S.size()) - 1; k >= 0; --k) {
            alpha[k] = rho[k] * dot(S[k], q);
            for (size_t j = 0; j < n3; ++j) q[j] -= alpha[k] * Y[k][j];
        }
        const double gamma = S.empty() ? 1.0 : dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
        for (size_t j = 0; j < n3; ++j) q[j] *= gamma;
        for (size_t k = 0; k < S.size(); ++k) {
            const double beta = rho[k] * dot(Y[k], q);
            for (size_t j = 0; j < n3; ++j) q[j] += (alpha[k] - beta) * S[k][j];
        }
        for (size_t j = 0; j < n3; ++j) d[j] = -q[j];
        double gd = dot(g, d);
        if (!(gd < 0.0)) {   // not a descent direction: restart from steepest descent
            S.clear(); Y.clear(); rho.clear();
            for (size_t j = 0; j < n3; ++j) d[j] = -g[j];
            gd = dot(g, d);
        }
        // cap the largest bead displacement
        double dmax2 = 0.0;
        for (size_t b = 0; b < nb; ++b)
            dmax2 = std::max(dmax2, d[b] * d[b] + d[nb + b] * d[nb + b] + d[2 * nb + b] * d[2 * nb + b]);
        double t = dmax2 > kMaxMove * kMaxMove ? kMaxMove / std::sqrt(dmax2) : 1.0;
        bool accepted = false;
        double en = e;
        for (int ls = 0; ls < 30; ++ls) {
            for (size_t j = 0; j < n3; ++j) xn[j] = x[j] + t * d[j];
            en = evaluate(xn, gn);
            if (std::isfinite(en) && en <= e + 1e-4 * t * gd) { accepted = true; break; }
            t *= 0.5;
        }
        ++it;
        if (!accepted) {
            // no decrease along d even for tiny steps: restart the history, or
            // stop if this already was steepest descent
            check_hip(hipMemcpy(d_.px, x.data(), n3 * sizeof(double), hipMemcpyHostToDevice), "positions");
            if (S.empty()) break;
            S.clear(); Y.clear(); rho.clear();
            continue;
        }
        std::vector<double> sv(n3), yv(n3);
        for (size_t j = 0; j < n3; ++j) { sv[j] = xn[j] - x[j]; yv[j] = gn[j] - g[j]; }
        const double sy = dot(sv, yv);
        if (sy > 1e-12 * std::sqrt(dot(sv, sv) * dot(yv, yv))) {
            if (static_cast<int>(S.size()) == kHistory) { S.erase(S.begin()); Y.erase(Y.begin()); rho.erase(rho.begin()); }
            S.push_back(std::move(sv));
            Y.push_back(std::move(yv));
            rho.push_back(1.0 / sy);
        }
        x.swap(xn);
        g.swap(gn);
        e = en;
    }
    // the device holds the last evaluated point; leave it at the accepted one
    check_hip(hipMemcpy(d_.px, x.data(), n3 * sizeof(double), hipMemcpyHostToDevice), "positions");
    // relaxation may pass through flagged overlaps; stepping starts from a
    // fresh evaluation at the relaxed positions
    check_hip(hipMemset(d_.err, 0, sizeof(int) * d_.nsys), "memset");
    std::fill(err_carry_.begin(), err_carry_.end(), 0);
    forces_dirty_ = true;
    return it;
}

void Engine::enqueue_position_copy(double* dst) {
    const size_t bytes = static_cast<size_t>(d_.nb) * sizeof(double);
    check_hip(hipMemcpyAsync(dst, d_.px, bytes, hipMemcpyDeviceToDevice, stream_), "position copy");
    check_hip(hipMemcpyAsync(dst + d_.nb, d_.py, bytes, hipMemcpyDeviceToDevice, stream_), "position copy");
    check_hip(hipMemcpyAsync(dst + 2 * static_cast<size_t>(d_.nb), d_.pz, bytes, hipMemcpyDeviceToDevice, stream_),
              "position copy");
}

void Engine::set_steps_done(int s, long long steps_done) {
    sync();
    long long gs[2] = {0, 0};
    check_hip(hipMemcpy(gs, d_.gstep, sizeof(gs), hipMemcpyDeviceToHost), "gstep");
    sys_step_origin_.at(s) = gs[0] - steps_done;
    DevSys D;
    check_hip(hipMemcpy(&D, d_.sys + s, sizeof(DevSys), hipMemcpyDeviceToHost), "sys");
    D.step_origin = sys_step_origin_[s];
    check_hip(hipMemcpy(d_.sys + s, &D, sizeof(DevSys), hipMemcpyHostToDevice), "sys");
    if (s == 0) d_.sys0 = D;
    drop_graphs();   // captured integrator launches carry system 0's record by value
}

// ------------------------------------------------------------ state I/O
void Engine::set_positions(int s, const double* xyz) {
    const HostSystem& H = sys_.at(s);
    std::vector<double> x(H.n), y(H.n), z(H.n);
    for (int i = 0; i < H.n; ++i) {
        x[i] = xyz[3 * i]; y[i] = xyz[3 * i + 1]; z[i] = xyz[3 * i + 2];
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i])) invalid("position is not finite");
    }
    const size_t off = bead_begin_[s];
    check_hip(hipMemcpy(d_.px + off, x.data(), H.n * sizeof(double), hipMemcpyHostToDevice), "positions");
    check_hip(hipMemcpy(d_.py + off, y.data(), H.n * sizeof(double), hipMemcpyHostToDevice), "positions");
    check_hip(hipMemcpy(d_.pz + off, z.data(), H.n * sizeof(double), hipMemcpyHostToDevice), "positions");
    forces_dirty_ = true;
    sys_status_[s] = OCM_OK;
}

void Engine::set_velocities(int s, const double* xyz) {
    const HostSystem& H = sys_.at(s);
    std::vector<double> x(H.n), y(H.n), z(H.n);
    for (int i = 0; i < H.n; ++i) { x[i] = xyz[3 * i]; y[i] = xyz[3 * i + 1]; z[i] = xyz[3 * i + 2]; }
    const size_t off = bead_begin_[s];
    check_hip(hipMemcpy(d_.vx + off, x.data(), H.n * sizeof(double), hipMemcpyHostToDevice), "velocities");
    check_hip(hipMemcpy(d_.vy + off, y.data(), H.n * sizeof(double), hipMemcpyHostToDevice), "velocities");
    check_hip(hipMemcpy(d_.vz + off, z.data(), H.n * sizeof(double), hipMemcpyHostToDevice), "velocities");
}

void Engine::set_velocities_to_temperature(int s, double T, uint64_t seed) {
    if (!(T >= 0)) invalid("temperature must be >= 0");
    const double kB = 1.380649e-23 * 6.02214076e23 / 1000.0;
    launch_velocity_noise(d_, stream_, s, kB * T, seed);
    sync();
}

static void to_host3(const double* dx, const double* dy, const double* dz, int n, double* out) {
    std::vector<double> x(n), y(n), z(n);
    (void)hipMemcpy(x.data(), dx, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipMemcpy(y.data(), dy, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipMemcpy(z.data(), dz, n * sizeof(double), hipMemcpyDeviceToHost);
    for (int i = 0; i < n; ++i) { out[3 * i] = x[i]; out[3 * i + 1] = y[i]; out[3 * i + 2] = z[i]; }
}

void Engine::get_state(int s, int what, uint32_t groups, ocm_state* out) {
    const TraceRange trace("ocm.get_state");
    const HostSystem& H = sys_.at(s);
    const size_t off = bead_begin_[s];
    sync();
    if ((what & OCM_STATE_POSITIONS) && out->positions) to_host3(d_.px + off, d_.py + off, d_.pz + off, H.n, out->positions);
    if ((what & OCM_STATE_VELOCITIES) && out->velocities) to_host3(d_.vx + off, d_.vy + off, d_.vz + off, H.n, out->velocities);
    long long gs[2] = {0, 0};
    check_hip(hipMemcpy(gs, d_.gstep, sizeof(gs), hipMemcpyDeviceToHost), "gstep");
    out->step = gs[0] - sys_step_origin_[s];
    if ((what & OCM_STATE_FORCES) && out->forces) {
        const uint32_t present = sys_group_mask_[s];
        ForceBuf src = d_.f;
        ensure_forces();   // builds (and if needed widens) the list for the current positions
        if ((present & groups) != present) {
            launch_zero_forces(d_, d_.f2, stream_);
            enqueue_forces(d_.f2, false, groups);
            src = d_.f2;
        }
        sync();
        // a read of forces raises no step error (the full evaluation's flags are carried)
        check_hip(hipMemsetAsync(d_.err, 0, sizeof(int) * d_.nsys, stream_), "memset");
        std::vector<long long> fx(H.n), fy(H.n), fz(H.n);
        check_hip(hipMemcpy(fx.data(), src.fx + off, H.n * sizeof(long long), hipMemcpyDeviceToHost), "forces");
        check_hip(hipMemcpy(fy.data(), src.fy + off, H.n * sizeof(long long), hipMemcpyDeviceToHost), "forces");
        check_hip(hipMemcpy(fz.data(), src.fz + off, H.n * sizeof(long long), hipMemcpyDeviceToHost), "forces");
        for (int i = 0; i < H.n; ++i) {
            out->forces[3 * i] = fx[i] * kFixInv;
            out->forces[3 * i + 1] = fy[i] * kFixInv;
            out->forces[3 * i + 2] = fz[i] * kFixInv;
        }
    }
    if (what & OCM_STATE_ENERGY) {
        launch_zero_energy(d_, stream_);
        if (d_.nm > 0) { launch_rebuild(d_, stream_); ++rebuilds_; stale_lists(); }
        launch_bonded(d_, stream_, true, d_.f2, groups);
        launch_pair(d_, stream_, true, d_.f2, groups);
        launch_kinetic(d_, stream_);
        sync();
        check_hip(hipMemsetAsync(d_.err, 0, sizeof(int) * d_.nsys, stream_), "memset");   // energies raise no step error
        int flags = 0;
        check_hip(hipMemcpy(&flags, d_.flags, sizeof(int), hipMemcpyDeviceToHost), "flags");
        if (flags & kFlagListOverflow) {
            grow_rows();
            check_hip(hipMemset(d_.flags, 0, sizeof(int)), "memset");
            get_state(s, OCM_STATE_ENERGY, groups, out);
            return;
        }
        std::vector<long long> e(32);
        long long ke = 0;
        check_hip(hipMemcpy(e.data(), d_.energy + static_cast<size_t>(s) * 32, 32 * sizeof(long long), hipMemcpyDeviceToHost), "energy");
        check_hip(hipMemcpy(&ke, d_.ke + s, sizeof(long long), hipMemcpyDeviceToHost), "ke");
        out->potential_energy = 0;
        for (int g = 0; g < 32; ++g) {
            out->energy_by_group[g] = e[g] * kFixInv;
            if ((groups >> g) & 1u) out->potential_energy += out->energy_by_group[g];
        }
        out->kinetic_energy = ke * kFixInv;
    }
}

// Pair group of nonbonded force `force_index` of system s.
int Engine::pair_group_of(int s, int force_index) const {
    const HostSystem& H = sys_.at(s);
    if (force_index < 0 || force_index >= static_cast<int>(H.forces.size())) invalid("force index out of range");
    const int g = force_group_of_[s][force_index];
    if (g < 0) invalid("force is not a nonbonded force");
    return g;
}

// Rebuilds the outer list at the current positions; returns the flag word.
int Engine::rebuild_now() {
    lists_stale_ = false;
    launch_fill_i32(stream_, d_.flags, 0, 1);
    launch_rebuild(d_, stream_);
    ++rebuilds_;
    stale_lists();
    sync();
    int flags = 0;
    check_hip(hipMemcpy(&flags, d_.flags, sizeof(int), hipMemcpyDeviceToHost), "flags");
    return flags;
}

long long Engine::neighbor_pairs(int s, int force_index, int32_t* out, long long cap) {
    const int g = pair_group_of(s, force_index);
    ensure_forces();
    // the list at the current positions, widened if a row no longer fits
    for (int attempt = 0;; ++attempt) {
        if (!(rebuild_now() & kFlagListOverflow)) break;
        if (attempt == 8) throw Error(OCM_ERR_OOM, "neighbor list did not fit after repeated growth");
        grow_rows();
    }
    // count, then dump every pair, so the first `cap` of the sorted set are returned
    long long want = std::max<long long>(cap, 1024);
    std::vector<int2> host;
    int cnt = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int2* buf = dalloc<int2>(static_cast<size_t>(want));
        launch_neighbor_dump(d_, stream_, g, want, group_member_count_[g], d_.dump_count, buf);
        sync();
        check_hip(hipMemcpy(&cnt, d_.dump_count, sizeof(int), hipMemcpyDeviceToHost), "dump");
        const bool fits = cnt <= want;
        if (fits) {
            host.resize(static_cast<size_t>(cnt));
            if (cnt > 0) check_hip(hipMemcpy(host.data(), buf, cnt * sizeof(int2), hipMemcpyDeviceToHost), "dump");
        }
        auto it = std::find(allocs_.begin(), allocs_.end(), static_cast<void*>(buf));
        if (it != allocs_.end()) { (void)hipFree(*it); allocs_.erase(it); }
        if (fits) break;
        want = cnt;
    }
    std::vector<std::pair<int, int>> v;
    v.reserve(host.size());
    for (const int2& p : host) v.emplace_back(p.x - bead_begin_[s], p.y - bead_begin_[s]);
    std::sort(v.begin(), v.end());
    const long long ncopy = std::min<long long>(cnt, cap);
    for (long long k = 0; k < ncopy; ++k) { out[2 * k] = v[k].first; out[2 * k + 1] = v[k].second; }
    return cnt;
}

static_assert(kNoAllocMaxPairs == OCM_NOALLOC_MAX_PAIRS, "scratch size and the header's limit must agree");

void Engine::neighbor_pairs_noalloc(int s, int force_index, int32_t* out, long long cap, long long* count) {
    const int g = pair_group_of(s, force_index);
    check_hip(hipSetDevice(opt_.device), "hipSetDevice");
    int* dcount = nullptr;
    int2* dpairs = nullptr;
    check_hip(noalloc_scratch(opt_.device, &dcount, &dpairs), "hipGetSymbolAddress(scratch)");
    // the list at the current positions, in the rows it has
    if (rebuild_now() & kFlagListOverflow)
        throw Error(OCM_ERR_OOM, "a neighbor-list row no longer fits its slots and this call does not widen "
                                 "them; ocm_step or ocm_get_neighbor_pairs widens the rows");
    launch_neighbor_dump(d_, stream_, g, kNoAllocMaxPairs, group_member_count_[g], dcount, dpairs);
    sync();
    int cnt = 0;
    check_hip(hipMemcpy(&cnt, dcount, sizeof(int), hipMemcpyDeviceToHost), "dump");
    *count = cnt;
    if (cnt > kNoAllocMaxPairs)
        throw Error(OCM_ERR_INVALID, "the force has " + std::to_string(cnt) + " pairs, more than the " +
                                         std::to_string(kNoAllocMaxPairs) +
                                         " the no-allocation dump holds; use ocm_get_neighbor_pairs");
    if (cap == 0) return;
    if (cap < cnt)
        throw Error(OCM_ERR_INVALID, "cap " + std::to_string(cap) + " is smaller than the pair count " +
                                         std::to_string(cnt));
    if (cnt == 0) return;
    check_hip(hipMemcpy(out, dpairs, static_cast<size_t>(cnt) * sizeof(int2), hipMemcpyDeviceToHost), "dump");
    // local indices, sorted in place in the caller's array
    struct Pair {
        int32_t i, j;
    };
    Pair* v = reinterpret_cast<Pair*>(out);
    const int base = bead_begin_[s];
    for (int k = 0; k < cnt; ++k) {
        v[k].i -= base;
        v[k].j -= base;
    }
    std::sort(v, v + cnt, [](const Pair& a, const Pair& b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
}

void Engine::inner_stats(long long* entries, int* prune) {
    sync();
    long long total = 0;
    if (d_.row_count_in) {
        std::vector<int> cnt(d_.nm, 0);
        check_hip(hipMemcpy(cnt.data(), d_.row_count_in, sizeof(int) * d_.nm, hipMemcpyDeviceToHost), "rows");
        for (int c : cnt) total += c;
    }
    *entries = total;
    *prune = d_.col_in ? prune_ : 0;
}

// Exact counts of both lists: entries that hold a partner (each pair appears
// once in each of its two rows) and entries including the -1 padding.
void Engine::list_counts(long long out[4]) {
    sync();
    for (int k = 0; k < 4; ++k) out[k] = 0;
    if (d_.nm == 0) return;
    std::vector<int> start(d_.nm), cnt(d_.nm);
    std::vector<int> col(static_cast<size_t>(d_.list_capacity));
    check_hip(hipMemcpy(start.data(), d_.row_start, sizeof(int) * d_.nm, hipMemcpyDeviceToHost), "rows");
    for (int inner = 0; inner < 2; ++inner) {
        const int* rc = inner ? d_.row_count_in : d_.row_count;
        const int* cl = inner ? d_.col_in : d_.col;
        if (!rc || !cl) continue;
        check_hip(hipMemcpy(cnt.data(), rc, sizeof(int) * d_.nm, hipMemcpyDeviceToHost), "rows");
        check_hip(hipMemcpy(col.data(), cl, sizeof(int) * col.size(), hipMemcpyDeviceToHost), "list");
        for (int r = 0; r < d_.nm; ++r) {
            out[2 * inner + 1] += cnt[r];
            for (int e = 0; e < cnt[r]; ++e) out[2 * inner] += col[static_cast<size_t>(start[r]) + e] >= 0;
        }
    }
}

void Engine::launch_shape(int* team, int* u, int* grid, int* prune_team) {
    ensure_forces();   // the shape is tuned at the first force evaluation
    *team = d_.team;
    *u = d_.pair_u;
    *grid = pair_grid_size(d_);
    *prune_team = d_.col_in ? d_.prune_team : 0;
}

void Engine::list_stats(long long* entries, long long* members, long long* rebuilds) {
    sync();
    std::vector<int> cnt(std::max(d_.nm, 1), 0);
    if (d_.nm > 0) check_hip(hipMemcpy(cnt.data(), d_.row_count, sizeof(int) * d_.nm, hipMemcpyDeviceToHost), "rows");
    long long total = 0;
    for (int k = 0; k < d_.nm; ++k) total += cnt[k];
    if (entries) *entries = total;
    if (members) *members = d_.nm;
    if (rebuilds) *rebuilds = rebuilds_;
}

}  // namespace ocm
