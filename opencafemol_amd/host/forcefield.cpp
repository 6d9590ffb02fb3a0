// forcefield.cpp -- generators, topology, system and integrator assembly.
#include "forcefield.hpp"

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <random>
#include <sstream>

namespace ocmhost {

namespace constant {
double eps0() { return 8.854187817e-12 * 1.0e3 * 1.0e-9 / Na; }
double cafetime() { return std::sqrt(1.0 / kj_per_kcal) * 0.1; }
}  // namespace constant

// ---------------------------------------------------------------- Topology
void Topology::link(std::size_t a, std::size_t b, const std::string& type) {
    nodes_.at(a).emplace_back(b, type);
    nodes_.at(b).emplace_back(a, type);
    molecule_updated_ = false;
}

void Topology::add_edges(const std::vector<index_pair>& edges, const std::string& type) {
    for (const auto& e : edges) link(e.first, e.second, type);
}

// Molecules as connected components over `type` edges, found depth first
// from the lowest unvisited bead (src/Topology.cpp:7-48).
void Topology::make_molecule(const std::string& type) {
    molecules_.clear();
    std::vector<bool> seen(nodes_.size(), false);
    for (std::size_t start = 0; start < nodes_.size(); ++start) {
        if (seen[start]) continue;
        molecules_.emplace_back();
        std::vector<std::size_t> stack{start};
        while (!stack.empty()) {
            const std::size_t v = stack.back();
            stack.pop_back();
            if (seen[v]) continue;
            seen[v] = true;
            molecules_.back().push_back(v);
            for (const auto& e : nodes_[v])
                if (e.second == type && !seen[e.first]) stack.push_back(e.first);
        }
    }
    molecule_updated_ = true;
}

index_pairs_type Topology::ignore_list_within_molecule() const {
    if (!molecule_updated_)
        throw std::runtime_error("[error] ignore_list_within_molecule called before make_molecule");
    index_pairs_type out;
    for (const auto& mol : molecules_)
        for (std::size_t a = 0; a + 1 < mol.size(); ++a)
            for (std::size_t b = a + 1; b < mol.size(); ++b) out.emplace_back(mol[a], mol[b]);
    return out;
}

// Depth-first walk that stops `dist` edges out and never re-enters a bead it
// already listed; the same walk as Topology::list_adjacent_within
// (src/Topology.cpp:114-135), so graphs with cycles give the same sets.
void Topology::adjacent_within(std::size_t node, std::size_t dist, const std::string& type,
                               std::vector<std::size_t>& out) const {
    out.push_back(node);
    if (dist == 0) return;
    for (const auto& e : nodes_[node]) {
        if (e.second != type) continue;
        if (std::find(out.begin(), out.end(), e.first) == out.end()) adjacent_within(e.first, dist - 1, type, out);
    }
}

index_pairs_type Topology::ignore_list_within_edge(std::size_t dist, const std::string& type) const {
    index_pairs_type out;
    for (std::size_t i = 0; i < nodes_.size(); ++i) {
        std::vector<std::size_t> near;
        adjacent_within(i, dist, type, near);
        for (std::size_t j : near)
            if (j != i) out.emplace_back(std::min(i, j), std::max(i, j));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

// -------------------------------------------------------- descriptors
ocm_force_desc ForceDescriptor::view() const {
    ocm_force_desc d{};
    d.kind = kind;
    d.group = group;
    d.use_periodic = use_periodic ? 1 : 0;
    d.n_terms = n_terms;
    d.indices = indices.empty() ? nullptr : indices.data();
    d.params = params.empty() ? nullptr : params.data();
    d.n_globals = static_cast<int32_t>(globals.size());
    d.globals = globals.empty() ? nullptr : globals.data();
    d.particle_params = particle_params.empty() ? nullptr : particle_params.data();
    d.category = category.empty() ? nullptr : category.data();
    d.n_categories = n_categories;
    d.category_allow = allow.empty() ? nullptr : allow.data();
    d.cutoff = cutoff;
    d.switch_distance = switch_distance;
    d.n_exclusions = static_cast<int64_t>(exclusions.size() / 2);
    d.exclusions = exclusions.empty() ? nullptr : exclusions.data();
    return d;
}

void interaction_categories(const std::vector<bool>& has_param, const group_pairs_type& ignore_group_pairs,
                            const std::vector<std::optional<std::string>>& group_vec, ForceDescriptor& out) {
    const std::size_t n = has_param.size();
    out.category.clear();
    out.allow.clear();
    if (ignore_group_pairs.empty()) {
        if (std::all_of(has_param.begin(), has_param.end(), [](bool b) { return b; })) {
            out.n_categories = 0;  // no interaction group: every bead takes part
            return;
        }
        out.n_categories = 1;
        out.category.resize(n);
        for (std::size_t i = 0; i < n; ++i) out.category[i] = has_param[i] ? 0 : -1;
        return;
    }
    std::set<std::string> names;
    for (const auto& p : ignore_group_pairs) { names.insert(p.first); names.insert(p.second); }
    const std::vector<std::string> related(names.begin(), names.end());
    if (related.size() + 1 > 32) throw std::runtime_error("[error] more than 31 groups named in ignore.group");
    out.n_categories = static_cast<int>(related.size()) + 1;
    out.category.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        if (!has_param[i]) { out.category[i] = -1; continue; }
        int c = 0;
        if (i < group_vec.size() && group_vec[i]) {
            const auto it = std::find(related.begin(), related.end(), *group_vec[i]);
            if (it != related.end()) c = 1 + static_cast<int>(it - related.begin());
        }
        out.category[i] = c;
    }
    const int nc = out.n_categories;
    out.allow.assign(static_cast<std::size_t>(nc) * nc, 1);
    for (std::size_t a = 0; a < related.size(); ++a)
        for (std::size_t b = a; b < related.size(); ++b) {
            // Utility::contains (src/util/Utility.hpp:163-178) matches a pair in either order
            const bool ignored =
                std::find(ignore_group_pairs.begin(), ignore_group_pairs.end(), std::make_pair(related[a], related[b])) !=
                    ignore_group_pairs.end() ||
                std::find(ignore_group_pairs.begin(), ignore_group_pairs.end(), std::make_pair(related[b], related[a])) !=
                    ignore_group_pairs.end();
            out.allow[(a + 1) * nc + (b + 1)] = out.allow[(b + 1) * nc + (a + 1)] = ignored ? 0 : 1;
        }
}

static void put_exclusions(const index_pairs_type& ignore, ForceDescriptor& d) {
    d.exclusions.reserve(2 * ignore.size());
    for (const auto& p : ignore) {
        d.exclusions.push_back(static_cast<int32_t>(p.first));
        d.exclusions.push_back(static_cast<int32_t>(p.second));
    }
}

template <typename T> static void check_sizes(const char* who, std::size_t n, const std::vector<T>& v, const char* what) {
    if (v.size() != n) {
        std::ostringstream o;
        o << "[error] " << who << ": parameter number of indices_vec (" << n << ") and " << what << " (" << v.size()
          << ") is not matched.";
        throw std::runtime_error(o.str());
    }
}

// ------------------------------------------------------ local generators
HarmonicBondForceFieldGenerator::HarmonicBondForceFieldGenerator(const index_pairs_type& indices,
                                                                 const std::vector<double>& v0s,
                                                                 const std::vector<double>& ks, bool use_periodic)
    : indices_(indices), v0s_(v0s), ks_(ks), use_periodic_(use_periodic) {
    check_sizes("HarmonicBondForceFieldGenerator", indices.size(), v0s, "v0");
    check_sizes("HarmonicBondForceFieldGenerator", indices.size(), ks, "k");
}
ForceDescriptor HarmonicBondForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_HARMONIC_BOND;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        d.indices.push_back(static_cast<int32_t>(indices_[t].first));
        d.indices.push_back(static_cast<int32_t>(indices_[t].second));
        d.params.push_back(v0s_[t]);
        d.params.push_back(ks_[t]);
    }
    return d;
}

GaussianBondForceFieldGenerator::GaussianBondForceFieldGenerator(const index_pairs_type& indices,
                                                                 const std::vector<double>& ks,
                                                                 const std::vector<double>& v0s,
                                                                 const std::vector<double>& sigmas, bool use_periodic)
    : indices_(indices), ks_(ks), v0s_(v0s), sigmas_(sigmas), use_periodic_(use_periodic) {
    check_sizes("GaussianBondForceFieldGenerator", indices.size(), ks, "k");
    check_sizes("GaussianBondForceFieldGenerator", indices.size(), v0s, "v0");
    check_sizes("GaussianBondForceFieldGenerator", indices.size(), sigmas, "sigma");
}
ForceDescriptor GaussianBondForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_GAUSSIAN_BOND;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        d.indices.push_back(static_cast<int32_t>(indices_[t].first));
        d.indices.push_back(static_cast<int32_t>(indices_[t].second));
        d.params.insert(d.params.end(), {ks_[t], v0s_[t], sigmas_[t]});
    }
    return d;
}

GoContactForceFieldGenerator::GoContactForceFieldGenerator(const index_pairs_type& indices,
                                                           const std::vector<double>& ks,
                                                           const std::vector<double>& r0s, bool use_periodic)
    : indices_(indices), ks_(ks), r0s_(r0s), use_periodic_(use_periodic) {
    check_sizes("GoContactForceFieldGenerator", indices.size(), ks, "k");
    check_sizes("GoContactForceFieldGenerator", indices.size(), r0s, "r0");
}
ForceDescriptor GoContactForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_GO_CONTACT;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        d.indices.push_back(static_cast<int32_t>(indices_[t].first));
        d.indices.push_back(static_cast<int32_t>(indices_[t].second));
        d.params.insert(d.params.end(), {ks_[t], r0s_[t]});
    }
    return d;
}

HarmonicAngleForceFieldGenerator::HarmonicAngleForceFieldGenerator(const std::vector<indices_type>& indices,
                                                                   const std::vector<double>& v0s,
                                                                   const std::vector<double>& ks, bool use_periodic)
    : indices_(indices), v0s_(v0s), ks_(ks), use_periodic_(use_periodic) {
    check_sizes("HarmonicAngleForceFieldGenerator", indices.size(), v0s, "v0");
    check_sizes("HarmonicAngleForceFieldGenerator", indices.size(), ks, "k");
}
ForceDescriptor HarmonicAngleForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_HARMONIC_ANGLE;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        for (auto i : indices_[t]) d.indices.push_back(static_cast<int32_t>(i));
        d.params.insert(d.params.end(), {v0s_[t], ks_[t]});
    }
    return d;
}

FlexibleLocalAngleForceFieldGenerator::FlexibleLocalAngleForceFieldGenerator(
    const std::vector<indices_type>& indices, const std::vector<double>& ks, const std::array<double, 10>& table,
    const std::string& aa_name, bool use_periodic)
    : indices_(indices), ks_(ks), table_(table), aa_name_(aa_name), use_periodic_(use_periodic) {
    check_sizes("FlexibleLocalAngleForceFieldGenerator", indices.size(), ks, "k");
    if (!constant::fla_spline_second_derivative_table().count(aa_name))
        throw std::runtime_error("[error] FlexibleLocalAngle: unknown residue " + aa_name);
}

// FlexibleLocalAngleForceFieldGenerator.cpp:11-19,47-68: minimum of the
// reference's spline evaluator (its rounded knot list and y''-table) over a
// sweep from min_theta in 1e-4 steps while below max_theta.
double FlexibleLocalAngleForceFieldGenerator::min_energy() const {
    static const double knots[10] = {1.30900, 1.48353, 1.65806, 1.83260, 2.00713,
                                     2.18166, 2.35619, 2.53073, 2.70526, 2.87979};
    const auto& y2 = constant::fla_spline_second_derivative_table().at(aa_name_);
    const double tmin = constant::fla_min_theta, tmax = constant::fla_max_theta;
    const double dth = (tmax - tmin) / 9.0, rdth = 1.0 / dth;
    double emin = table_[0];
    for (double th = tmin; th < tmax; th += 1e-4) {
        long n = std::min<long>(8, static_cast<long>(std::floor((th - tmin) * rdth)));
        if (knots[n + 1] < th) ++n;
        else if (th < knots[n]) --n;
        const double a = (knots[n + 1] - th) * rdth, b = (th - knots[n]) * rdth;
        const double e = a * table_[n] + b * table_[n + 1] +
                         ((a * a * a - a) * y2[n] + (b * b * b - b) * y2[n + 1]) * dth * dth / 6.0;
        emin = std::min(emin, e);
    }
    return emin;
}

ForceDescriptor FlexibleLocalAngleForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_FLEXIBLE_LOCAL_ANGLE;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        for (auto i : indices_[t]) d.indices.push_back(static_cast<int32_t>(i));
        d.params.push_back(ks_[t]);
    }
    d.globals.assign(table_.begin(), table_.end());
    d.globals.insert(d.globals.end(), {constant::fla_min_theta, constant::fla_max_theta, min_energy()});
    return d;
}

GaussianDihedralForceFieldGenerator::GaussianDihedralForceFieldGenerator(const std::vector<indices_type>& indices,
                                                                         const std::vector<double>& ks,
                                                                         const std::vector<double>& theta0s,
                                                                         const std::vector<double>& sigmas,
                                                                         bool use_periodic)
    : indices_(indices), ks_(ks), theta0s_(theta0s), sigmas_(sigmas), use_periodic_(use_periodic) {
    check_sizes("GaussianDihedralForceFieldGenerator", indices.size(), ks, "k");
    check_sizes("GaussianDihedralForceFieldGenerator", indices.size(), theta0s, "theta0");
    check_sizes("GaussianDihedralForceFieldGenerator", indices.size(), sigmas, "sigma");
}
ForceDescriptor GaussianDihedralForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_GAUSSIAN_DIHEDRAL;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        for (auto i : indices_[t]) d.indices.push_back(static_cast<int32_t>(i));
        d.params.insert(d.params.end(), {ks_[t], theta0s_[t], sigmas_[t]});
    }
    return d;
}

FlexibleLocalDihedralForceFieldGenerator::FlexibleLocalDihedralForceFieldGenerator(
    const std::vector<indices_type>& indices, const std::vector<double>& ks, const std::array<double, 7>& table,
    const std::string& aa_pair_name, bool use_periodic)
    : indices_(indices), ks_(ks), table_(table), aa_pair_name_(aa_pair_name), use_periodic_(use_periodic) {
    check_sizes("FlexibleLocalDihedralForceFieldGenerator", indices.size(), ks, "k");
}

// FlexibleLocalDihedralForceFieldGenerator.cpp:7-26: minimum of the Fourier
// sum over phi in [-pi, pi) in 1e-4 steps.
double FlexibleLocalDihedralForceFieldGenerator::min_energy() const {
    double emin = std::numeric_limits<double>::max();
    for (double phi = -constant::pi; phi < constant::pi; phi += 1e-4) {
        const double s1 = std::sin(phi), c1 = std::cos(phi);
        const double ss = s1 * s1, cc = c1 * c1;
        const double s2 = 2 * s1 * c1, c2 = cc - ss;
        const double s3 = s1 * (3 - 4 * ss), c3 = c1 * (4 * cc - 3);
        emin = std::min(emin, table_[1] * c1 + table_[2] * s1 + table_[3] * c2 + table_[4] * s2 + table_[5] * c3 +
                                  table_[6] * s3 + table_[0]);
    }
    return emin;
}

ForceDescriptor FlexibleLocalDihedralForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_FLEXIBLE_LOCAL_DIHEDRAL;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        for (auto i : indices_[t]) d.indices.push_back(static_cast<int32_t>(i));
        d.params.push_back(ks_[t]);
    }
    d.globals.assign(table_.begin(), table_.end());
    d.globals.push_back(min_energy());
    return d;
}

// ----------------------------------------------------- global generators
ExcludedVolumeForceFieldGenerator::ExcludedVolumeForceFieldGenerator(
    double eps, double cutoff_ratio, const std::vector<std::optional<double>>& radii, const index_pairs_type& ignore,
    bool use_periodic, const group_pairs_type& ignore_group_pairs, const std::vector<std::optional<std::string>>& group_vec)
    : eps_(eps), cutoff_(cutoff_ratio), radii_(radii), ignore_list_(ignore), use_periodic_(use_periodic),
      ignore_group_pairs_(ignore_group_pairs), group_vec_(group_vec) {}

// ExcludedVolumeForceFieldGenerator.cpp:31-83: cutoff = ratio x (largest +
// second largest radius), the second tracked with a strict comparison.
double ExcludedVolumeForceFieldGenerator::cutoff_distance() const {
    double mx = std::numeric_limits<double>::min(), second = std::numeric_limits<double>::min();
    for (const auto& r : radii_) {
        if (!r) continue;
        if (mx < *r) { second = mx; mx = *r; }
        else if (second < *r) second = *r;
    }
    return (mx + second) * cutoff_;
}

ForceDescriptor ExcludedVolumeForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_EXCLUDED_VOLUME;
    d.use_periodic = use_periodic_;
    std::vector<bool> has(radii_.size());
    for (std::size_t i = 0; i < radii_.size(); ++i) {
        has[i] = radii_[i].has_value();
        d.particle_params.push_back(has[i] ? *radii_[i] : 0.0);
    }
    interaction_categories(has, ignore_group_pairs_, group_vec_, d);
    d.cutoff = cutoff_distance();
    d.globals = {eps_};
    std::cerr << "    ExcludedVolume                : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

double DebyeHuckelForceFieldGenerator::dielectric_water(double T, double C) {
    return (249.4 - 0.788 * T + 7.2e-4 * T * T) * (1. - 2.551e-1 * C + 5.151e-2 * C * C - 6.889e-3 * C * C * C);
}

// DebyeHuckelForceFieldGenerator.cpp:10-51
DebyeHuckelForceFieldGenerator::DebyeHuckelForceFieldGenerator(
    double ionic_strength, double temperature, double cutoff_ratio, const std::vector<std::optional<double>>& charges,
    const index_pairs_type& ignore, bool use_periodic, const group_pairs_type& ignore_group_pairs,
    const std::vector<std::optional<std::string>>& group_vec)
    : ionic_strength_(ionic_strength), temperature_(temperature), cutoff_ratio_(cutoff_ratio), charges_(charges),
      ignore_list_(ignore), use_periodic_(use_periodic), ignore_group_pairs_(ignore_group_pairs), group_vec_(group_vec) {
    const double epsk = dielectric_water(temperature_, ionic_strength_);
    const double eps0 = constant::eps0() / constant::elementary_charge / constant::elementary_charge;  // mol/kJ/nm
    inv_4_pi_eps0_epsk_ = 1.0 / (4.0 * constant::pi * eps0 * epsk);
    if (!(ionic_strength_ > 0)) throw std::runtime_error("[error] ionic strength must be larger than zero");
    const double I = ionic_strength_ * 1.0e-24;  // mol/nm^3
    debye_length_ = std::sqrt((eps0 * epsk * constant::kB * 1.0e-3 * temperature_) / (2. * I));
    abs_cutoff_ = debye_length_ * cutoff_ratio_;
    cutoff_correction_ = std::exp(-cutoff_ratio_) / abs_cutoff_;
    std::cerr << std::scientific << std::setprecision(2) << "        eps0 is " << eps0 << " mol/KJ/nm" << std::endl
              << std::fixed << "        epsk is " << epsk << " (dimensionless)" << std::endl
              << "        debye length is " << debye_length_ << " nm" << std::endl;
}

ForceDescriptor DebyeHuckelForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_DEBYE_HUCKEL;
    d.use_periodic = use_periodic_;
    std::vector<bool> has(charges_.size());
    for (std::size_t i = 0; i < charges_.size(); ++i) {
        has[i] = charges_[i].has_value();
        d.particle_params.push_back(has[i] ? *charges_[i] : 0.0);
    }
    interaction_categories(has, ignore_group_pairs_, group_vec_, d);
    d.cutoff = abs_cutoff_;
    d.globals = {inv_4_pi_eps0_epsk_, debye_length_, cutoff_correction_};
    std::cerr << "    DebyeHuckel                   : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

static void check_lj(const char* who, const std::vector<std::optional<double>>& e, const std::vector<std::optional<double>>& s) {
    if (e.size() != s.size()) throw std::runtime_error(std::string("[error] ") + who + ": sigma and epsilon sizes differ");
    for (std::size_t i = 0; i < s.size(); ++i)
        if (s[i].has_value() != e[i].has_value())
            throw std::runtime_error(std::string("[error] ") + who +
                                     " : The parameter set is epsilon and sigma. incomplete parameter set was given for "
                                     "particle idx " + std::to_string(i) + ".");
}

// LennardJones*ForceFieldGenerator.cpp: cutoff = ratio x (largest + second
// largest sigma) / 2, both tracked with <=.
static double lj_cutoff(const std::vector<std::optional<double>>& sig, double ratio) {
    double mx = std::numeric_limits<double>::min(), second = std::numeric_limits<double>::min();
    for (const auto& s : sig) {
        if (!s) continue;
        if (mx <= *s) { second = mx; mx = *s; }
        else if (second <= *s) second = *s;
    }
    return (mx + second) * 0.5 * ratio;
}

static void lj_params(const std::vector<std::optional<double>>& e, const std::vector<std::optional<double>>& s,
                      ForceDescriptor& d) {
    for (std::size_t i = 0; i < s.size(); ++i) {
        d.particle_params.push_back(e[i] ? *e[i] : 0.0);
        d.particle_params.push_back(s[i] ? *s[i] : 0.0);
    }
}

LennardJonesRepulsiveForceFieldGenerator::LennardJonesRepulsiveForceFieldGenerator(
    double cutoff_ratio, const std::vector<std::optional<double>>& epsilons,
    const std::vector<std::optional<double>>& sigmas, const index_pairs_type& ignore, bool use_periodic,
    const group_pairs_type& ignore_group_pairs, const std::vector<std::optional<std::string>>& group_vec)
    : cutoff_ratio_(cutoff_ratio), epsilons_(epsilons), sigmas_(sigmas), ignore_list_(ignore),
      use_periodic_(use_periodic), ignore_group_pairs_(ignore_group_pairs), group_vec_(group_vec) {
    check_lj("LennardJonesRepulsiveForceFieldGenerator", epsilons_, sigmas_);
}
ForceDescriptor LennardJonesRepulsiveForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_LJ_REPULSIVE;
    d.use_periodic = use_periodic_;
    lj_params(epsilons_, sigmas_, d);
    std::vector<bool> has(sigmas_.size());
    for (std::size_t i = 0; i < sigmas_.size(); ++i) has[i] = sigmas_[i].has_value();
    interaction_categories(has, ignore_group_pairs_, group_vec_, d);
    d.cutoff = lj_cutoff(sigmas_, cutoff_ratio_);
    std::cerr << "    LennardJonesRepulsive         : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

LennardJonesAttractiveForceFieldGenerator::LennardJonesAttractiveForceFieldGenerator(
    double cutoff_ratio, const std::vector<std::optional<double>>& epsilons,
    const std::vector<std::optional<double>>& sigmas, const index_pairs_type& ignore, bool use_periodic,
    const group_pairs_type& ignore_group_pairs, const std::vector<std::optional<std::string>>& group_vec)
    : cutoff_ratio_(cutoff_ratio), epsilons_(epsilons), sigmas_(sigmas), ignore_list_(ignore),
      use_periodic_(use_periodic), ignore_group_pairs_(ignore_group_pairs), group_vec_(group_vec) {
    check_lj("LennardJonesAttractiveForceFieldGenerator", epsilons_, sigmas_);
}
ForceDescriptor LennardJonesAttractiveForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_LJ_ATTRACTIVE;
    d.use_periodic = use_periodic_;
    lj_params(epsilons_, sigmas_, d);
    std::vector<bool> has(epsilons_.size());
    for (std::size_t i = 0; i < epsilons_.size(); ++i) has[i] = epsilons_[i].has_value();
    interaction_categories(has, ignore_group_pairs_, group_vec_, d);
    d.cutoff = lj_cutoff(sigmas_, cutoff_ratio_);
    d.switch_distance = 0.9 * d.cutoff;
    std::cerr << "    LennardJonesAttractive        : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

// ------------------------------------------- membrane and table potentials
CappedGoContactForceFieldGenerator::CappedGoContactForceFieldGenerator(const index_pairs_type& indices,
                                                                       const std::vector<double>& ks,
                                                                       const std::vector<double>& r0s,
                                                                       double capping_ratio, bool use_periodic)
    : indices_(indices), ks_(ks), r0s_(r0s), capping_ratio_(capping_ratio), use_periodic_(use_periodic) {
    check_sizes("CappedGoContactForceFieldGenerator", indices.size(), ks, "k");
    check_sizes("CappedGoContactForceFieldGenerator", indices.size(), r0s, "r0");
    if (!(capping_ratio > 0.0)) throw std::runtime_error("[error] CappedGoContactForceFieldGenerator: capping_ratio must be positive");
}
ForceDescriptor CappedGoContactForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_CAPPED_GO_CONTACT;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        d.indices.push_back(static_cast<int32_t>(indices_[t].first));
        d.indices.push_back(static_cast<int32_t>(indices_[t].second));
        d.params.insert(d.params.end(), {ks_[t], r0s_[t], capping_ratio_});
    }
    return d;
}

CosineDihedralForceFieldGenerator::CosineDihedralForceFieldGenerator(const std::vector<indices_type>& indices,
                                                                     const std::vector<double>& ks,
                                                                     const std::vector<double>& theta0s,
                                                                     const std::vector<double>& ns, bool use_periodic)
    : indices_(indices), ks_(ks), theta0s_(theta0s), ns_(ns), use_periodic_(use_periodic) {
    check_sizes("CosineDihedralForceFieldGenerator", indices.size(), ks, "k");
    check_sizes("CosineDihedralForceFieldGenerator", indices.size(), theta0s, "theta0");
    check_sizes("CosineDihedralForceFieldGenerator", indices.size(), ns, "n");
}
ForceDescriptor CosineDihedralForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_COSINE_DIHEDRAL;
    d.use_periodic = use_periodic_;
    d.n_terms = static_cast<int>(indices_.size());
    for (std::size_t t = 0; t < indices_.size(); ++t) {
        for (auto i : indices_[t]) d.indices.push_back(static_cast<int32_t>(i));
        d.params.insert(d.params.end(), {ks_[t], theta0s_[t], ns_[t]});
    }
    return d;
}

// (largest + second largest) of the values that are present, both tracked
// with <= from the smallest positive double, as the WCA and iSoLF generators
// do (WeeksChandlerAndersenForceFieldGenerator.cpp:46-66).
static double two_largest_sum(const std::vector<std::optional<double>>& v) {
    double mx = std::numeric_limits<double>::min(), second = std::numeric_limits<double>::min();
    for (const auto& x : v) {
        if (!x) continue;
        if (mx <= *x) { second = mx; mx = *x; }
        else if (second <= *x) second = *x;
    }
    return mx + second;
}

static const double kTwoSixth = std::pow(2.0, 1.0 / 6.0);

WeeksChandlerAndersenForceFieldGenerator::WeeksChandlerAndersenForceFieldGenerator(
    const std::vector<std::optional<double>>& sigmas, const std::vector<std::optional<double>>& epsilons,
    const index_pairs_type& ignore, bool use_periodic, const group_pairs_type& ignore_group_pairs,
    const std::vector<std::optional<std::string>>& group_vec)
    : sigmas_(sigmas), epsilons_(epsilons), ignore_list_(ignore), use_periodic_(use_periodic),
      ignore_group_pairs_(ignore_group_pairs), group_vec_(group_vec) {
    if (sigmas_.size() != epsilons_.size())
        throw std::runtime_error("[error] WCA: sigma and epsilon sizes differ");
    for (std::size_t i = 0; i < sigmas_.size(); ++i)
        if (sigmas_[i].has_value() != epsilons_[i].has_value())
            throw std::runtime_error("[error] WeeksChandlerAndersenForceFieldGenerator : The parameter set is sigma and "
                                     "epsilon. incomplete parameter set was given for particle idx " +
                                     std::to_string(i) + ".");
}
ForceDescriptor WeeksChandlerAndersenForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_WCA;
    d.use_periodic = use_periodic_;
    std::vector<bool> has(sigmas_.size());
    for (std::size_t i = 0; i < sigmas_.size(); ++i) {
        d.particle_params.push_back(sigmas_[i] ? *sigmas_[i] : 0.0);
        d.particle_params.push_back(epsilons_[i] ? *epsilons_[i] : 0.0);
        has[i] = sigmas_[i].has_value();
    }
    interaction_categories(has, ignore_group_pairs_, group_vec_, d);
    d.cutoff = two_largest_sum(sigmas_) * 0.5 * kTwoSixth;
    std::cerr << "    WeeksChandlerAndersen         : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

set_pairs_type uniform_interaction_groups(const std::vector<std::size_t>& former, const std::vector<std::size_t>& latter,
                                          const group_pairs_type& ignore_group_pairs,
                                          const std::vector<std::optional<std::string>>& group_vec) {
    set_pairs_type out;
    if (ignore_group_pairs.empty()) {
        out.push_back({std::set<int>(former.begin(), former.end()), std::set<int>(latter.begin(), latter.end())});
        return out;
    }
    // beads whose group is named in an ignore pair go to that name's set,
    // the rest to "others"; names are visited in sorted order
    std::set<std::string> names;
    for (const auto& p : ignore_group_pairs) { names.insert(p.first); names.insert(p.second); }
    auto split = [&](const std::vector<std::size_t>& members, std::map<std::string, std::set<int>>& named,
                     std::set<int>& others) {
        for (const auto& nm : names) named[nm];
        for (std::size_t i : members) {
            const bool grouped = i < group_vec.size() && group_vec[i] && names.count(*group_vec[i]);
            if (grouped) named[*group_vec[i]].insert(static_cast<int>(i));
            else others.insert(static_cast<int>(i));
        }
    };
    std::map<std::string, std::set<int>> f_named, l_named;
    std::set<int> f_others, l_others;
    split(former, f_named, f_others);
    split(latter, l_named, l_others);
    auto ignored = [&](const std::string& a, const std::string& b) {
        return std::find(ignore_group_pairs.begin(), ignore_group_pairs.end(), std::make_pair(a, b)) !=
                   ignore_group_pairs.end() ||
               std::find(ignore_group_pairs.begin(), ignore_group_pairs.end(), std::make_pair(b, a)) !=
                   ignore_group_pairs.end();
    };
    if (!f_others.empty()) {
        for (const auto& kv : l_named)
            if (!kv.second.empty()) out.push_back({f_others, kv.second});
        if (!l_others.empty()) out.push_back({f_others, l_others});
    }
    for (const auto& fk : f_named) {
        if (fk.second.empty()) continue;
        if (!l_others.empty()) out.push_back({fk.second, l_others});
        for (const auto& lk : l_named)
            if (!ignored(fk.first, lk.first) && !lk.second.empty()) out.push_back({fk.second, lk.second});
    }
    return out;
}

// Set pairs as categories: beads in the same sets share a category (-1: in
// no set), and categories a, b interact when some pair holds a in its first
// set and b in its second, or the reverse.  A pair of beads interacts once
// however many set pairs name it.
static void categories_from_set_pairs(std::size_t n, const set_pairs_type& groups, ForceDescriptor& d) {
    const std::size_t ns = 2 * groups.size();
    std::vector<std::vector<bool>> sig(n, std::vector<bool>(ns, false));
    for (std::size_t k = 0; k < groups.size(); ++k) {
        for (int i : groups[k].first) sig[static_cast<std::size_t>(i)][2 * k] = true;
        for (int i : groups[k].second) sig[static_cast<std::size_t>(i)][2 * k + 1] = true;
    }
    std::vector<std::vector<bool>> cats;
    d.category.assign(n, -1);
    for (std::size_t i = 0; i < n; ++i) {
        if (std::none_of(sig[i].begin(), sig[i].end(), [](bool b) { return b; })) continue;
        auto it = std::find(cats.begin(), cats.end(), sig[i]);
        if (it == cats.end()) {
            if (cats.size() == 32) throw std::runtime_error("[error] interaction groups need more than 32 categories");
            cats.push_back(sig[i]);
            it = cats.end() - 1;
        }
        d.category[i] = static_cast<int32_t>(it - cats.begin());
    }
    const std::size_t nc = std::max<std::size_t>(cats.size(), 1);
    d.n_categories = static_cast<int>(nc);
    d.allow.assign(nc * nc, 0);
    for (std::size_t a = 0; a < cats.size(); ++a)
        for (std::size_t b = 0; b < cats.size(); ++b)
            for (std::size_t k = 0; k < groups.size(); ++k)
                if ((cats[a][2 * k] && cats[b][2 * k + 1]) || (cats[a][2 * k + 1] && cats[b][2 * k])) {
                    d.allow[a * nc + b] = 1;
                    break;
                }
}

UniformWeeksChandlerAndersenForceFieldGenerator::UniformWeeksChandlerAndersenForceFieldGenerator(
    std::size_t system_size, double eps, double sigma, const std::vector<std::size_t>& former_group,
    const std::vector<std::size_t>& latter_group, const index_pairs_type& ignore, bool use_periodic,
    const group_pairs_type& ignore_group_pairs, const std::vector<std::optional<std::string>>& group_vec)
    : system_size_(system_size), eps_(eps), sigma_(sigma), ignore_list_(ignore), use_periodic_(use_periodic),
      former_size_(former_group.size()), latter_size_(latter_group.size()),
      groups_(uniform_interaction_groups(former_group, latter_group, ignore_group_pairs, group_vec)) {}
ForceDescriptor UniformWeeksChandlerAndersenForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_UNIFORM_WCA;
    d.use_periodic = use_periodic_;
    d.globals = {sigma_, eps_};
    categories_from_set_pairs(system_size_, groups_, d);
    d.cutoff = sigma_ * kTwoSixth;
    std::cerr << "    UniformWeeksChandlerAndersen  : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

UniformLennardJonesAttractiveForceFieldGenerator::UniformLennardJonesAttractiveForceFieldGenerator(
    std::size_t system_size, double eps, double sigma, double cutoff_ratio, const std::vector<std::size_t>& former_group,
    const std::vector<std::size_t>& latter_group, const index_pairs_type& ignore, bool use_periodic,
    const group_pairs_type& ignore_group_pairs, const std::vector<std::optional<std::string>>& group_vec)
    : system_size_(system_size), eps_(eps), sigma_(sigma), abs_cutoff_(sigma * cutoff_ratio), ignore_list_(ignore),
      use_periodic_(use_periodic), former_size_(former_group.size()), latter_size_(latter_group.size()),
      groups_(uniform_interaction_groups(former_group, latter_group, ignore_group_pairs, group_vec)) {}
ForceDescriptor UniformLennardJonesAttractiveForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_UNIFORM_LJ_ATTRACTIVE;
    d.use_periodic = use_periodic_;
    d.globals = {sigma_, eps_};
    categories_from_set_pairs(system_size_, groups_, d);
    d.cutoff = abs_cutoff_;
    d.switch_distance = 0.9 * abs_cutoff_;
    std::cerr << "    UniformLennardJonesAttractive : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

iSoLFAttractiveForceFieldGenerator::iSoLFAttractiveForceFieldGenerator(
    const std::vector<std::optional<double>>& sigmas, const std::vector<std::optional<double>>& epsilons,
    const std::vector<std::optional<double>>& omegas, const index_pairs_type& ignore, bool use_periodic,
    const group_pairs_type& ignore_group_pairs, const std::vector<std::optional<std::string>>& group_vec)
    : sigmas_(sigmas), epsilons_(epsilons), omegas_(omegas), ignore_list_(ignore), use_periodic_(use_periodic),
      ignore_group_pairs_(ignore_group_pairs), group_vec_(group_vec) {
    if (sigmas_.size() != epsilons_.size() || sigmas_.size() != omegas_.size())
        throw std::runtime_error("[error] iSoLFAttractive: sigma, epsilon and omega sizes differ");
    for (std::size_t i = 0; i < sigmas_.size(); ++i)
        if (sigmas_[i].has_value() != epsilons_[i].has_value() || sigmas_[i].has_value() != omegas_[i].has_value())
            throw std::runtime_error("[error] iSoLFAttractiveForceFieldGenerator : The parameter set is sigma, epsilon "
                                     "and omega. incomplete parameter set was given for particle index " +
                                     std::to_string(i) + ".");
}
ForceDescriptor iSoLFAttractiveForceFieldGenerator::generate() const {
    ForceDescriptor d;
    d.kind = OCM_ISOLF_ATTRACTIVE;
    d.use_periodic = use_periodic_;
    std::vector<bool> has(sigmas_.size());
    for (std::size_t i = 0; i < sigmas_.size(); ++i) {
        d.particle_params.push_back(sigmas_[i] ? *sigmas_[i] : 0.0);
        d.particle_params.push_back(epsilons_[i] ? *epsilons_[i] : 0.0);
        d.particle_params.push_back(omegas_[i] ? *omegas_[i] : 0.0);
        has[i] = sigmas_[i].has_value();
    }
    interaction_categories(has, ignore_group_pairs_, group_vec_, d);
    d.cutoff = two_largest_sum(sigmas_) * 0.5 * kTwoSixth + two_largest_sum(omegas_) * 0.5;
    std::cerr << "    iSoLFAttractive               : cutoff distance is " << d.cutoff << " nm" << std::endl;
    put_exclusions(ignore_list_, d);
    return d;
}

// ------------------------------------------------------ system, integrator
ocm_system_desc SystemDescriptor::view() {
    views.clear();
    for (const auto& f : forces) views.push_back(f.view());
    ocm_system_desc d{};
    d.n_particles = static_cast<int32_t>(masses.size());
    d.n_forces = static_cast<int32_t>(views.size());
    d.masses = masses.data();
    d.use_periodic_box = periodic ? 1 : 0;
    for (int c = 0; c < 3; ++c) d.box[c] = box[c];
    d.forces = views.empty() ? nullptr : views.data();
    return d;
}

std::string SystemGenerator::group_key(const std::string& name) {
    for (const char* k : {"FlexibleLocalAngle", "FlexibleLocalDihedral", "CombinatorialGoContact"})
        if (name.find(k) != std::string::npos) return k;
    return name;
}

void SystemGenerator::add_ff_generator(std::unique_ptr<ForceFieldGeneratorBase>&& gen) {
    const std::size_t id = groups_.size();
    groups_.insert(std::make_pair(group_key(gen->name()), id));  // an existing name keeps its id
    gens_.push_back(std::move(gen));
}

SystemDescriptor SystemGenerator::generate() const {
    SystemDescriptor s;
    s.masses = mass_vec_;
    if (edge_lengths_) {
        s.periodic = true;
        s.box = *edge_lengths_;
    }
    for (const auto& g : gens_) {
        ForceDescriptor f = g->generate();
        f.group = static_cast<int>(groups_.at(group_key(g->name())));
        if (f.group > 31) throw std::runtime_error("[error] more than 32 force groups");
        s.forces.push_back(std::move(f));
    }
    return s;
}

ocm_integrator_desc LangevinIntegratorGenerator::generate() const {
    ocm_integrator_desc d{};
    d.temperature = temperature_;
    d.friction = friction_;
    d.dt = delta_t_;
    if (seed_ != 0) {
        d.seed = static_cast<uint64_t>(static_cast<uint32_t>(seed_));
    } else {
        std::random_device rd;
        d.seed = (static_cast<uint64_t>(rd()) << 32) | rd();
    }
    return d;
}

std::string LangevinIntegratorGenerator::dump_info() const {
    std::ostringstream o;
    o << std::fixed << std::setprecision(2) << "    temperature : " << std::setw(7) << temperature_ << " K\n"
      << std::setprecision(3) << "    delta_t     : " << std::setw(7) << delta_t_ << " ps\n"
      << "    gamma       : " << std::setw(7) << friction_ << " ps^-1\n    seed        : ";
    if (seed_ == 0) o << "not specified or 0. random seed will be chosen.";
    else o << std::setw(7) << seed_;
    return o.str();
}

}  // namespace ocmhost
