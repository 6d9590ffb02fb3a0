// forcefield.hpp -- force-field generators of the MD front end.
//
// The class shapes follow the reference's generator interface
// (src/forcefield/ForceFieldGeneratorBase.hpp:9-16): each generator is built
// from OpenMM-unit parameters by a reader and its generate() returns the
// force.  Here generate() returns a ForceDescriptor -- owned arrays viewed as
// an ocm_force_desc (include/ocm.h) -- instead of an OpenMM::Force.
#pragma once

#include "../../include/ocm.h"

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ocmhost {

using index_pair = std::pair<std::size_t, std::size_t>;
using index_pairs_type = std::vector<index_pair>;
using group_pairs_type = std::vector<std::pair<std::string, std::string>>;

// ---------------------------------------------------------------- Topology
// Typed-edge graph of the beads (src/Topology.hpp:10-82, Topology.cpp:7-135).
class Topology {
public:
    explicit Topology(std::size_t n) : nodes_(n) {}
    void add_edges(const std::vector<index_pair>& edges, const std::string& type);
    template <std::size_t N> void add_edges(const std::vector<std::array<std::size_t, N>>& sets, const std::string& type) {
        for (const auto& s : sets)
            for (std::size_t a = 0; a + 1 < N; ++a)
                for (std::size_t b = a + 1; b < N; ++b) link(s[a], s[b], type);
    }
    void make_molecule(const std::string& type);
    index_pairs_type ignore_list_within_molecule() const;
    // sorted unique pairs (i<j) of beads within `dist` edges of type `type`
    index_pairs_type ignore_list_within_edge(std::size_t dist, const std::string& type) const;
    std::size_t size() const { return nodes_.size(); }

private:
    void link(std::size_t a, std::size_t b, const std::string& type);
    void adjacent_within(std::size_t node, std::size_t dist, const std::string& type, std::vector<std::size_t>& out) const;
    std::vector<std::vector<std::pair<std::size_t, std::string>>> nodes_;
    std::vector<std::vector<std::size_t>> molecules_;
    bool molecule_updated_ = false;
};

// --------------------------------------------------------- descriptors
struct ForceDescriptor {
    int kind = 0;
    int group = 0;
    bool use_periodic = false;
    int n_terms = 0;
    std::vector<int32_t> indices;
    std::vector<double> params;
    std::vector<double> globals;
    std::vector<double> particle_params;
    std::vector<int32_t> category;   // empty: every bead is category 0
    int n_categories = 0;
    std::vector<uint8_t> allow;      // empty: all category pairs interact
    double cutoff = 0.0;
    double switch_distance = 0.0;
    std::vector<int32_t> exclusions;  // flattened pairs
    ocm_force_desc view() const;
};

// Interaction groups of InteractionGroup.hpp:17-101 expressed as categories:
// 0 = "others", 1.. = the named groups that appear in ignore pairs (sorted by
// name), -1 = non-participant; allow[a][b] follows the set pairs it builds.
void interaction_categories(const std::vector<bool>& has_param, const group_pairs_type& ignore_group_pairs,
                            const std::vector<std::optional<std::string>>& group_vec, ForceDescriptor& out);

class ForceFieldGeneratorBase {
public:
    virtual ~ForceFieldGeneratorBase() = default;
    virtual ForceDescriptor generate() const = 0;
    virtual std::string name() const = 0;
};

// ------------------------------------------------------ local generators
class HarmonicBondForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    HarmonicBondForceFieldGenerator(const index_pairs_type& indices, const std::vector<double>& v0s,
                                    const std::vector<double>& ks, bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "HarmonicBond"; }
    std::size_t size() const { return indices_.size(); }
    const index_pairs_type& indices() const { return indices_; }
private:
    index_pairs_type indices_;
    std::vector<double> v0s_, ks_;
    bool use_periodic_;
};

class GaussianBondForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    GaussianBondForceFieldGenerator(const index_pairs_type& indices, const std::vector<double>& ks,
                                    const std::vector<double>& v0s, const std::vector<double>& sigmas,
                                    bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "GaussianBond"; }
    std::size_t size() const { return indices_.size(); }
private:
    index_pairs_type indices_;
    std::vector<double> ks_, v0s_, sigmas_;
    bool use_periodic_;
};

class GoContactForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    GoContactForceFieldGenerator(const index_pairs_type& indices, const std::vector<double>& ks,
                                 const std::vector<double>& r0s, bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "GoContact"; }
    std::size_t size() const { return indices_.size(); }
private:
    index_pairs_type indices_;
    std::vector<double> ks_, r0s_;
    bool use_periodic_;
};

// CappedGoContactForceFieldGenerator.hpp:20-52: Go 12-10 contacts with a
// linear cap below capping_ratio * r0.
class CappedGoContactForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    CappedGoContactForceFieldGenerator(const index_pairs_type& indices, const std::vector<double>& ks,
                                       const std::vector<double>& r0s, double capping_ratio, bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "CappedGoContact"; }
    std::size_t size() const { return indices_.size(); }
private:
    index_pairs_type indices_;
    std::vector<double> ks_, r0s_;
    double capping_ratio_;
    bool use_periodic_;
};

class HarmonicAngleForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    using indices_type = std::array<std::size_t, 3>;
    HarmonicAngleForceFieldGenerator(const std::vector<indices_type>& indices, const std::vector<double>& v0s,
                                     const std::vector<double>& ks, bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "HarmonicAngle"; }
    std::size_t size() const { return indices_.size(); }
private:
    std::vector<indices_type> indices_;
    std::vector<double> v0s_, ks_;
    bool use_periodic_;
};

class FlexibleLocalAngleForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    using indices_type = std::array<std::size_t, 3>;
    FlexibleLocalAngleForceFieldGenerator(const std::vector<indices_type>& indices, const std::vector<double>& ks,
                                          const std::array<double, 10>& spline_table, const std::string& aa_name,
                                          bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "FlexibleLocalAngle (" + aa_name_ + ")"; }
    std::size_t size() const { return indices_.size(); }
    double min_energy() const;  // the reference's 1e-4 sweep of its own spline evaluator
private:
    std::vector<indices_type> indices_;
    std::vector<double> ks_;
    std::array<double, 10> table_;
    std::string aa_name_;
    bool use_periodic_;
};

class GaussianDihedralForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    using indices_type = std::array<std::size_t, 4>;
    GaussianDihedralForceFieldGenerator(const std::vector<indices_type>& indices, const std::vector<double>& ks,
                                        const std::vector<double>& theta0s, const std::vector<double>& sigmas,
                                        bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "GaussianDihedral"; }
    std::size_t size() const { return indices_.size(); }
private:
    std::vector<indices_type> indices_;
    std::vector<double> ks_, theta0s_, sigmas_;
    bool use_periodic_;
};

// CosineDihedralForceFieldGenerator.hpp:22-60: k (1 - cos(n (phi - t0))).
class CosineDihedralForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    using indices_type = std::array<std::size_t, 4>;
    CosineDihedralForceFieldGenerator(const std::vector<indices_type>& indices, const std::vector<double>& ks,
                                      const std::vector<double>& theta0s, const std::vector<double>& ns,
                                      bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "CosineDihedral"; }
    std::size_t size() const { return indices_.size(); }
private:
    std::vector<indices_type> indices_;
    std::vector<double> ks_, theta0s_, ns_;
    bool use_periodic_;
};

class FlexibleLocalDihedralForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    using indices_type = std::array<std::size_t, 4>;
    FlexibleLocalDihedralForceFieldGenerator(const std::vector<indices_type>& indices, const std::vector<double>& ks,
                                             const std::array<double, 7>& fourier_table, const std::string& aa_pair_name,
                                             bool use_periodic);
    ForceDescriptor generate() const override;
    std::string name() const override { return "FlexibleLocalDihedral (" + aa_pair_name_ + ")"; }
    std::size_t size() const { return indices_.size(); }
    double min_energy() const;
private:
    std::vector<indices_type> indices_;
    std::vector<double> ks_;
    std::array<double, 7> table_;
    std::string aa_pair_name_;
    bool use_periodic_;
};

// ----------------------------------------------------- global generators
class ExcludedVolumeForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    ExcludedVolumeForceFieldGenerator(double eps, double cutoff_ratio, const std::vector<std::optional<double>>& radii,
                                      const index_pairs_type& ignore_list, bool use_periodic,
                                      const group_pairs_type& ignore_group_pairs = {},
                                      const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "ExcludedVolume"; }
    double cutoff_distance() const;
private:
    double eps_, cutoff_;
    std::vector<std::optional<double>> radii_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    group_pairs_type ignore_group_pairs_;
    std::vector<std::optional<std::string>> group_vec_;
};

class DebyeHuckelForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    DebyeHuckelForceFieldGenerator(double ionic_strength, double temperature, double cutoff_ratio,
                                   const std::vector<std::optional<double>>& charges,
                                   const index_pairs_type& ignore_list, bool use_periodic,
                                   const group_pairs_type& ignore_group_pairs = {},
                                   const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "DebyeHuckel"; }
    double debye_length() const { return debye_length_; }
    double inv_4_pi_eps0_epsk() const { return inv_4_pi_eps0_epsk_; }
    double abs_cutoff() const { return abs_cutoff_; }
    static double dielectric_water(double T, double C);
private:
    double ionic_strength_, temperature_, cutoff_ratio_;
    std::vector<std::optional<double>> charges_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    group_pairs_type ignore_group_pairs_;
    std::vector<std::optional<std::string>> group_vec_;
    double debye_length_, inv_4_pi_eps0_epsk_, cutoff_correction_, abs_cutoff_;
};

class LennardJonesRepulsiveForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    LennardJonesRepulsiveForceFieldGenerator(double cutoff_ratio, const std::vector<std::optional<double>>& epsilons,
                                             const std::vector<std::optional<double>>& sigmas,
                                             const index_pairs_type& ignore_list, bool use_periodic,
                                             const group_pairs_type& ignore_group_pairs = {},
                                             const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "LennardJonesRepulsive"; }
private:
    double cutoff_ratio_;
    std::vector<std::optional<double>> epsilons_, sigmas_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    group_pairs_type ignore_group_pairs_;
    std::vector<std::optional<std::string>> group_vec_;
};

class LennardJonesAttractiveForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    LennardJonesAttractiveForceFieldGenerator(double cutoff_ratio, const std::vector<std::optional<double>>& epsilons,
                                              const std::vector<std::optional<double>>& sigmas,
                                              const index_pairs_type& ignore_list, bool use_periodic,
                                              const group_pairs_type& ignore_group_pairs = {},
                                              const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "LennardJonesAttractive"; }
private:
    double cutoff_ratio_;
    std::vector<std::optional<double>> epsilons_, sigmas_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    group_pairs_type ignore_group_pairs_;
    std::vector<std::optional<std::string>> group_vec_;
};

// WeeksChandlerAndersenForceFieldGenerator.hpp:15-45: per-bead sigma and
// epsilon, cutoff 2^(1/6) (largest + second largest sigma) / 2.
class WeeksChandlerAndersenForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    WeeksChandlerAndersenForceFieldGenerator(const std::vector<std::optional<double>>& sigmas,
                                             const std::vector<std::optional<double>>& epsilons,
                                             const index_pairs_type& ignore_list, bool use_periodic,
                                             const group_pairs_type& ignore_group_pairs = {},
                                             const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "WeeksChandlerAndersen"; }
private:
    std::vector<std::optional<double>> sigmas_, epsilons_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    group_pairs_type ignore_group_pairs_;
    std::vector<std::optional<std::string>> group_vec_;
};

// Interaction groups as pairs of bead sets, each pair meaning "every bead of
// the first set with every bead of the second" (OpenMM addInteractionGroup).
using set_pairs_type = std::vector<std::pair<std::set<int>, std::set<int>>>;

// The set pairs of a table potential between two named bead kinds, split by
// the groups `ignore.group` names (UniformWeeksChandlerAndersen-
// ForceFieldGenerator.cpp:21-131; the same rule in UniformLennardJones-
// AttractiveForceFieldGenerator.cpp:36-139).
set_pairs_type uniform_interaction_groups(const std::vector<std::size_t>& former, const std::vector<std::size_t>& latter,
                                          const group_pairs_type& ignore_group_pairs,
                                          const std::vector<std::optional<std::string>>& group_vec);

// UniformWeeksChandlerAndersenForceFieldGenerator.hpp:17-52: one sigma and
// epsilon between the beads of two named kinds, cutoff 2^(1/6) sigma.
class UniformWeeksChandlerAndersenForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    UniformWeeksChandlerAndersenForceFieldGenerator(std::size_t system_size, double eps, double sigma,
                                                    const std::vector<std::size_t>& former_group,
                                                    const std::vector<std::size_t>& latter_group,
                                                    const index_pairs_type& ignore_list, bool use_periodic,
                                                    const group_pairs_type& ignore_group_pairs = {},
                                                    const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    // the reference spells this name so (.hpp:36); force groups follow it
    std::string name() const override { return "UniformWeeksChadlerAndersen"; }
    std::size_t former_group_size() const { return former_size_; }
    std::size_t latter_group_size() const { return latter_size_; }
private:
    std::size_t system_size_;
    double eps_, sigma_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    std::size_t former_size_, latter_size_;
    set_pairs_type groups_;
};

// UniformLennardJonesAttractiveForceFieldGenerator.hpp:17-55: switched LJ
// attraction with one sigma and epsilon between two named kinds; cutoff
// cutoff_ratio * sigma, switching from 0.9 of it.
class UniformLennardJonesAttractiveForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    UniformLennardJonesAttractiveForceFieldGenerator(std::size_t system_size, double eps, double sigma,
                                                     double cutoff_ratio, const std::vector<std::size_t>& former_group,
                                                     const std::vector<std::size_t>& latter_group,
                                                     const index_pairs_type& ignore_list, bool use_periodic,
                                                     const group_pairs_type& ignore_group_pairs = {},
                                                     const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "UniformLennardJonesAttractive"; }
    std::size_t former_group_size() const { return former_size_; }
    std::size_t latter_group_size() const { return latter_size_; }
private:
    std::size_t system_size_;
    double eps_, sigma_, abs_cutoff_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    std::size_t former_size_, latter_size_;
    set_pairs_type groups_;
};

// iSoLFAttractiveForceFieldGenerator.hpp:15-48: per-bead sigma, epsilon,
// omega; cutoff 2^(1/6) (two largest sigma)/2 + (two largest omega)/2.
class iSoLFAttractiveForceFieldGenerator final : public ForceFieldGeneratorBase {
public:
    iSoLFAttractiveForceFieldGenerator(const std::vector<std::optional<double>>& sigmas,
                                       const std::vector<std::optional<double>>& epsilons,
                                       const std::vector<std::optional<double>>& omegas,
                                       const index_pairs_type& ignore_list, bool use_periodic,
                                       const group_pairs_type& ignore_group_pairs = {},
                                       const std::vector<std::optional<std::string>>& group_vec = {});
    ForceDescriptor generate() const override;
    std::string name() const override { return "iSoLFAttractive"; }
private:
    std::vector<std::optional<double>> sigmas_, epsilons_, omegas_;
    index_pairs_type ignore_list_;
    bool use_periodic_;
    group_pairs_type ignore_group_pairs_;
    std::vector<std::optional<std::string>> group_vec_;
};

// ------------------------------------------------------ system, integrator
struct SystemDescriptor {
    std::vector<double> masses;
    bool periodic = false;
    std::array<double, 3> box{0, 0, 0};
    std::vector<ForceDescriptor> forces;
    std::vector<ocm_force_desc> views;     // filled by view()
    ocm_system_desc view();
};

// src/SystemGenerator.hpp:15-55: masses, optional cuboid box, one force per
// generator, one force group per generator name (all FlexibleLocalAngle /
// FlexibleLocalDihedral / CombinatorialGoContact generators share one each).
class SystemGenerator {
public:
    explicit SystemGenerator(std::vector<double> mass_vec) : mass_vec_(std::move(mass_vec)) {}
    void add_ff_generator(std::unique_ptr<ForceFieldGeneratorBase>&& gen);
    void set_pbc(double x, double y, double z) { edge_lengths_ = std::array<double, 3>{x, y, z}; }
    SystemDescriptor generate() const;
    std::map<std::string, std::size_t> ffname_groupid_map() const { return groups_; }
    std::size_t size() const { return mass_vec_.size(); }
    bool periodic() const { return edge_lengths_.has_value(); }
private:
    static std::string group_key(const std::string& name);
    std::vector<double> mass_vec_;
    std::vector<std::unique_ptr<ForceFieldGeneratorBase>> gens_;
    std::optional<std::array<double, 3>> edge_lengths_;
    std::map<std::string, std::size_t> groups_;
};

class IntegratorGeneratorBase {
public:
    virtual ~IntegratorGeneratorBase() = default;
    virtual ocm_integrator_desc generate() const = 0;
    virtual std::string dump_info() const = 0;
    virtual double temperature() const = 0;
};

// LangevinIntegratorGenerator (src/IntegratorGenerator.hpp:19-42); seed 0
// draws a random seed as OpenMM does.
class LangevinIntegratorGenerator final : public IntegratorGeneratorBase {
public:
    LangevinIntegratorGenerator(double temperature, double friction_coeff, double delta_t, int seed)
        : temperature_(temperature), friction_(friction_coeff), delta_t_(delta_t), seed_(seed) {}
    ocm_integrator_desc generate() const override;
    std::string dump_info() const override;
    double temperature() const override { return temperature_; }
private:
    double temperature_, friction_, delta_t_;
    int seed_;
};

// physical constants of src/util/Constants.hpp:16-23
namespace constant {
constexpr double Na = 6.02214076e23;
constexpr double kB = 1.380649e-23;
constexpr double pi = 3.1415926535897932385;
constexpr double elementary_charge = 1.6021766208e-19;
constexpr double kj_per_kcal = 4.184;
constexpr double nm_per_angstrom = 0.1;
constexpr double angstrom_per_nm = 10.0;
double eps0();      // C^2 mol / kJ / nm
double cafetime();  // ps
extern const double fla_min_theta, fla_max_theta;
const std::map<std::string, std::array<double, 10>>& fla_spline_table();
const std::map<std::string, std::array<double, 10>>& fla_spline_second_derivative_table();
const std::map<std::pair<std::string, std::string>, std::array<double, 7>>& fld_fourier_table();
}  // namespace constant

}  // namespace ocmhost
