// read_toml.hpp -- TOML input readers (src/input/ReadTOMLInput.cpp,
// src/input/ReadTOMLForceFieldGenerator.cpp) for the MD-step potentials.
#pragma once

#include "forcefield.hpp"
#include "toml.hpp"

#include <array>
#include <memory>
#include <optional>
#include <string>
#include <vector>

namespace ocmhost {

using ocmtoml::Value;

// `env` aliasing (src/util/Utility.hpp:47-158): a string parameter names an
// entry of the force field's env table.
const Value& find_parameter(const Value& params, const Value* env, const std::string& name);
const Value& find_parameter_either(const Value& params, const Value* env, const std::string& n1, const std::string& n2);
double find_double_or(const Value& params, const Value* env, const std::string& name, double fallback);
void expand_include(Value& v, const std::string& base_dir);
bool check_keys_available(const Value& table, std::initializer_list<std::string> keys);

index_pairs_type read_ignore_molecule_and_particles_within(const Value& ignore, const Topology& topology);
group_pairs_type read_ignore_group(const Value& ignore);

std::unique_ptr<HarmonicBondForceFieldGenerator> read_toml_harmonic_bond_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<GaussianBondForceFieldGenerator> read_toml_gaussian_bond_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<GoContactForceFieldGenerator> read_toml_go_contact_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<CappedGoContactForceFieldGenerator> read_toml_capped_go_contact_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<CosineDihedralForceFieldGenerator> read_toml_cosine_dihedral_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<WeeksChandlerAndersenForceFieldGenerator> read_toml_weeks_chandler_andersen_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic);
std::unique_ptr<iSoLFAttractiveForceFieldGenerator> read_toml_isolf_attractive_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic);
// One generator per (first, second) entry of the potential's `table`, in the
// table's order; pairs with an empty side are skipped with a warning
// (ReadTOMLInput.cpp:435-500 for WCA, :546-611 for LJ attraction).
std::vector<std::unique_ptr<UniformWeeksChandlerAndersenForceFieldGenerator>>
read_toml_uniform_weeks_chandler_andersen_ff_generators(const Value& ff, std::size_t n, const Topology& topology,
                                                        const std::vector<std::optional<std::string>>& groups,
                                                        bool use_periodic);
std::vector<std::unique_ptr<UniformLennardJonesAttractiveForceFieldGenerator>>
read_toml_uniform_lennard_jones_attractive_ff_generators(const Value& ff, std::size_t n, const Topology& topology,
                                                         const std::vector<std::optional<std::string>>& groups,
                                                         bool use_periodic);
std::unique_ptr<HarmonicAngleForceFieldGenerator> read_toml_harmonic_angle_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<FlexibleLocalAngleForceFieldGenerator> read_toml_flexible_local_angle_ff_generator(
    const Value& ff, const std::string& aa_type, const std::array<double, 10>& table, Topology& topology, bool use_periodic);
std::unique_ptr<GaussianDihedralForceFieldGenerator> read_toml_gaussian_dihedral_ff_generator(const Value& ff, Topology& topology, bool use_periodic);
std::unique_ptr<FlexibleLocalDihedralForceFieldGenerator> read_toml_flexible_local_dihedral_ff_generator(
    const Value& ff, const std::pair<std::string, std::string>& aa_pair, const std::array<double, 7>& table,
    Topology& topology, bool use_periodic);
std::unique_ptr<ExcludedVolumeForceFieldGenerator> read_toml_excluded_volume_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic);
std::unique_ptr<DebyeHuckelForceFieldGenerator> read_toml_debye_huckel_ff_generator(
    const Value& ff, std::size_t n, double ionic_strength, double temperature, const Topology& topology,
    const std::vector<std::optional<std::string>>& groups, bool use_periodic);
std::unique_ptr<LennardJonesAttractiveForceFieldGenerator> read_toml_lennard_jones_attractive_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic);
std::unique_ptr<LennardJonesRepulsiveForceFieldGenerator> read_toml_lennard_jones_repulsive_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic);

SystemGenerator read_toml_system(const Value& data);
std::unique_ptr<LangevinIntegratorGenerator> read_toml_integrator_gen(const Value& data);
std::vector<std::array<double, 3>> read_toml_initial_conf(const Value& data);   // nm
std::vector<std::array<double, 3>> read_toml_initial_vel(const Value& data);    // nm/ps; empty if absent

// The parsed [files]/[simulator] settings the run loop needs.
struct RunSettings {
    std::string output_prefix, output_format, output_path;
    bool progress_bar = true;
    std::size_t total_step = 0, save_step = 0;
    double delta_t_cafe = 0.0;
    bool energy_minimization = false;
    bool periodic = false;
    std::int64_t vel_seed = 0;
    std::vector<std::optional<std::string>> names;
};
RunSettings read_toml_run_settings(const Value& data);
Value load_toml_input(const std::string& path);   // parse + expand includes + suffix check

}  // namespace ocmhost
