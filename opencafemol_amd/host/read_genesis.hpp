// read_genesis.hpp -- GENESIS-format inputs: the `.inp` control file, the
// GROMACS-style `.top`/`.itp` topology and the `.gro` coordinates
// (src/input/ReadGenesisInput.cpp, src/input/ReadGenesisForceFieldGenerator.cpp).
#pragma once

#include "forcefield.hpp"

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace ocmhost {

// section name -> key -> value; the first value of a repeated key is kept
using InpData = std::map<std::string, std::map<std::string, std::string>>;
// section name -> content lines; repeated sections are concatenated
using TopData = std::map<std::string, std::vector<std::string>>;

InpData read_inp_file(const std::string& inp_file_name);
// blank and `;` lines dropped, trailing `;` comments cut, `#include "f"` expanded
// relative to `file_path` (the directory of the .inp file)
std::vector<std::string> preprocess_top_file(const std::string& top_file_name, const std::string& file_path);
TopData read_top_file(const std::string& top_file_name, const std::string& file_path);

// one reader per topology section; column layouts as in the reference
std::unique_ptr<HarmonicBondForceFieldGenerator> read_genesis_harmonic_bond_ff_generator(
    const std::vector<std::string>& bonds, Topology& topology, bool use_periodic);
std::unique_ptr<GaussianBondForceFieldGenerator> read_genesis_gaussian_bond_ff_generator(
    const std::vector<std::string>& angles, std::size_t n, bool use_periodic);
std::unique_ptr<GoContactForceFieldGenerator> read_genesis_go_contact_ff_generator(
    const std::vector<std::string>& pairs, Topology& topology, bool use_periodic);
std::unique_ptr<FlexibleLocalAngleForceFieldGenerator> read_genesis_flexible_local_angle_ff_generator(
    const std::vector<std::string>& angles, const std::vector<std::string>& atoms, const std::string& aa_type,
    bool use_periodic);
std::unique_ptr<GaussianDihedralForceFieldGenerator> read_genesis_gaussian_dihedral_ff_generator(
    const std::vector<std::string>& dihedrals, std::size_t n, bool use_periodic);
std::unique_ptr<FlexibleLocalDihedralForceFieldGenerator> read_genesis_flexible_local_dihedral_ff_generator(
    const std::vector<std::string>& dihedrals, const std::vector<std::string>& atoms,
    const std::pair<std::string, std::string>& aa_type_pair, bool use_periodic);
std::unique_ptr<ExcludedVolumeForceFieldGenerator> read_genesis_exv_ff_generator(
    const std::vector<std::string>& atomtypes, const std::vector<std::string>& atoms, const Topology& topology,
    bool use_periodic);

std::unique_ptr<LangevinIntegratorGenerator> read_genesis_integrator_gen(
    const std::map<std::string, std::string>& dynamics, const std::map<std::string, std::string>& ensemble);
std::vector<std::array<double, 3>> read_genesis_initial_conf(const std::string& gro_file_name,
                                                             const std::string& file_path);  // nm

// Everything make_simulator_from_genesis_inputs builds before the Simulator.
struct GenesisInput {
    SystemGenerator system{std::vector<double>{}};
    std::unique_ptr<LangevinIntegratorGenerator> integrator;
    std::vector<std::array<double, 3>> positions;       // nm
    std::vector<std::string> atom_names, res_names;     // 4- and 3-column fields of [ atoms ]
    std::string file_prefix;                            // dcdfile up to its last '.'
    std::string pdbfile;                                // empty if [OUTPUT] has none
    std::string platform;                               // empty if [PLATFORM] is absent
    std::size_t nsteps = 0, crdout_period = 0;
    double timestep = 0.0;                              // ps
    bool periodic = false;
};
GenesisInput read_genesis_input(const std::string& inp_file_name);

// The initial-state PDB model written when [OUTPUT] names a pdbfile
// (ReadGenesisInput.cpp:405-433).
void write_genesis_initial_pdb(const GenesisInput& in);

}  // namespace ocmhost
