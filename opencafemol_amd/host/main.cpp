// main.cpp -- command-line driver: ocm_run <input.toml> or <input.inp>
// (the reference's src/OpenAICG2Plus.cpp:16-94 with OpenMM plugin loading
// replaced by the gfx950 device step).
#include "read_genesis.hpp"
#include "read_toml.hpp"
#include "simulator.hpp"

#include <chrono>
#include <cstdio>
#include <iostream>

using namespace ocmhost;

static void note_platform(const std::string& name) {
    std::cerr << "    platform \"" << name << "\" requested; running on the gfx950 device step" << std::endl;
}

// ReadGenesisInput.cpp:217-503: no minimization and no velocity setting, so
// the run starts from zero velocities; DCD and .ene outputs share the
// dcdfile prefix, and the DCD header carries the timestep in ps.
static Simulator make_genesis_simulator(const std::string& path) {
    GenesisInput in = read_genesis_input(path);
    if (!in.pdbfile.empty()) write_genesis_initial_pdb(in);
    const SystemDescriptor probe = in.system.generate();
    std::vector<std::unique_ptr<ObserverBase>> obs;
    obs.push_back(std::make_unique<DCDObserver>(in.file_prefix, in.nsteps, in.crdout_period,
                                                static_cast<float>(in.timestep), in.periodic, probe.box));
    obs.push_back(std::make_unique<EnergyObserver>(in.file_prefix, in.system.ffname_groupid_map()));
    if (!in.platform.empty()) note_platform(in.platform);
    ocm_options opt;
    ocm_options_default(&opt);
    Simulator sim(in.system, std::move(in.integrator), opt, in.nsteps, in.crdout_period, obs);
    sim.initialize(in.positions);
    return sim;
}

static Simulator make_simulator(const std::string& path) {
    std::cerr << "parsing " << path << "..." << std::endl;
    const Value data = load_toml_input(path);
    const RunSettings rs = read_toml_run_settings(data);
    const auto positions = read_toml_initial_conf(data);
    SystemGenerator sys = read_toml_system(data);
    auto integ = read_toml_integrator_gen(data);
    std::vector<std::unique_ptr<ObserverBase>> obs;
    const std::string prefix = rs.output_path + rs.output_prefix;
    const SystemDescriptor probe = sys.generate();
    if (rs.output_format == "pdb") obs.push_back(std::make_unique<PDBObserver>(prefix, rs.names));
    else obs.push_back(std::make_unique<DCDObserver>(prefix, rs.total_step, rs.save_step,
                                                     static_cast<float>(rs.delta_t_cafe), rs.periodic, probe.box));
    obs.push_back(std::make_unique<EnergyObserver>(prefix, sys.ffname_groupid_map()));
    if (const Value* plat = data.find("platform"))
        if (const Value* name = plat->find("name"))
            note_platform(name->as_string());
    ocm_options opt;
    ocm_options_default(&opt);
    Simulator sim(sys, std::move(integ), opt, rs.total_step, rs.save_step, obs, rs.energy_minimization, rs.progress_bar);
    sim.initialize(positions);
    const auto vel = read_toml_initial_vel(data);
    if (!vel.empty()) {
        sim.set_velocity(vel);
        std::cerr << "initial velocity was set based on the given value." << std::endl;
    } else {
        if (rs.vel_seed == 0)
            std::cerr << "system does not provide initial velocity nor non-zero `vel_seed`. initial velocity becomes "
                         "completely random and cannot be reproduced." << std::endl;
        sim.set_velocity(rs.vel_seed);
    }
    return sim;
}

int main(int argc, char** argv) {
    std::cerr << "opencafemol_amd " << ocm_version() << std::endl;
    const std::string arg = argc == 2 ? argv[1] : "";
    const auto dot = arg.rfind('.');
    const std::string suffix = dot == std::string::npos ? "" : arg.substr(dot);
    if (suffix != ".toml" && suffix != ".inp") {
        std::fprintf(stderr, "Usage: %s <input.toml> or <input.inp>\n", argv[0]);
        return 1;
    }
    try {
        Simulator sim = suffix == ".inp" ? make_genesis_simulator(arg) : make_simulator(arg);
        const auto start = std::chrono::system_clock::now();
        std::cerr << "calculation start!" << std::endl;
        sim.run();
        const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::system_clock::now() - start).count();
        if (ms < 1000) std::fprintf(stderr, "elapsed time : %lld [msec]\n", static_cast<long long>(ms));
        else std::fprintf(stderr, "elapsed time : %.1f [sec]\n", ms * 0.001);
        return 0;
    } catch (const std::exception& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
}
