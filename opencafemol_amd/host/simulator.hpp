// simulator.hpp -- frame loop over the device step and its observers
// (src/Simulator.{hpp,cpp}, src/Observer.{hpp,cpp}).
#pragma once

#include "forcefield.hpp"

#include <array>
#include <cstdint>
#include <fstream>
#include <memory>
#include <optional>
#include <string>
#include <vector>

namespace ocmhost {

class ObserverBase {
public:
    virtual ~ObserverBase() = default;
    virtual void initialize(std::size_t n_particles) = 0;
    virtual void output(std::size_t step, ocm_context* ctx) = 0;
    virtual void finalize() const {}
    virtual std::string name() const = 0;
};

// PDB MODEL frames of CA atoms in Angstrom (src/Observer.cpp:7-62).
class PDBObserver final : public ObserverBase {
public:
    PDBObserver(const std::string& prefix, const std::vector<std::optional<std::string>>& names);
    void initialize(std::size_t) override {}
    void output(std::size_t step, ocm_context* ctx) override;
    std::string name() const override { return "PDBObserver"; }
private:
    std::string file_;
    std::vector<std::string> names_;
};

// DCD position (Angstrom) and velocity (Angstrom / cafetime) trajectories
// (src/Observer.cpp:66-298): 84-byte CORD header, 80-character title block,
// atom count, then per frame an optional unit cell block (A, gamma, B, beta,
// alpha, C) and three float32 coordinate blocks.
class DCDObserver final : public ObserverBase {
public:
    DCDObserver(const std::string& prefix, std::size_t total_step, std::size_t save_step, float delta_t,
                bool periodic, std::array<double, 3> box);
    void initialize(std::size_t n) override;
    void output(std::size_t step, ocm_context* ctx) override;
    std::string name() const override { return "DCDObserver"; }
private:
    void header(std::ofstream& ofs) const;
    std::string pos_file_, vel_file_;
    std::size_t total_step_, save_step_, n_ = 0;
    float delta_t_;
    bool periodic_;
    std::array<double, 3> box_;
};

// Per-group potential energies and kinetic energy in kcal/mol, one line per
// frame, columns in the name order of the group map (src/Observer.cpp:302-347).
class EnergyObserver final : public ObserverBase {
public:
    EnergyObserver(const std::string& prefix, const std::map<std::string, std::size_t>& groups);
    void initialize(std::size_t) override;
    void output(std::size_t step, ocm_context* ctx) override;
    std::string name() const override { return "EnergyObserver"; }
private:
    std::string file_;
    std::map<std::string, std::size_t> groups_;
};

class Simulator {
public:
    Simulator(const SystemGenerator& system_gen, std::unique_ptr<IntegratorGeneratorBase>&& integrator_gen,
              const ocm_options& options, std::size_t total_step, std::size_t save_step,
              std::vector<std::unique_ptr<ObserverBase>>& observers, bool energy_minimization = false,
              bool output_progress = true);
    ~Simulator();
    Simulator(Simulator&&) noexcept;
    Simulator(const Simulator&) = delete;

    void initialize(const std::vector<std::array<double, 3>>& positions);
    void set_velocity(std::int64_t vel_seed);
    void set_velocity(const std::vector<std::array<double, 3>>& velocities);
    void run();

    ocm_context* context() const { return ctx_; }
    std::size_t total_step() const { return total_step_; }
    std::size_t save_step() const { return save_step_; }

private:
    SystemDescriptor system_;
    std::unique_ptr<IntegratorGeneratorBase> integrator_gen_;
    ocm_context* ctx_ = nullptr;
    std::size_t total_step_, save_step_;
    std::vector<std::unique_ptr<ObserverBase>> observers_;
    bool energy_minimization_, output_progress_;
};

void check_ocm(int code, const char* what);

}  // namespace ocmhost
