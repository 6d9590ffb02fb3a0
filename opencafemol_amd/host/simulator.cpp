// simulator.cpp -- frame loop and trajectory/energy writers.
#include "simulator.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <random>

namespace ocmhost {

void check_ocm(int code, const char* what) {
    if (code != OCM_OK)
        throw std::runtime_error(std::string("[error] ") + what + ": " + ocm_last_error() + " (status " +
                                 std::to_string(code) + ")");
}

template <typename T> static void put(std::ofstream& o, const T& v) { o.write(reinterpret_cast<const char*>(&v), sizeof(T)); }

static std::vector<double> state_array(ocm_context* ctx, int what, std::size_t n) {
    std::vector<double> buf(3 * n);
    ocm_state st{};
    if (what == OCM_STATE_POSITIONS) st.positions = buf.data();
    else st.velocities = buf.data();
    check_ocm(ocm_get_state(ctx, what, 0xFFFFFFFFu, &st), "getState");
    return buf;
}

// ------------------------------------------------------------------- PDB
PDBObserver::PDBObserver(const std::string& prefix, const std::vector<std::optional<std::string>>& names)
    : file_(prefix + ".pdb") {
    std::ofstream(file_, std::ios::trunc);
    for (const auto& n : names) names_.push_back(n ? *n : "UNK");
    std::cerr << "    output trajectory file    : " << file_ << std::endl;
}

void PDBObserver::output(std::size_t step, ocm_context* ctx) {
    const auto x = state_array(ctx, OCM_STATE_POSITIONS, names_.size());
    std::ofstream fp(file_, std::ios::app);
    fp << "MODEL     " << step << "\n";
    for (std::size_t i = 0; i < names_.size(); ++i) {
        fp << "ATOM  " << std::setw(5) << i + 1 << " " << "  CA " << std::setw(3) << names_[i] << "  "
           << "   1    " << std::fixed << std::setprecision(3) << std::setw(8) << x[3 * i] * constant::angstrom_per_nm
           << std::setw(8) << x[3 * i + 1] * constant::angstrom_per_nm << std::setw(8)
           << x[3 * i + 2] * constant::angstrom_per_nm << "  1.00  0.00\n";
    }
    fp << "ENDMDL" << std::endl;
}

// ------------------------------------------------------------------- DCD
DCDObserver::DCDObserver(const std::string& prefix, std::size_t total_step, std::size_t save_step, float delta_t,
                         bool periodic, std::array<double, 3> box)
    : pos_file_(prefix + "_position.dcd"), vel_file_(prefix + "_velocity.dcd"), total_step_(total_step),
      save_step_(save_step), delta_t_(delta_t), periodic_(periodic), box_(box) {
    std::ofstream(pos_file_, std::ios::trunc);
    std::ofstream(vel_file_, std::ios::trunc);
    std::cerr << "    output trajectory file    : " << pos_file_ << "\n    output velocity file      : " << vel_file_ << std::endl;
}

void DCDObserver::header(std::ofstream& o) const {
    const int32_t b84 = 84, zero = 0;
    put(o, b84);
    o.write("CORD", 4);
    put(o, static_cast<int32_t>(total_step_ / save_step_));
    put(o, zero);
    put(o, static_cast<int32_t>(save_step_));
    put(o, static_cast<int32_t>(total_step_));
    put(o, zero);  // chains
    for (int k = 0; k < 4; ++k) put(o, zero);
    put(o, delta_t_);
    put(o, static_cast<int32_t>(periodic_ ? 1 : 0));
    for (int k = 0; k < 8; ++k) put(o, zero);
    put(o, static_cast<int32_t>(24));
    put(o, b84);
    put(o, b84);
    put(o, static_cast<int32_t>(1));
    char title[80];
    std::memset(title, 0, sizeof(title));
    std::snprintf(title, sizeof(title), "opencafemol_amd MD step on gfx950");
    o.write(title, 80);
    put(o, b84);
    const int32_t b4 = 4;
    put(o, b4);
    put(o, static_cast<int32_t>(n_));
    put(o, b4);
}

void DCDObserver::initialize(std::size_t n) {
    n_ = n;
    std::ofstream p(pos_file_, std::ios::binary | std::ios::app), v(vel_file_, std::ios::binary | std::ios::app);
    header(p);
    header(v);
}

static void write_blocks(std::ofstream& o, const std::vector<double>& a, std::size_t n, double scale) {
    std::vector<float> buf(n);
    const int32_t bs = static_cast<int32_t>(sizeof(float) * n);
    for (int c = 0; c < 3; ++c) {
        for (std::size_t i = 0; i < n; ++i) buf[i] = static_cast<float>(a[3 * i + c] * scale);
        put(o, bs);
        o.write(reinterpret_cast<const char*>(buf.data()), bs);
        put(o, bs);
    }
}

void DCDObserver::output(std::size_t, ocm_context* ctx) {
    const auto x = state_array(ctx, OCM_STATE_POSITIONS, n_);
    const auto v = state_array(ctx, OCM_STATE_VELOCITIES, n_);
    std::ofstream p(pos_file_, std::ios::binary | std::ios::app);
    if (periodic_) {
        const int32_t bs = sizeof(double) * 6;
        const double a = box_[0] * constant::angstrom_per_nm, b = box_[1] * constant::angstrom_per_nm;
        const double c = box_[2] * constant::angstrom_per_nm, right = 90.0;
        put(p, bs);
        put(p, a); put(p, right); put(p, b); put(p, right); put(p, right); put(p, c);
        put(p, bs);
    }
    write_blocks(p, x, n_, constant::angstrom_per_nm);
    std::ofstream vo(vel_file_, std::ios::binary | std::ios::app);
    write_blocks(vo, v, n_, constant::angstrom_per_nm * constant::cafetime());
}

// ---------------------------------------------------------------- energy
EnergyObserver::EnergyObserver(const std::string& prefix, const std::map<std::string, std::size_t>& groups)
    : file_(prefix + ".ene"), groups_(groups) {
    std::ofstream(file_, std::ios::trunc);
    std::cerr << "    output energy file        : " << file_ << std::endl;
}

void EnergyObserver::initialize(std::size_t) {
    std::ofstream o(file_, std::ios::out);
    o << "# unit of energy : kcal/mol\n# timestep";
    for (const auto& kv : groups_) o << " " << std::setw(24) << kv.first;
    o << " kinetic_energy" << std::endl;
}

void EnergyObserver::output(std::size_t step, ocm_context* ctx) {
    ocm_state st{};
    check_ocm(ocm_get_state(ctx, OCM_STATE_ENERGY, 0xFFFFFFFFu, &st), "getState(Energy)");
    std::ofstream o(file_, std::ios::app);
    o << std::setw(10) << std::left << step;
    for (const auto& kv : groups_)
        o << ' ' << std::setw(24) << std::right << std::fixed << st.energy_by_group[kv.second] / constant::kj_per_kcal;
    o << ' ' << std::setw(14) << std::right << std::fixed << st.kinetic_energy / constant::kj_per_kcal << std::endl;
}

// ------------------------------------------------------------- Simulator
Simulator::Simulator(const SystemGenerator& system_gen, std::unique_ptr<IntegratorGeneratorBase>&& integrator_gen,
                     const ocm_options& options, std::size_t total_step, std::size_t save_step,
                     std::vector<std::unique_ptr<ObserverBase>>& observers, bool energy_minimization,
                     bool output_progress)
    : system_(system_gen.generate()), integrator_gen_(std::move(integrator_gen)), total_step_(total_step),
      save_step_(save_step), energy_minimization_(energy_minimization), output_progress_(output_progress) {
    const ocm_system_desc sd = system_.view();
    const ocm_integrator_desc id = integrator_gen_->generate();
    check_ocm(ocm_context_create(&sd, &id, &options, &ctx_), "context creation");
    std::cerr << "initializing simulator...\n    total step : " << total_step_ << " step\n    save step  : "
              << save_step_ << " step\nintegrator information\n" << integrator_gen_->dump_info()
              << "\ninitializing observers..." << std::endl;
    for (auto& o : observers) {
        observers_.push_back(std::move(o));
        observers_.back()->initialize(system_.masses.size());
        std::cerr << "    " << observers_.back()->name() << std::endl;
    }
}

Simulator::Simulator(Simulator&& o) noexcept
    : system_(std::move(o.system_)), integrator_gen_(std::move(o.integrator_gen_)), ctx_(o.ctx_),
      total_step_(o.total_step_), save_step_(o.save_step_), observers_(std::move(o.observers_)),
      energy_minimization_(o.energy_minimization_), output_progress_(o.output_progress_) {
    o.ctx_ = nullptr;
}

Simulator::~Simulator() {
    if (ctx_) ocm_context_destroy(ctx_);
}

void Simulator::initialize(const std::vector<std::array<double, 3>>& positions) {
    std::vector<double> x;
    for (const auto& p : positions) x.insert(x.end(), p.begin(), p.end());
    check_ocm(ocm_set_positions(ctx_, x.data(), static_cast<int32_t>(positions.size())), "setPositions");
    if (energy_minimization_) {
        std::cerr << "energy minimization in progress..." << std::endl;
        check_ocm(ocm_minimize(ctx_, 10.0, 0), "energy minimization");
    }
}

void Simulator::set_velocity(std::int64_t vel_seed) {
    uint64_t seed = static_cast<uint64_t>(vel_seed);
    if (vel_seed == 0) {
        std::random_device rd;
        seed = (static_cast<uint64_t>(rd()) << 32) | rd();
    }
    check_ocm(ocm_set_velocities_to_temperature(ctx_, integrator_gen_->temperature(), seed), "setVelocitiesToTemperature");
}

void Simulator::set_velocity(const std::vector<std::array<double, 3>>& velocities) {
    std::vector<double> v;
    for (const auto& p : velocities) v.insert(v.end(), p.begin(), p.end());
    check_ocm(ocm_set_velocities(ctx_, v.data(), static_cast<int32_t>(velocities.size())), "setVelocities");
}

// src/Simulator.cpp:49-76: frames are written before each block of save_step
// steps, so the state after the final block is not written.
void Simulator::run() {
    const std::size_t frames = total_step_ / save_step_;
    for (std::size_t f = 0; f < frames; ++f) {
        for (auto& o : observers_) o->output(f * save_step_, ctx_);
        check_ocm(ocm_step(ctx_, static_cast<int64_t>(save_step_)), "step");
        if (output_progress_) {
            const double ratio = std::min(1.0, double(f * save_step_) / double(total_step_));
            std::fprintf(stderr, "\r%5.1f%%", ratio * 100.0);
        }
    }
    for (auto& o : observers_) o->finalize();
    if (output_progress_) std::fprintf(stderr, "\r%5.1f%%\n", 100.0);
}

}  // namespace ocmhost
