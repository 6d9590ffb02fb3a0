// host_capi.cpp -- C entry points of include/ocm_host.h.
#include "../../include/ocm_host.h"
#include "read_genesis.hpp"
#include "read_toml.hpp"

#include <string>

struct ocmh_input {
    ocmhost::SystemDescriptor system;
    ocm_system_desc view;
    ocm_integrator_desc integ;
    std::vector<double> positions, velocities;
    std::map<std::string, std::size_t> groups;
};

namespace {
thread_local std::string g_err;
}

extern "C" {

int ocmh_load_toml(const char* path, ocmh_input** out) {
    try {
        g_err.clear();
        if (!path || !out) throw std::runtime_error("[error] NULL argument");
        const ocmhost::Value data = ocmhost::load_toml_input(path);
        auto* in = new ocmh_input();
        try {
            ocmhost::SystemGenerator gen = ocmhost::read_toml_system(data);
            in->system = gen.generate();
            in->groups = gen.ffname_groupid_map();
            in->view = in->system.view();
            in->integ = ocmhost::read_toml_integrator_gen(data)->generate();
            for (const auto& p : ocmhost::read_toml_initial_conf(data)) in->positions.insert(in->positions.end(), p.begin(), p.end());
            for (const auto& v : ocmhost::read_toml_initial_vel(data)) in->velocities.insert(in->velocities.end(), v.begin(), v.end());
        } catch (...) {
            delete in;
            throw;
        }
        *out = in;
        return OCM_OK;
    } catch (const std::exception& e) {
        g_err = e.what();
        return OCM_ERR_INVALID;
    }
}

int ocmh_load_genesis(const char* inp_path, ocmh_input** out) {
    try {
        g_err.clear();
        if (!inp_path || !out) throw std::runtime_error("[error] NULL argument");
        ocmhost::GenesisInput g = ocmhost::read_genesis_input(inp_path);
        auto* in = new ocmh_input();
        try {
            in->system = g.system.generate();
            in->groups = g.system.ffname_groupid_map();
            in->view = in->system.view();
            in->integ = g.integrator->generate();
            for (const auto& p : g.positions) in->positions.insert(in->positions.end(), p.begin(), p.end());
        } catch (...) {
            delete in;
            throw;
        }
        *out = in;
        return OCM_OK;
    } catch (const std::exception& e) {
        g_err = e.what();
        return OCM_ERR_INVALID;
    }
}

void ocmh_free(ocmh_input* in) { delete in; }
const ocm_system_desc* ocmh_system(ocmh_input* in) { return in ? &in->view : nullptr; }
const ocm_integrator_desc* ocmh_integrator(ocmh_input* in) { return in ? &in->integ : nullptr; }
const double* ocmh_positions(ocmh_input* in) { return in && !in->positions.empty() ? in->positions.data() : nullptr; }
const double* ocmh_velocities(ocmh_input* in) { return in && !in->velocities.empty() ? in->velocities.data() : nullptr; }
int32_t ocmh_group_of(ocmh_input* in, const char* name) {
    if (!in || !name) return -1;
    const auto it = in->groups.find(name);
    return it == in->groups.end() ? -1 : static_cast<int32_t>(it->second);
}
const char* ocmh_last_error(void) { return g_err.c_str(); }

}  // extern "C"
