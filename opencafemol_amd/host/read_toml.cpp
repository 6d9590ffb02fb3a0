// read_toml.cpp -- TOML readers for the MD-step potentials.  Unit
// conversions and defaults follow src/input/ReadTOMLForceFieldGenerator.cpp:
// positions and lengths A -> nm, energies kcal/mol -> kJ/mol, harmonic k
// doubled (the TOML convention is k (x - x0)^2, OpenMM's 1/2 k (x - x0)^2).
#include "read_toml.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>

namespace ocmhost {

namespace {
constexpr double KJ = constant::kj_per_kcal;
constexpr double NM = constant::nm_per_angstrom;

const Value* env_of(const Value& ff) { return ff.find("env"); }

std::string kind_name(const Value& v) {
    switch (v.kind) {
    case Value::Bool: return "boolean";
    case Value::Int: return "integer";
    case Value::Float: return "floating";
    case Value::String: return "string";
    case Value::ArrayK: return "array";
    case Value::TableK: return "table";
    default: return "value";
    }
}

void range_check(const char* who, std::size_t idx, std::size_t n) {
    if (idx >= n)
        throw std::runtime_error(std::string("[error] ") + who + " : index " + std::to_string(idx) +
                                 " exceeds the system's largest index " + std::to_string(n - 1) + ".");
}

const Value& resolve(const Value& p, const Value* env, const std::string& name) {
    if (!p.is_string()) return p;
    const std::string& var = p.as_string();
    if (!env) throw std::out_of_range("[error] named variable \"" + var + "\" used but no env is defined" + p.where());
    const Value* v = env->find(var);
    if (!v) throw std::out_of_range("[error] named variable \"" + var + "\" does not exist in env" + p.where());
    (void)name;
    return *v;
}

long long offset_of(const Value& param, const Value* env, std::size_t k) {
    const Value* o = param.find("offset");
    if (!o) return 0;
    const Value& v = resolve(*o, env, "offset");
    if (v.is_array()) {
        if (k >= v.as_array().size()) throw std::runtime_error("[error] add_offset: offset array is too short" + v.where());
        return v.as_array()[k].as_int();
    }
    return v.as_int();
}

template <std::size_t N>
std::array<std::size_t, N> read_indices(const Value& param, const Value* env, const char* who, std::size_t n) {
    const Value& arr = resolve(param.at("indices"), env, "indices");
    const auto& a = arr.as_array();
    if (a.size() != N) throw std::runtime_error(std::string("[error] ") + who + ": indices needs " + std::to_string(N) + " entries" + arr.where());
    const Value* o = param.find("offset");
    if (o && resolve(*o, env, "offset").is_array() && resolve(*o, env, "offset").as_array().size() != N)
        throw std::runtime_error("[error] add_offset: invalid size of offset array here. the size of offset array must much to the size of indices array.");
    std::array<std::size_t, N> out{};
    for (std::size_t k = 0; k < N; ++k) {
        const long long v = a[k].as_int() + offset_of(param, env, k);
        if (v < 0) throw std::runtime_error(std::string("[error] ") + who + " : negative index");
        out[k] = static_cast<std::size_t>(v);
        range_check(who, out[k], n);
    }
    return out;
}

std::size_t read_index(const Value& param, const Value* env, const char* who, std::size_t n) {
    const Value* o = param.find("offset");
    if (o && resolve(*o, env, "offset").is_array())
        throw std::runtime_error("[error] add_offset: offset array is invalid here. expected type is scalar.");
    const long long v = resolve(param.at("index"), env, "index").as_int() + offset_of(param, env, 0);
    if (v < 0) throw std::runtime_error(std::string("[error] ") + who + " : negative index");
    range_check(who, static_cast<std::size_t>(v), n);
    return static_cast<std::size_t>(v);
}

double num(const Value& param, const Value* env, const std::string& name) {
    return find_parameter(param, env, name).as_double();
}

std::string topology_type(const Value& ff) {
    const Value* t = ff.find("topology");
    return t ? t->as_string() : std::string();
}

void warn_unknown_key(const std::string& key, const Value& v) {
    std::cerr << "\033[33m[warning]\033[m unknown value \"" << key << "\" found. this " << kind_name(v)
              << " will never be used." << std::endl;
}
}  // namespace

// ------------------------------------------------------------ helpers
const Value& find_parameter(const Value& params, const Value* env, const std::string& name) {
    const Value* p = params.find(name);
    if (!p) throw std::out_of_range("[error] value " + name + " does not exists" + params.where());
    return resolve(*p, env, name);
}

const Value& find_parameter_either(const Value& params, const Value* env, const std::string& n1, const std::string& n2) {
    const Value* a = params.find(n1);
    const Value* b = params.find(n2);
    if (a && b) throw std::out_of_range("[error] key duplicates: " + n1 + " and " + n2 + params.where());
    if (!a && !b) throw std::out_of_range("[error] value " + n1 + " and " + n2 + " does not exists" + params.where());
    return resolve(a ? *a : *b, env, a ? n1 : n2);
}

double find_double_or(const Value& params, const Value* env, const std::string& name, double fallback) {
    const Value* p = params.find(name);
    if (!p) return fallback;
    if (p->is_string()) {
        if (!env) return fallback;
        const Value* v = env->find(p->as_string());
        return v ? v->as_double() : fallback;
    }
    return p->as_double();
}

bool check_keys_available(const Value& table, std::initializer_list<std::string> keys) {
    bool ok = true;
    for (const auto& kv : table.as_table()) {
        if (std::find(keys.begin(), keys.end(), kv.first) == keys.end()) {
            warn_unknown_key(kv.first, kv.second);
            ok = false;
        }
    }
    return ok;
}

// src/input/Utility.cpp:7-80: "include" files are merged into the table that
// names them; arrays are concatenated, other duplicates are errors.
static void merge_tables(Value& table, const Value& other) {
    for (const auto& kv : other.as_table()) {
        Value* mine = table.find_mut(kv.first);
        if (!mine) {
            table.insert(kv.first, kv.second);
        } else if (mine->is_table() && kv.second.is_table()) {
            merge_tables(*mine, kv.second);
        } else if (mine->is_array() && kv.second.is_array()) {
            auto merged = std::make_shared<ocmtoml::Array>(*mine->a);
            merged->insert(merged->end(), kv.second.a->begin(), kv.second.a->end());
            mine->a = merged;
        } else {
            throw std::runtime_error("[error] value \"" + kv.first + "\" duplicates.");
        }
    }
}

void expand_include(Value& v, const std::string& base_dir) {
    if (v.is_array()) {
        for (auto& e : *v.a) expand_include(e, base_dir);
        return;
    }
    if (!v.is_table()) return;
    for (auto& kv : *v.t) expand_include(kv.second, base_dir);
    Value* inc = v.find_mut("include");
    if (!inc) return;
    std::vector<std::string> files;
    if (inc->is_array()) for (const auto& f : inc->as_array()) files.push_back(f.as_string());
    else files.push_back(inc->as_string());
    auto& t = *v.t;
    t.erase(std::remove_if(t.begin(), t.end(), [](const auto& kv) { return kv.first == "include"; }), t.end());
    for (const auto& f : files) {
        std::string path = f;
        if (!std::ifstream(path) && !base_dir.empty()) path = base_dir + "/" + f;
        std::cerr << "    expanding file " << path << std::endl;
        merge_tables(v, ocmtoml::parse_file(path));
    }
}

index_pairs_type read_ignore_molecule_and_particles_within(const Value& ignore, const Topology& topology) {
    index_pairs_type out;
    bool molecule = false;
    if (const Value* m = ignore.find("molecule")) {
        const std::string& name = m->as_string();
        if (name == "Intra" || name == "Self") {
            molecule = true;
            const auto l = topology.ignore_list_within_molecule();
            out.insert(out.end(), l.begin(), l.end());
        } else if (name == "Others" || name == "Inter") {
            throw std::runtime_error("[error] ignore molecule do not support \"Others\" or \"Inter\".");
        }
    }
    if (const Value* pw = ignore.find("particles_within")) {
        if (molecule)
            std::cerr << "\033[33m[warning]\033[m ignore molecule \"Intra\" or \"Self\" was defined,"
                         "so this ignore particle within bond will be ignored." << std::endl;
        for (const auto& kv : pw->as_table()) {
            const auto l = topology.ignore_list_within_edge(static_cast<std::size_t>(kv.second.as_int()), kv.first);
            out.insert(out.end(), l.begin(), l.end());
        }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

group_pairs_type read_ignore_group(const Value& ignore) {
    group_pairs_type out;
    const Value* g = ignore.find("group");
    if (!g) return out;
    if (const Value* inter = g->find("inter")) {
        for (const auto& p : inter->as_array()) {
            const auto& a = p.as_array();
            if (a.size() != 2) throw std::runtime_error("[error] ignore.group.inter entries need two names");
            out.emplace_back(a[0].as_string(), a[1].as_string());
        }
    }
    if (const Value* intra = g->find("intra"))
        for (const auto& p : intra->as_array()) out.emplace_back(p.as_string(), p.as_string());
    return out;
}

// ------------------------------------------------------------ local readers
std::unique_ptr<HarmonicBondForceFieldGenerator> read_toml_harmonic_bond_ff_generator(const Value& ff, Topology& topology,
                                                                                       bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    index_pairs_type idx;
    std::vector<double> v0s, ks;
    for (const auto& p : ff.at("parameters").as_array()) {
        const auto i = read_indices<2>(p, env, "read_toml_harmonic_bond_ff_generator", topology.size());
        idx.emplace_back(i[0], i[1]);
        v0s.push_back(num(p, env, "v0") * NM);
        ks.push_back(num(p, env, "k") * KJ * 100.0 * 2.0);
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    BondLength    : Harmonic (" << idx.size() << " found)" << std::endl;
    return std::make_unique<HarmonicBondForceFieldGenerator>(idx, v0s, ks, use_periodic);
}

std::unique_ptr<GaussianBondForceFieldGenerator> read_toml_gaussian_bond_ff_generator(const Value& ff, Topology& topology,
                                                                                       bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    index_pairs_type idx;
    std::vector<double> ks, v0s, sigmas;
    for (const auto& p : ff.at("parameters").as_array()) {
        const auto i = read_indices<2>(p, env, "read_toml_gaussian_bond_ff_generator", topology.size());
        idx.emplace_back(i[0], i[1]);
        ks.push_back(num(p, env, "k") * KJ);
        v0s.push_back(num(p, env, "v0") * NM);
        sigmas.push_back(find_parameter_either(p, env, "sigma", "\xCF\x83").as_double() * NM);
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    BondLength    : Gaussian (" << idx.size() << " found)" << std::endl;
    return std::make_unique<GaussianBondForceFieldGenerator>(idx, ks, v0s, sigmas, use_periodic);
}

std::unique_ptr<GoContactForceFieldGenerator> read_toml_go_contact_ff_generator(const Value& ff, Topology& topology,
                                                                                 bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    index_pairs_type idx;
    std::vector<double> ks, r0s;
    for (const auto& p : ff.at("parameters").as_array()) {
        const auto i = read_indices<2>(p, env, "read_toml_go_contact_ff_generator", topology.size());
        idx.emplace_back(i[0], i[1]);
        ks.push_back(num(p, env, "k") * KJ);
        r0s.push_back(num(p, env, "v0") * NM);
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    BondLength    : GoContact (" << idx.size() << " found)" << std::endl;
    return std::make_unique<GoContactForceFieldGenerator>(idx, ks, r0s, use_periodic);
}

// ReadTOMLForceFieldGenerator.cpp:216-275: the Go contact's keys plus
// capping_ratio (default 0.75).
std::unique_ptr<CappedGoContactForceFieldGenerator> read_toml_capped_go_contact_ff_generator(const Value& ff,
                                                                                             Topology& topology,
                                                                                             bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "capping_ratio", "env"});
    const Value* env = env_of(ff);
    const double capping_ratio = ff.contains("capping_ratio") ? ff.at("capping_ratio").as_double() : 0.75;
    index_pairs_type idx;
    std::vector<double> ks, r0s;
    for (const auto& p : ff.at("parameters").as_array()) {
        const auto i = read_indices<2>(p, env, "read_toml_capped_go_contact_ff_generator", topology.size());
        idx.emplace_back(i[0], i[1]);
        ks.push_back(num(p, env, "k") * KJ);
        r0s.push_back(num(p, env, "v0") * NM);
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    BondLength    : CappedGoContact (" << idx.size() << " found)" << std::endl;
    return std::make_unique<CappedGoContactForceFieldGenerator>(idx, ks, r0s, capping_ratio, use_periodic);
}

// ReadTOMLForceFieldGenerator.cpp:549-622: k in kcal/mol, t0 = v0 + pi (the
// input's convention is k (1 + cos(n (phi - v0)))), integer n.
std::unique_ptr<CosineDihedralForceFieldGenerator> read_toml_cosine_dihedral_ff_generator(const Value& ff,
                                                                                           Topology& topology,
                                                                                           bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::array<std::size_t, 4>> idx;
    std::vector<double> ks, t0s, ns;
    for (const auto& p : ff.at("parameters").as_array()) {
        idx.push_back(read_indices<4>(p, env, "read_toml_cosine_dihedral_ff_generator", topology.size()));
        ks.push_back(num(p, env, "k") * KJ);
        t0s.push_back(num(p, env, "v0") + constant::pi);
        const long long n = find_parameter(p, env, "n").as_int();
        if (n < 0) throw std::runtime_error("[error] read_toml_cosine_dihedral_ff_generator: n must not be negative" + p.where());
        ns.push_back(static_cast<double>(n));
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    DihedralAngle : Cosine (" << idx.size() << " found)" << std::endl;
    return std::make_unique<CosineDihedralForceFieldGenerator>(idx, ks, t0s, ns, use_periodic);
}

std::unique_ptr<HarmonicAngleForceFieldGenerator> read_toml_harmonic_angle_ff_generator(const Value& ff, Topology& topology,
                                                                                         bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::array<std::size_t, 3>> idx;
    std::vector<double> v0s, ks;
    for (const auto& p : ff.at("parameters").as_array()) {
        idx.push_back(read_indices<3>(p, env, "read_toml_harmonic_angle_ff_generator", topology.size()));
        double v0 = num(p, env, "v0");
        if (v0 < 0 || constant::pi * 2 < v0)
            throw std::runtime_error("[error] read_toml_harmonic_angle_ff_generator: v0 must be between 0 and 2pi");
        if (constant::pi < v0) v0 = 2.0 * constant::pi - v0;
        v0s.push_back(v0);
        ks.push_back(num(p, env, "k") * KJ * 2.0);
    }
    std::cerr << "    BondAngle     : Harmonic (" << idx.size() << " found)" << std::endl;
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    return std::make_unique<HarmonicAngleForceFieldGenerator>(idx, v0s, ks, use_periodic);
}

std::unique_ptr<FlexibleLocalAngleForceFieldGenerator> read_toml_flexible_local_angle_ff_generator(
    const Value& ff, const std::string& aa_type, const std::array<double, 10>& table, Topology& topology,
    bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::array<std::size_t, 3>> idx;
    std::vector<double> ks;
    for (const auto& p : ff.at("parameters").as_array()) {
        const std::string& y = p.at("y").as_string();  // like "y1_PHE"
        if (y.size() < 6 || y.substr(3, 3) != aa_type) continue;
        idx.push_back(read_indices<3>(p, env, "read_toml_flexible_local_angle_ff_generator", topology.size()));
        ks.push_back(num(p, env, "k") * KJ);
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    BondAngle     : FlexibleLocalAngle - " << aa_type << " (" << idx.size() << " found)" << std::endl;
    return std::make_unique<FlexibleLocalAngleForceFieldGenerator>(idx, ks, table, aa_type, use_periodic);
}

std::unique_ptr<GaussianDihedralForceFieldGenerator> read_toml_gaussian_dihedral_ff_generator(const Value& ff,
                                                                                               Topology& topology,
                                                                                               bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::array<std::size_t, 4>> idx;
    std::vector<double> ks, th0s, sigmas;
    for (const auto& p : ff.at("parameters").as_array()) {
        idx.push_back(read_indices<4>(p, env, "read_toml_gaussian_dihedral_ff_generator", topology.size()));
        ks.push_back(num(p, env, "k") * KJ);
        th0s.push_back(num(p, env, "v0"));
        sigmas.push_back(find_parameter_either(p, env, "sigma", "\xCF\x83").as_double());
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    std::cerr << "    DihedralAngle : Gaussian (" << idx.size() << " found)" << std::endl;
    return std::make_unique<GaussianDihedralForceFieldGenerator>(idx, ks, th0s, sigmas, use_periodic);
}

// The residue-pair class of a dihedral follows its coef ("AAA-BBB"):
// R1-GLY takes every X-GLY; GLY-PRO takes coef "GLY-PRO"; R2-PRO takes every
// X-PRO (so GLY-PRO dihedrals carry both PRO terms, as in the reference);
// otherwise the first residue's class with R3
// (ReadTOMLForceFieldGenerator.cpp:636-766).
std::unique_ptr<FlexibleLocalDihedralForceFieldGenerator> read_toml_flexible_local_dihedral_ff_generator(
    const Value& ff, const std::pair<std::string, std::string>& aa_pair, const std::array<double, 7>& table,
    Topology& topology, bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "topology", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::array<std::size_t, 4>> idx;
    std::vector<double> ks;
    for (const auto& p : ff.at("parameters").as_array()) {
        const std::string& coef = p.at("coef").as_string();
        if (coef.size() < 7) throw std::runtime_error("[error] FlexibleLocalDihedral coef must look like AAA-BBB");
        const std::string first = coef.substr(0, 3), second = coef.substr(4, 3);
        bool take;
        if (aa_pair.second == "GLY") take = second == "GLY";
        else if (aa_pair.second == "PRO") take = aa_pair.first == "GLY" ? coef == "GLY-PRO" : second == "PRO";
        else take = second != "GLY" && second != "PRO" && first == aa_pair.first;
        if (!take) continue;
        idx.push_back(read_indices<4>(p, env, "read_toml_flexible_local_dihedral_ff_generator", topology.size()));
        ks.push_back(num(p, env, "k") * KJ);
    }
    if (!topology_type(ff).empty()) topology.add_edges(idx, topology_type(ff));
    const std::string name = aa_pair.first + "-" + aa_pair.second;
    std::cerr << "    DihedralAngle : FlexibleLocalDihedral - " << name << " (" << idx.size() << " found)" << std::endl;
    return std::make_unique<FlexibleLocalDihedralForceFieldGenerator>(idx, ks, table, name, use_periodic);
}

// ----------------------------------------------------------- global readers
std::unique_ptr<ExcludedVolumeForceFieldGenerator> read_toml_excluded_volume_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "env", "cutoff", "parameters", "epsilon"});
    const Value* env = env_of(ff);
    const double eps = ff.at("epsilon").as_double() * KJ;
    const double cutoff = ff.contains("cutoff") ? ff.at("cutoff").as_double() : 2.0;
    std::vector<std::optional<double>> radii(n);
    const auto& params = ff.at("parameters").as_array();
    for (const auto& p : params)
        radii[read_index(p, env, "read_toml_excluded_volume_ff_generator", topology.size())] = num(p, env, "radius") * NM;
    std::cerr << "    Global        : ExcludedVolume (" << params.size() << " found)" << std::endl;
    index_pairs_type ignore;
    group_pairs_type gp;
    if (const Value* ig = ff.find("ignore")) {
        ignore = read_ignore_molecule_and_particles_within(*ig, topology);
        gp = read_ignore_group(*ig);
    }
    return std::make_unique<ExcludedVolumeForceFieldGenerator>(eps, cutoff, radii, ignore, use_periodic, gp, groups);
}

std::unique_ptr<DebyeHuckelForceFieldGenerator> read_toml_debye_huckel_ff_generator(
    const Value& ff, std::size_t n, double ionic_strength, double temperature, const Topology& topology,
    const std::vector<std::optional<std::string>>& groups, bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "cutoff", "parameters", "env"});
    const Value* env = env_of(ff);
    const double cutoff = ff.contains("cutoff") ? ff.at("cutoff").as_double() : 5.5;
    std::vector<std::optional<double>> charges(n);
    const auto& params = ff.at("parameters").as_array();
    for (const auto& p : params)
        charges[read_index(p, env, "read_toml_debye_huckel_ff_generator", topology.size())] = num(p, env, "charge");
    std::cerr << "    Global        : DebyeHuckel (" << params.size() << " found)" << std::endl;
    index_pairs_type ignore;
    group_pairs_type gp;
    if (const Value* ig = ff.find("ignore")) {
        ignore = read_ignore_molecule_and_particles_within(*ig, topology);
        gp = read_ignore_group(*ig);
    }
    return std::make_unique<DebyeHuckelForceFieldGenerator>(ionic_strength, temperature, cutoff, charges, ignore,
                                                            use_periodic, gp, groups);
}

template <typename Gen>
static std::unique_ptr<Gen> read_lj(const Value& ff, std::size_t n, const Topology& topology,
                                    const std::vector<std::optional<std::string>>& groups, bool use_periodic,
                                    const char* who, const char* label) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "cutoff", "parameters", "env"});
    const Value* env = env_of(ff);
    const double ratio = find_double_or(ff, env, "cutoff", 2.5);
    std::vector<std::optional<double>> eps(n), sig(n);
    const auto& params = ff.at("parameters").as_array();
    for (const auto& p : params) {
        const std::size_t i = read_index(p, env, who, topology.size());
        sig[i] = num(p, env, "sigma") * NM;
        eps[i] = num(p, env, "epsilon") * KJ;
    }
    std::cerr << "    Global        : " << label << " (" << params.size() << " found)" << std::endl;
    index_pairs_type ignore;
    group_pairs_type gp;
    if (const Value* ig = ff.find("ignore")) {
        ignore = read_ignore_molecule_and_particles_within(*ig, topology);
        gp = read_ignore_group(*ig);
    }
    return std::make_unique<Gen>(ratio, eps, sig, ignore, use_periodic, gp, groups);
}

std::unique_ptr<LennardJonesAttractiveForceFieldGenerator> read_toml_lennard_jones_attractive_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic) {
    return read_lj<LennardJonesAttractiveForceFieldGenerator>(ff, n, topology, groups, use_periodic,
                                                              "read_toml_lennard_jones_attractive_ff_generator",
                                                              "LennardJonesAttractive");
}

std::unique_ptr<LennardJonesRepulsiveForceFieldGenerator> read_toml_lennard_jones_repulsive_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic) {
    return read_lj<LennardJonesRepulsiveForceFieldGenerator>(ff, n, topology, groups, use_periodic,
                                                             "read_toml_lennard_jones_repulsive_ff_generator",
                                                             "LennardJonesRepulsive");
}

static void read_ignore(const Value& ff, const Topology& topology, index_pairs_type& ignore, group_pairs_type& gp) {
    if (const Value* ig = ff.find("ignore")) {
        ignore = read_ignore_molecule_and_particles_within(*ig, topology);
        gp = read_ignore_group(*ig);
    }
}

// ReadTOMLForceFieldGenerator.cpp:1225-1296: per-bead sigma (or "σ", Å) and
// epsilon (or "ε", kcal/mol).
std::unique_ptr<WeeksChandlerAndersenForceFieldGenerator> read_toml_weeks_chandler_andersen_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::optional<double>> sig(n), eps(n);
    const auto& params = ff.at("parameters").as_array();
    for (const auto& p : params) {
        const std::size_t i = read_index(p, env, "read_toml_weeks_chandler_andersen_ff_generator", topology.size());
        sig[i] = find_parameter_either(p, env, "sigma", "\xCF\x83").as_double() * NM;
        eps[i] = find_parameter_either(p, env, "epsilon", "\xCE\xB5").as_double() * KJ;
    }
    std::cerr << "    Global        : WeeksChandlerAndersen (" << params.size() << " found)" << std::endl;
    index_pairs_type ignore;
    group_pairs_type gp;
    read_ignore(ff, topology, ignore, gp);
    return std::make_unique<WeeksChandlerAndersenForceFieldGenerator>(sig, eps, ignore, use_periodic, gp, groups);
}

// ReadTOMLForceFieldGenerator.cpp:1423-1494: per-bead sigma, omega (Å) and
// epsilon (kcal/mol).
std::unique_ptr<iSoLFAttractiveForceFieldGenerator> read_toml_isolf_attractive_ff_generator(
    const Value& ff, std::size_t n, const Topology& topology, const std::vector<std::optional<std::string>>& groups,
    bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "parameters", "env"});
    const Value* env = env_of(ff);
    std::vector<std::optional<double>> sig(n), eps(n), om(n);
    const auto& params = ff.at("parameters").as_array();
    for (const auto& p : params) {
        const std::size_t i = read_index(p, env, "read_toml_isolf_attractive_ff_generator", topology.size());
        sig[i] = num(p, env, "sigma") * NM;
        eps[i] = num(p, env, "epsilon") * KJ;
        om[i] = num(p, env, "omega") * NM;
    }
    std::cerr << "    Global        : iSoLFAttractive (" << params.size() << " found)" << std::endl;
    index_pairs_type ignore;
    group_pairs_type gp;
    read_ignore(ff, topology, ignore, gp);
    return std::make_unique<iSoLFAttractiveForceFieldGenerator>(sig, eps, om, ignore, use_periodic, gp, groups);
}

// The table potentials: for every entry table.<first>.<second> = {sigma,
// epsilon}, the beads named <first> and <second> in `parameters`.  A pair
// listed twice is an error; a pair with no bead on one side is skipped.
template <typename Gen, typename Make>
static std::vector<std::unique_ptr<Gen>> read_uniform(const Value& ff, std::size_t n, const Topology& topology,
                                                      const char* who, const char* label, Make make) {
    const Value* env = env_of(ff);
    index_pairs_type ignore;
    group_pairs_type gp;
    read_ignore(ff, topology, ignore, gp);
    std::cerr << "    Global        : " << label << std::endl;
    std::vector<std::pair<std::string, std::string>> treated;
    std::vector<std::unique_ptr<Gen>> out;
    const auto& params = ff.at("parameters").as_array();
    for (const auto& first : ff.at("table").as_table()) {
        for (const auto& second : first.second.as_table()) {
            const auto name_pair = std::make_pair(first.first, second.first);
            if (std::find(treated.begin(), treated.end(), name_pair) != treated.end())
                throw std::runtime_error("[error] parameter table for " + name_pair.first + "-" + name_pair.second +
                                         " is defined in dupulicate.");
            const double sigma = find_parameter_either(second.second, env, "sigma", "\xCF\x83").as_double() * NM;
            const double epsilon = find_parameter_either(second.second, env, "epsilon", "\xCE\xB5").as_double() * KJ;
            std::vector<std::size_t> former, latter;
            for (const auto& p : params) {
                const std::string& name = p.at("name").as_string();
                const std::size_t i = read_index(p, env, who, topology.size());
                if (name == name_pair.first) former.push_back(i);
                if (name == name_pair.second) latter.push_back(i);
            }
            std::cerr << "        " << name_pair.first << "-" << name_pair.second << " (" << former.size() << "-"
                      << latter.size() << " found)" << std::endl;
            if (former.empty() || latter.empty()) {
                std::cerr << "        \033[33m[warning]\033[m this force field generation will be skipped" << std::endl;
                continue;
            }
            out.push_back(make(n, epsilon, sigma, former, latter, ignore, gp));
            treated.push_back(name_pair);
        }
    }
    return out;
}

std::vector<std::unique_ptr<UniformWeeksChandlerAndersenForceFieldGenerator>>
read_toml_uniform_weeks_chandler_andersen_ff_generators(const Value& ff, std::size_t n, const Topology& topology,
                                                        const std::vector<std::optional<std::string>>& groups,
                                                        bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "parameters", "table", "env"});
    return read_uniform<UniformWeeksChandlerAndersenForceFieldGenerator>(
        ff, n, topology, "read_toml_uniform_weeks_chandler_andersen_ff_generator", "UniformWeeksChandlerAndersen",
        [&](std::size_t sz, double e, double s, const std::vector<std::size_t>& f, const std::vector<std::size_t>& l,
            const index_pairs_type& ig, const group_pairs_type& gp) {
            return std::make_unique<UniformWeeksChandlerAndersenForceFieldGenerator>(sz, e, s, f, l, ig, use_periodic,
                                                                                    gp, groups);
        });
}

// cutoff: a ratio on sigma, default 2.5 (ReadTOMLForceFieldGenerator.cpp:1509-1530)
std::vector<std::unique_ptr<UniformLennardJonesAttractiveForceFieldGenerator>>
read_toml_uniform_lennard_jones_attractive_ff_generators(const Value& ff, std::size_t n, const Topology& topology,
                                                         const std::vector<std::optional<std::string>>& groups,
                                                         bool use_periodic) {
    check_keys_available(ff, {"interaction", "potential", "ignore", "table", "cutoff", "parameters", "env"});
    const double ratio = find_double_or(ff, env_of(ff), "cutoff", 2.5);
    return read_uniform<UniformLennardJonesAttractiveForceFieldGenerator>(
        ff, n, topology, "read_toml_uniform_lennard_jones_attractive_ff_generator", "UniformLennardJonesAttractive",
        [&](std::size_t sz, double e, double s, const std::vector<std::size_t>& f, const std::vector<std::size_t>& l,
            const index_pairs_type& ig, const group_pairs_type& gp) {
            return std::make_unique<UniformLennardJonesAttractiveForceFieldGenerator>(sz, e, s, ratio, f, l, ig,
                                                                                     use_periodic, gp, groups);
        });
}

// The Gaussian+Cosine dihedral form: each parameter holds a Gaussian and a
// Cosine table beside its indices; each half becomes its own generator
// (ReadTOMLInput.cpp:262-292).
static Value split_dihedral_half(const Value& ff, const char* which) {
    Value arr = Value::array();
    for (const auto& elem : ff.at("parameters").as_array()) {
        Value v = Value::table();
        for (const auto& kv : elem.at(which).as_table()) v.insert(kv.first, kv.second);
        v.insert("indices", elem.at("indices"));
        arr.a->push_back(v);
    }
    Value out = Value::table();
    out.insert("parameters", arr);
    return out;
}

// ------------------------------------------------------------ system
static const Value& first_system(const Value& data) {
    const auto& systems = data.at("systems").as_array();
    if (systems.empty()) throw std::runtime_error("[error] [[systems]] is empty");
    return systems[0];
}

static bool periodic_of(const Value& data) {
    const std::string& b = data.at("simulator").at("boundary_type").as_string();
    if (b == "Periodic" || b == "PeriodicCuboid") return true;
    if (b == "Unlimited") return false;
    throw std::runtime_error("[error] unknown boundary_type \"" + b + "\"");
}

// src/input/ReadTOMLInput.cpp:21-875, restricted to the MD-step potentials.
SystemGenerator read_toml_system(const Value& data) {
    const Value& sys = first_system(data);
    const auto& particles = sys.at("particles").as_array();
    const std::size_t n = particles.size();
    Topology topology(n);
    std::vector<std::optional<std::string>> group_vec(n);
    std::vector<double> masses(n);
    for (std::size_t i = 0; i < n; ++i) {
        const Value& p = particles[i];
        masses[i] = find_parameter_either(p, nullptr, "m", "mass").as_double();
        if (const Value* g = p.find("group")) group_vec[i] = g->as_string();
    }
    SystemGenerator gen(masses);
    const bool use_periodic = periodic_of(data);
    if (use_periodic) {
        const Value& shape = sys.at("boundary_shape");
        const auto& lo = shape.at("lower").as_array();
        const auto& up = shape.at("upper").as_array();
        if (lo.size() != 3 || up.size() != 3) throw std::runtime_error("[error] boundary_shape needs 3 components");
        if (lo[0].as_double() != 0.0 || lo[1].as_double() != 0.0 || lo[2].as_double() != 0.0)
            std::cerr << "\033[33m[warning]\033[m Lower bound of periodic boundary box is not (0.0, 0.0, 0.0). "
                         "Only a box whose lower bound is the origin is supported, so the simulation box will be "
                         "moved to satisfy this condition." << std::endl;
        gen.set_pbc((up[0].as_double() - lo[0].as_double()) * NM, (up[1].as_double() - lo[1].as_double()) * NM,
                    (up[2].as_double() - lo[2].as_double()) * NM);
    }
    const Value& attr = sys.at("attributes");
    if (const Value* ens = attr.find("ensemble")) {
        const std::string& type = ens->at("type").as_string();
        if (type != "NVT")
            throw std::runtime_error("[error] ensemble type " + type +
                                     " needs a barostat, which is outside this build's MD-step scope.");
    }
    const auto& ffs = data.at("forcefields").as_array();
    if (ffs.empty()) throw std::runtime_error("[error] [[forcefields]] is empty");
    const Value& ff = ffs[0];
    if (const Value* locals = ff.find("local")) {
        for (const auto& lf : locals->as_array()) {
            const std::string& inter = lf.at("interaction").as_string();
            const std::string& pot = lf.at("potential").as_string();
            if (inter == "BondLength" && pot == "Harmonic")
                gen.add_ff_generator(read_toml_harmonic_bond_ff_generator(lf, topology, use_periodic));
            else if (inter == "BondLength" && pot == "Gaussian")
                gen.add_ff_generator(read_toml_gaussian_bond_ff_generator(lf, topology, use_periodic));
            else if (inter == "BondLength" && pot == "GoContact")
                gen.add_ff_generator(read_toml_go_contact_ff_generator(lf, topology, use_periodic));
            else if (inter == "BondAngle" && pot == "Harmonic")
                gen.add_ff_generator(read_toml_harmonic_angle_ff_generator(lf, topology, use_periodic));
            else if (inter == "BondAngle" && pot == "FlexibleLocalAngle")
                for (const auto& kv : constant::fla_spline_table())
                    gen.add_ff_generator(read_toml_flexible_local_angle_ff_generator(lf, kv.first, kv.second, topology, use_periodic));
            else if (inter == "DihedralAngle" && pot == "Gaussian")
                gen.add_ff_generator(read_toml_gaussian_dihedral_ff_generator(lf, topology, use_periodic));
            else if (inter == "DihedralAngle" && pot == "FlexibleLocalDihedral")
                for (const auto& kv : constant::fld_fourier_table())
                    gen.add_ff_generator(read_toml_flexible_local_dihedral_ff_generator(lf, kv.first, kv.second, topology, use_periodic));
            else if (inter == "BondLength" && pot == "CappedGoContact")
                gen.add_ff_generator(read_toml_capped_go_contact_ff_generator(lf, topology, use_periodic));
            else if (inter == "DihedralAngle" && pot == "Cosine")
                gen.add_ff_generator(read_toml_cosine_dihedral_ff_generator(lf, topology, use_periodic));
            else if (inter == "DihedralAngle" && (pot == "Gaussian+Cosine" || pot == "Cosine+Gaussian")) {
                gen.add_ff_generator(read_toml_gaussian_dihedral_ff_generator(split_dihedral_half(lf, "Gaussian"), topology,
                                                                              use_periodic));
                gen.add_ff_generator(read_toml_cosine_dihedral_ff_generator(split_dihedral_half(lf, "Cosine"), topology,
                                                                            use_periodic));
            } else
                throw std::runtime_error("[error] local potential " + inter + "/" + pot +
                                         " is outside this build's MD-step scope.");
        }
    }
    topology.make_molecule("bond");
    if (const Value* globals = ff.find("global")) {
        for (const auto& gf : globals->as_array()) {
            const std::string& pot = gf.at("potential").as_string();
            if (pot == "ExcludedVolume") {
                gen.add_ff_generator(read_toml_excluded_volume_ff_generator(gf, n, topology, group_vec, use_periodic));
            } else if (pot == "DebyeHuckel") {
                if (!attr.contains("temperature"))
                    throw std::runtime_error("[error] attributes table must contains temperature for DebyeHuckel potential.");
                if (!attr.contains("ionic_strength"))
                    throw std::runtime_error("[error] attributes table must contains ionic_strength for DebyeHuckel potential.");
                gen.add_ff_generator(read_toml_debye_huckel_ff_generator(gf, n, attr.at("ionic_strength").as_double(),
                                                                         attr.at("temperature").as_double(), topology,
                                                                         group_vec, use_periodic));
            } else if (pot == "LennardJonesAttractive") {
                if (gf.contains("table"))
                    for (auto& g : read_toml_uniform_lennard_jones_attractive_ff_generators(gf, n, topology, group_vec,
                                                                                           use_periodic))
                        gen.add_ff_generator(std::move(g));
                else
                    gen.add_ff_generator(read_toml_lennard_jones_attractive_ff_generator(gf, n, topology, group_vec, use_periodic));
            } else if (pot == "WCA") {
                if (gf.contains("table"))
                    for (auto& g : read_toml_uniform_weeks_chandler_andersen_ff_generators(gf, n, topology, group_vec,
                                                                                          use_periodic))
                        gen.add_ff_generator(std::move(g));
                else
                    gen.add_ff_generator(read_toml_weeks_chandler_andersen_ff_generator(gf, n, topology, group_vec, use_periodic));
            } else if (pot == "iSoLFAttractive") {
                gen.add_ff_generator(read_toml_isolf_attractive_ff_generator(gf, n, topology, group_vec, use_periodic));
            } else if (pot == "LennardJonesRepulsive") {
                gen.add_ff_generator(read_toml_lennard_jones_repulsive_ff_generator(gf, n, topology, group_vec, use_periodic));
            } else {
                throw std::runtime_error("[error] global potential " + pot + " is outside this build's MD-step scope.");
            }
        }
    }
    if (ff.contains("external"))
        throw std::runtime_error("[error] external potentials are outside this build's MD-step scope.");
    return gen;
}

// src/input/ReadTOMLInput.cpp:877-906
std::unique_ptr<LangevinIntegratorGenerator> read_toml_integrator_gen(const Value& data) {
    const Value& attr = first_system(data).at("attributes");
    const double temperature = attr.at("temperature").as_double();
    const Value& sim = data.at("simulator");
    const double delta_t = sim.at("delta_t").as_double() * constant::cafetime();
    const int seed = sim.contains("seed") ? static_cast<int>(sim.at("seed").as_int()) : 0;
    const double gamma_t = sim.contains("gamma_t") ? sim.at("gamma_t").as_double() : 0.01;  // 1/cafetime
    return std::make_unique<LangevinIntegratorGenerator>(temperature, gamma_t / constant::cafetime(), delta_t, seed);
}

std::vector<std::array<double, 3>> read_toml_initial_conf(const Value& data) {
    const auto& particles = first_system(data).at("particles").as_array();
    std::vector<std::array<double, 3>> out(particles.size());
    for (std::size_t i = 0; i < particles.size(); ++i) {
        const auto& a = find_parameter_either(particles[i], nullptr, "pos", "position").as_array();
        if (a.size() != 3) throw std::runtime_error("[error] particle position needs 3 components");
        for (int c = 0; c < 3; ++c) out[i][c] = a[c].as_double() * NM;
    }
    return out;
}

std::vector<std::array<double, 3>> read_toml_initial_vel(const Value& data) {
    const auto& particles = first_system(data).at("particles").as_array();
    std::vector<std::array<double, 3>> out;
    if (particles.empty() || !(particles[0].contains("vel") || particles[0].contains("velocity"))) return out;
    out.resize(particles.size());
    for (std::size_t i = 0; i < particles.size(); ++i) {
        const Value& p = particles[i];
        if (!(p.contains("vel") || p.contains("velocity")))
            throw std::runtime_error("[error] All particle need to have velocity, because first particle in system has velocity.");
        const auto& a = find_parameter_either(p, nullptr, "vel", "velocity").as_array();
        for (int c = 0; c < 3; ++c) out[i][c] = a[c].as_double() * NM / constant::cafetime();
    }
    return out;
}

RunSettings read_toml_run_settings(const Value& data) {
    RunSettings r;
    const Value& out = data.at("files").at("output");
    r.output_prefix = out.at("prefix").as_string();
    r.output_format = out.at("format").as_string();
    r.output_path = out.contains("path") ? out.at("path").as_string() : "";
    if (r.output_path.empty()) r.output_path = "./";
    if (r.output_path.back() != '/') r.output_path += '/';
    r.progress_bar = out.contains("progress_bar") ? out.at("progress_bar").as_bool() : true;
    const Value& sim = data.at("simulator");
    r.total_step = static_cast<std::size_t>(sim.at("total_step").as_int());
    r.save_step = static_cast<std::size_t>(sim.at("save_step").as_int());
    if (r.save_step == 0) throw std::runtime_error("[error] save_step must be positive");
    r.delta_t_cafe = sim.at("delta_t").as_double();
    r.energy_minimization = sim.contains("energy_minimization") ? sim.at("energy_minimization").as_bool() : false;
    r.periodic = periodic_of(data);
    const Value& sys = first_system(data);
    r.vel_seed = sys.contains("vel_seed") ? sys.at("vel_seed").as_int() : 0;
    for (const auto& p : sys.at("particles").as_array())
        r.names.push_back(p.contains("name") ? std::optional<std::string>(p.at("name").as_string()) : std::nullopt);
    if (r.output_format != "pdb" && r.output_format != "dcd")
        throw std::runtime_error("[error] output file format `" + r.output_format + "` is not supported.");
    return r;
}

Value load_toml_input(const std::string& path) {
    const auto dot = path.rfind('.');
    const std::string suffix = dot == std::string::npos ? "" : path.substr(dot);
    if (suffix != ".toml")
        throw std::runtime_error("[error] File suffix is `" + suffix + "`. Toml mode needs `.toml` file for toml.");
    Value v = ocmtoml::parse_file(path);
    const auto slash = path.rfind('/');
    expand_include(v, slash == std::string::npos ? "" : path.substr(0, slash));
    return v;
}

}  // namespace ocmhost
