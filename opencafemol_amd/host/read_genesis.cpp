// read_genesis.cpp -- GENESIS-format readers.
//
// The topology columns are fixed offsets into each line, as the reference
// reads them (a field starts at its offset, runs the given width or to the end
// of the line, leading blanks are skipped and the number ends where the text
// stops being one).  Unit handling follows the reference line by line:
//   [ bonds ]            v0 nm, k x 0.5                  ReadGenesisForceFieldGenerator.cpp:19-22
//   [ angles ] f=21      v0 nm, k negated, sigma nm      :45-51
//   [ angles ] f=22      FLA table of the middle atom's type, k = 1 kcal/mol   :106-124
//   [ dihedrals ] f=41   theta0 degrees -> rad, k negated, sigma rad   :139-147
//   [ dihedrals ] f=52   FLD class rules below, k = 1 kcal/mol         :177-233
//   [ pairs ]            r0 nm, k kJ/mol                 :76-79
//   [ atomtypes ]        radius = rmin / 2, one eps for all (compared as text)   :244-267
#include "read_genesis.hpp"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>

namespace ocmhost {

namespace {

std::string trim(const std::string& s) {
    std::size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

bool has_space(const std::string& s) {
    return std::any_of(s.begin(), s.end(), [](char c) { return std::isspace(static_cast<unsigned char>(c)); });
}

bool is_blank(const std::string& s) { return trim(s).empty(); }

// Utility::erase_space (src/util/Utility.cpp:23-29): every blank removed.
std::string erase_space(std::string s) {
    s.erase(std::remove_if(s.begin(), s.end(), [](char c) { return std::isspace(static_cast<unsigned char>(c)); }),
            s.end());
    return s;
}

// `[ name ]` with optional blanks around and inside the brackets.
bool section_header(const std::string& line, std::string& name) {
    const std::string t = trim(line);
    if (t.size() < 2 || t.front() != '[' || t.back() != ']') return false;
    const std::string inner = trim(t.substr(1, t.size() - 2));
    if (inner.empty()) return false;
    name = inner;
    return true;
}

// `key = value`, each one blank-free token; with several '=' the split that
// leaves the longest key wins.
bool key_value(const std::string& text, std::string& key, std::string& value) {
    for (std::size_t p = text.rfind('='); p != std::string::npos; p = p ? text.rfind('=', p - 1) : std::string::npos) {
        const std::string k = trim(text.substr(0, p)), v = trim(text.substr(p + 1));
        if (!k.empty() && !v.empty() && !has_space(k) && !has_space(v)) {
            key = k;
            value = v;
            return true;
        }
        if (p == 0) break;
    }
    return false;
}

// `section` names a topology section, or is "gro" for the coordinate file
std::string where(const std::string& line, const char* section) {
    if (std::string(section) == "gro") return "the line \"" + line + "\" of the .gro file";
    return "the line \"" + line + "\" of `[ " + section + " ]`";
}

std::string column(const std::string& line, std::size_t pos, std::size_t len, const char* section) {
    if (pos > line.size())
        throw std::runtime_error("[error] " + where(line, section) + " ends before column " + std::to_string(pos + 1) +
                                 ".");
    return line.substr(pos, len);
}

int col_int(const std::string& line, std::size_t pos, std::size_t len, const char* section) {
    const std::string f = column(line, pos, len, section);
    try {
        return std::stoi(f);
    } catch (const std::logic_error&) {
        throw std::runtime_error("[error] " + where(line, section) + " has no integer at columns " +
                                 std::to_string(pos + 1) + "-" + std::to_string(pos + len) + ".");
    }
}

double col_double(const std::string& line, std::size_t pos, std::size_t len, const char* section) {
    const std::string f = column(line, pos, len, section);
    try {
        return std::stod(f);
    } catch (const std::logic_error&) {
        throw std::runtime_error("[error] " + where(line, section) + " has no number at columns " +
                                 std::to_string(pos + 1) + "-" + std::to_string(pos + len) + ".");
    }
}

// 1-based atom number in a 10-column field -> 0-based index below n
std::size_t col_atom(const std::string& line, std::size_t pos, std::size_t n, const char* section) {
    const int v = col_int(line, pos, 10, section);
    if (v < 1 || static_cast<std::size_t>(v) > n)
        throw std::runtime_error("[error] " + where(line, section) + " names atom " + std::to_string(v) +
                                 ", outside 1.." + std::to_string(n) + ".");
    return static_cast<std::size_t>(v - 1);
}

std::vector<std::string> atom_types(const std::vector<std::string>& atoms) {
    std::vector<std::string> out;
    out.reserve(atoms.size());
    for (const auto& a : atoms) out.push_back(erase_space(column(a, 10, 5, "atoms")));
    return out;
}

const std::string& entry(const std::map<std::string, std::string>& table, const std::string& key,
                         const std::string& section) {
    const auto it = table.find(key);
    if (it == table.end())
        throw std::runtime_error("[error] `" + key + "` is missing in the [" + section + "] section.");
    return it->second;
}

const std::map<std::string, std::string>& inp_section(const InpData& inp, const std::string& name) {
    const auto it = inp.find(name);
    if (it == inp.end()) throw std::runtime_error("[error] There is no [" + name + "] section in the inp file.");
    return it->second;
}

int to_int(const std::string& s, const std::string& what) {
    try {
        return std::stoi(s);
    } catch (const std::logic_error&) {
        throw std::runtime_error("[error] `" + what + " = " + s + "` is not an integer.");
    }
}

double to_double(const std::string& s, const std::string& what) {
    try {
        return std::stod(s);
    } catch (const std::logic_error&) {
        throw std::runtime_error("[error] `" + what + " = " + s + "` is not a number.");
    }
}

}  // namespace

// ---------------------------------------------------------------- files
// ReadGenesisInput.cpp:18-70: `[NAME]` sections of `key = value` lines; `#`
// starts a comment; lines before the first section are ignored.
InpData read_inp_file(const std::string& inp_file_name) {
    std::cerr << "reading " << inp_file_name << "..." << std::endl;
    std::ifstream ifs(inp_file_name);
    if (!ifs.good()) throw std::runtime_error("[error] file open error ->" + inp_file_name);
    InpData out;
    std::string line, current, name, key, value;
    while (std::getline(ifs, line)) {
        if (section_header(line, name)) {
            current = name;
            out[current];
            continue;
        }
        if (current.empty() || is_blank(line)) continue;
        if (key_value(line.substr(0, line.find('#')), key, value)) out[current].emplace(key, value);
    }
    return out;
}

// ReadGenesisInput.cpp:72-115.  A line holding a ';' is cut there before any
// other test, so an `#include` with a trailing comment stays a plain line.
std::vector<std::string> preprocess_top_file(const std::string& top_file_name, const std::string& file_path) {
    std::ifstream ifs(file_path + top_file_name);
    if (!ifs.good()) throw std::runtime_error("[error] file open error ->" + top_file_name);
    static const std::string kInclude = "#include \"";
    std::vector<std::string> out;
    std::string line;
    while (std::getline(ifs, line)) {
        const std::string t = trim(line);
        if (t.empty() || t.front() == ';') continue;
        if (const auto semi = line.find(';'); semi != std::string::npos) {
            out.push_back(line.substr(0, semi));
            continue;
        }
        if (line.compare(0, kInclude.size(), kInclude) == 0) {
            const std::string rest = trim(line.substr(kInclude.size()));   // leading blank cannot occur: `"` precedes
            if (!rest.empty() && rest.back() == '"' && !has_space(rest.substr(0, rest.size() - 1)) &&
                line.substr(kInclude.size()).find_first_not_of(" \t\r\n\v\f") == 0) {
                const std::string inc = rest.substr(0, rest.size() - 1);
                std::cerr << "    reading include file " << file_path << inc << "..." << std::endl;
                for (auto& l : preprocess_top_file(inc, file_path)) out.push_back(std::move(l));
                continue;
            }
        }
        out.push_back(line);
    }
    return out;
}

// ReadGenesisInput.cpp:117-169: content lines belong to the latest `[ name ]`;
// a content line before the first header is an error.
TopData read_top_file(const std::string& top_file_name, const std::string& file_path) {
    std::cerr << "reading " << file_path << top_file_name << "..." << std::endl;
    TopData out;
    std::vector<std::string>* current = nullptr;
    std::string name;
    for (auto& line : preprocess_top_file(top_file_name, file_path)) {
        if (section_header(line, name)) {
            current = &out[name];
            continue;
        }
        if (!current)
            throw std::runtime_error("[error] the line \"" + line + "\" does not belong to any section in " +
                                     top_file_name + " file.");
        current->push_back(line);
    }
    return out;
}

// ReadGenesisInput.cpp:192-215: title line, atom count, then x/y/z in nm at
// columns 21-29, 30-38, 39-47.
std::vector<std::array<double, 3>> read_genesis_initial_conf(const std::string& gro_file_name,
                                                             const std::string& file_path) {
    std::ifstream ifs(file_path + gro_file_name);
    if (!ifs.good()) throw std::runtime_error("[error] file open error ->" + gro_file_name);
    std::string line;
    std::getline(ifs, line);
    if (!std::getline(ifs, line)) throw std::runtime_error("[error] " + gro_file_name + " has no atom count line.");
    const int n = to_int(line, "atom count of " + gro_file_name);
    if (n < 0) throw std::runtime_error("[error] negative atom count in " + gro_file_name + ".");
    std::vector<std::array<double, 3>> out(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) {
        if (!std::getline(ifs, line))
            throw std::runtime_error("[error] " + gro_file_name + " ends after " + std::to_string(i) + " of " +
                                     std::to_string(n) + " atoms.");
        for (int c = 0; c < 3; ++c) out[i][c] = col_double(line, 20 + 9 * c, 9, "gro");
    }
    return out;
}

// ReadGenesisInput.cpp:171-190
std::unique_ptr<LangevinIntegratorGenerator> read_genesis_integrator_gen(
    const std::map<std::string, std::string>& dynamics, const std::map<std::string, std::string>& ensemble) {
    const double dt = to_double(entry(dynamics, "timestep", "DYNAMICS"), "timestep");   // ps
    const auto s = dynamics.find("iseed");
    const int seed = s == dynamics.end() ? 0 : to_int(s->second, "iseed");
    const double temperature = to_double(entry(ensemble, "temperature", "ENSEMBLE"), "temperature");
    const double gamma = to_double(entry(ensemble, "gamma_t", "ENSEMBLE"), "gamma_t");   // 1/ps
    return std::make_unique<LangevinIntegratorGenerator>(temperature, gamma, dt, seed);
}

// ----------------------------------------------------------- local terms
// ReadGenesisForceFieldGenerator.cpp:9-32.  v0 is the first number from
// column 26 on and k the first from column 44 on.
std::unique_ptr<HarmonicBondForceFieldGenerator> read_genesis_harmonic_bond_ff_generator(
    const std::vector<std::string>& bonds, Topology& topology, bool use_periodic) {
    index_pairs_type idx;
    std::vector<double> v0s, ks;
    for (const auto& l : bonds) {
        idx.emplace_back(col_atom(l, 0, topology.size(), "bonds"), col_atom(l, 10, topology.size(), "bonds"));
        v0s.push_back(col_double(l, 25, 42, "bonds"));
        ks.push_back(col_double(l, 43, 60, "bonds") * 0.5);
    }
    topology.add_edges(idx, "bond");
    std::cerr << "    BondLength    : Harmonic (" << idx.size() << " found)" << std::endl;
    return std::make_unique<HarmonicBondForceFieldGenerator>(idx, v0s, ks, use_periodic);
}

// ReadGenesisForceFieldGenerator.cpp:34-63: function 21 in columns 31-35.
std::unique_ptr<GaussianBondForceFieldGenerator> read_genesis_gaussian_bond_ff_generator(
    const std::vector<std::string>& angles, std::size_t n, bool use_periodic) {
    index_pairs_type idx;
    std::vector<double> v0s, ks, sigmas;
    for (const auto& l : angles) {
        if (col_int(l, 30, 5, "angles") != 21) continue;
        idx.emplace_back(col_atom(l, 0, n, "angles"), col_atom(l, 20, n, "angles"));
        v0s.push_back(col_double(l, 35, 15, "angles"));
        ks.push_back(-col_double(l, 50, 15, "angles"));
        sigmas.push_back(col_double(l, 65, 15, "angles"));
    }
    std::cerr << "    BondLength    : Gaussian (" << idx.size() << " found)" << std::endl;
    return std::make_unique<GaussianBondForceFieldGenerator>(idx, ks, v0s, sigmas, use_periodic);
}

// ReadGenesisForceFieldGenerator.cpp:65-89
std::unique_ptr<GoContactForceFieldGenerator> read_genesis_go_contact_ff_generator(
    const std::vector<std::string>& pairs, Topology& topology, bool use_periodic) {
    index_pairs_type idx;
    std::vector<double> ks, r0s;
    for (const auto& l : pairs) {
        idx.emplace_back(col_atom(l, 0, topology.size(), "pairs"), col_atom(l, 10, topology.size(), "pairs"));
        r0s.push_back(col_double(l, 30, 15, "pairs"));
        ks.push_back(col_double(l, 45, 15, "pairs"));
    }
    topology.add_edges(idx, "contact");
    std::cerr << "    BondLength    : GoContact (" << idx.size() << " found)" << std::endl;
    return std::make_unique<GoContactForceFieldGenerator>(idx, ks, r0s, use_periodic);
}

// ReadGenesisForceFieldGenerator.cpp:91-125: function 22 whose middle atom has
// type `aa_type` (columns 11-15 of [ atoms ], blanks removed).
std::unique_ptr<FlexibleLocalAngleForceFieldGenerator> read_genesis_flexible_local_angle_ff_generator(
    const std::vector<std::string>& angles, const std::vector<std::string>& atoms, const std::string& aa_type,
    bool use_periodic) {
    const auto types = atom_types(atoms);
    std::vector<FlexibleLocalAngleForceFieldGenerator::indices_type> idx;
    for (const auto& l : angles) {
        if (col_int(l, 30, 5, "angles") != 22) continue;
        const std::size_t j = col_atom(l, 10, types.size(), "angles");
        if (types[j] != aa_type) continue;
        idx.push_back({col_atom(l, 0, types.size(), "angles"), j, col_atom(l, 20, types.size(), "angles")});
    }
    std::cerr << "    BondAngle     : FlexibleLocalAngle - " << aa_type << " (" << idx.size() << " found)"
              << std::endl;
    return std::make_unique<FlexibleLocalAngleForceFieldGenerator>(
        idx, std::vector<double>(idx.size(), constant::kj_per_kcal), constant::fla_spline_table().at(aa_type), aa_type,
        use_periodic);
}

// ReadGenesisForceFieldGenerator.cpp:127-159: function 41 in columns 41-45.
std::unique_ptr<GaussianDihedralForceFieldGenerator> read_genesis_gaussian_dihedral_ff_generator(
    const std::vector<std::string>& dihedrals, std::size_t n, bool use_periodic) {
    std::vector<GaussianDihedralForceFieldGenerator::indices_type> idx;
    std::vector<double> ks, theta0s, sigmas;
    for (const auto& l : dihedrals) {
        if (col_int(l, 40, 5, "dihedrals") != 41) continue;
        idx.push_back({col_atom(l, 0, n, "dihedrals"), col_atom(l, 10, n, "dihedrals"), col_atom(l, 20, n, "dihedrals"),
                       col_atom(l, 30, n, "dihedrals")});
        theta0s.push_back(col_double(l, 45, 15, "dihedrals") / 180.0 * constant::pi);
        ks.push_back(-col_double(l, 60, 15, "dihedrals"));
        sigmas.push_back(col_double(l, 75, 15, "dihedrals"));
    }
    std::cerr << "    DihedralAngle : Gaussian (" << idx.size() << " found)" << std::endl;
    return std::make_unique<GaussianDihedralForceFieldGenerator>(idx, ks, theta0s, sigmas, use_periodic);
}

// ReadGenesisForceFieldGenerator.cpp:161-234: function 52, selected by the
// types t_j, t_k of the two middle atoms:
//   ("R1", "GLY")   t_k is GLY
//   ("R2", "PRO")   t_j is not GLY and t_k is PRO
//   ("GLY", "PRO")  t_j is GLY and t_k is PRO
//   (X, "R3")       t_j is X, whatever t_k is
// A term can therefore match two classes (e.g. GLY-PRO and GLY-R3); each
// matching class adds it, as the reference does.
std::unique_ptr<FlexibleLocalDihedralForceFieldGenerator> read_genesis_flexible_local_dihedral_ff_generator(
    const std::vector<std::string>& dihedrals, const std::vector<std::string>& atoms,
    const std::pair<std::string, std::string>& aa_type_pair, bool use_periodic) {
    const auto types = atom_types(atoms);
    const std::size_t n = types.size();
    std::vector<FlexibleLocalDihedralForceFieldGenerator::indices_type> idx;
    for (const auto& l : dihedrals) {
        if (col_int(l, 40, 5, "dihedrals") != 52) continue;
        const std::size_t j = col_atom(l, 10, n, "dihedrals"), k = col_atom(l, 20, n, "dihedrals");
        bool take;
        if (aa_type_pair.first == "R1") take = types[k] == "GLY";
        else if (aa_type_pair.first == "R2") take = types[j] != "GLY" && types[k] == "PRO";
        else if (aa_type_pair.first == "GLY" && aa_type_pair.second == "PRO") take = types[j] == "GLY" && types[k] == "PRO";
        else take = types[j] == aa_type_pair.first;
        if (take) idx.push_back({col_atom(l, 0, n, "dihedrals"), j, k, col_atom(l, 30, n, "dihedrals")});
    }
    const std::string name = aa_type_pair.first + "-" + aa_type_pair.second;
    std::cerr << "    DihedralAngle : FlexibleLocalDihedral - " << name << " (" << idx.size() << " found)" << std::endl;
    return std::make_unique<FlexibleLocalDihedralForceFieldGenerator>(
        idx, std::vector<double>(idx.size(), constant::kj_per_kcal), constant::fld_fourier_table().at(aa_type_pair),
        name, use_periodic);
}

// ---------------------------------------------------------- global terms
// ReadGenesisForceFieldGenerator.cpp:236-292: radius = rmin/2 (columns 31-40)
// of the atom's type; eps (columns 42-50) must be the same text on every
// atomtypes line; beads within 3 bonds or 1 contact are excluded; cutoff
// ratio 2.
std::unique_ptr<ExcludedVolumeForceFieldGenerator> read_genesis_exv_ff_generator(
    const std::vector<std::string>& atomtypes, const std::vector<std::string>& atoms, const Topology& topology,
    bool use_periodic) {
    std::map<std::string, double> rmin_half;
    std::string eps_text;
    for (const auto& l : atomtypes) {
        const std::string name = erase_space(column(l, 0, 5, "atomtypes"));
        if (!rmin_half.emplace(name, col_double(l, 30, 10, "atomtypes") * 0.5).second)
            throw std::runtime_error("[error] `" + name + "` key was duplicated in `[ atomtypes ]` section.");
        const std::string e = erase_space(column(l, 41, 9, "atomtypes"));
        if (eps_text.empty()) eps_text = e;
        else if (eps_text != e)
            throw std::runtime_error("[error] ExcludedVolumeForceField only support uniform eps for all particle "
                                     "case. Some different value were defined in `[ atomtypes ]` section.");
    }
    double eps = 0.0;
    try {
        eps = std::stod(eps_text);
    } catch (const std::logic_error&) {
        throw std::runtime_error("[error] `[ atomtypes ]` gives no eps value.");
    }
    std::vector<std::optional<double>> radii;
    for (const auto& t : atom_types(atoms)) {
        const auto it = rmin_half.find(t);
        if (it == rmin_half.end())
            throw std::runtime_error("[error] atom type `" + t + "` of `[ atoms ]` is not in `[ atomtypes ]`.");
        radii.emplace_back(it->second);
    }
    std::cerr << "    Global        : ExcludedVolume (" << radii.size() << " found)" << std::endl;
    index_pairs_type ignore = topology.ignore_list_within_edge(3, "bond");
    const index_pairs_type contact = topology.ignore_list_within_edge(1, "contact");
    ignore.insert(ignore.end(), contact.begin(), contact.end());
    std::sort(ignore.begin(), ignore.end());
    ignore.erase(std::unique(ignore.begin(), ignore.end()), ignore.end());
    return std::make_unique<ExcludedVolumeForceFieldGenerator>(eps, 2.0, radii, ignore, use_periodic);
}

// ------------------------------------------------------------- assembly
// ReadGenesisInput.cpp:217-436 (system, outputs, dynamics) and :489-520
// (paths: .top, .gro and any includes are relative to the .inp directory).
GenesisInput read_genesis_input(const std::string& inp_file_name) {
    const auto slash = inp_file_name.rfind('/');
    const std::string file_path = slash == std::string::npos ? "" : inp_file_name.substr(0, slash + 1);
    const InpData inp = read_inp_file(inp_file_name);
    const auto& input = inp_section(inp, "INPUT");
    const std::string& top_name = entry(input, "grotopfile", "INPUT");
    const std::string& gro_name = entry(input, "grocrdfile", "INPUT");

    const TopData top = read_top_file(top_name, file_path);
    const auto atoms_it = top.find("atoms");
    if (atoms_it == top.end())
        throw std::runtime_error("[error] There is no atoms section. Top file must contain one atoms section.");
    const auto& atoms = atoms_it->second;

    GenesisInput out;
    std::vector<double> masses;
    for (const auto& l : atoms) {
        masses.push_back(col_double(l, 49, 9, "atoms"));   // amu
        out.res_names.push_back(column(l, 27, 3, "atoms"));
        out.atom_names.push_back(column(l, 31, 4, "atoms"));
    }
    const std::size_t n = masses.size();
    out.system = SystemGenerator(masses);

    const auto& boundary = inp_section(inp, "BOUNDARY");
    out.periodic = entry(boundary, "type", "BOUNDARY") == "PBC";
    if (out.periodic) {
        // The reference passes no box in this mode; GENESIS itself takes the
        // cell edges (Angstrom) from these keys, which are read here.
        std::array<double, 3> box{};
        const char* keys[3] = {"box_size_x", "box_size_y", "box_size_z"};
        for (int c = 0; c < 3; ++c) {
            const auto it = boundary.find(keys[c]);
            if (it == boundary.end())
                throw std::runtime_error(std::string("[error] type = PBC needs `") + keys[c] +
                                         "` (Angstrom) in the [BOUNDARY] section.");
            box[c] = to_double(it->second, keys[c]) * constant::nm_per_angstrom;
        }
        out.system.set_pbc(box[0], box[1], box[2]);
        std::cerr << "    boundary type is periodic boundary condition" << std::endl;
    }

    std::cerr << "reading forcefield tables..." << std::endl;
    Topology topology(n);
    auto skip = [] { std::cerr << "        -> skip this forcefield generation" << std::endl; };
    if (const auto it = top.find("bonds"); it != top.end()) {
        auto g = read_genesis_harmonic_bond_ff_generator(it->second, topology, out.periodic);
        if (!g->indices().empty()) out.system.add_ff_generator(std::move(g));
        else skip();
    }
    if (const auto it = top.find("angles"); it != top.end()) {
        auto g = read_genesis_gaussian_bond_ff_generator(it->second, n, out.periodic);
        if (g->size()) out.system.add_ff_generator(std::move(g));
        else skip();
        for (const auto& kv : constant::fla_spline_table()) {
            auto f = read_genesis_flexible_local_angle_ff_generator(it->second, atoms, kv.first, out.periodic);
            if (f->size()) out.system.add_ff_generator(std::move(f));
            else skip();
        }
    }
    if (const auto it = top.find("dihedrals"); it != top.end()) {
        auto g = read_genesis_gaussian_dihedral_ff_generator(it->second, n, out.periodic);
        if (g->size()) out.system.add_ff_generator(std::move(g));
        else skip();
        for (const auto& kv : constant::fld_fourier_table()) {
            auto f = read_genesis_flexible_local_dihedral_ff_generator(it->second, atoms, kv.first, out.periodic);
            if (f->size()) out.system.add_ff_generator(std::move(f));
            else skip();
        }
    }
    if (const auto it = top.find("pairs"); it != top.end()) {
        auto g = read_genesis_go_contact_ff_generator(it->second, topology, out.periodic);
        if (g->size()) out.system.add_ff_generator(std::move(g));
        else skip();
    }
    const auto types_it = top.find("atomtypes");
    if (types_it == top.end())
        throw std::runtime_error(
            "[error] There is no `[ atomtypes ]` section. Genesis input mode needs this section.");
    const auto mol_it = top.find("moleculetype");
    if (mol_it == top.end() || mol_it->second.empty())
        throw std::runtime_error(
            "[error] There is no `[ moleculetype ]` section. Genesis input mode needs this section.");
    // nrexcl (columns 14-23) must agree across molecules; the exclusions
    // themselves are fixed at 3 bonds and 1 contact whatever its value.
    const int nrexcl = col_int(mol_it->second.front(), 13, 10, "moleculetype");
    for (const auto& l : mol_it->second)
        if (col_int(l, 13, 10, "moleculetype") != nrexcl)
            throw std::runtime_error("[error] all nrexcl need to be same in `[moleculetype]` section.");
    out.system.add_ff_generator(read_genesis_exv_ff_generator(types_it->second, atoms, topology, out.periodic));

    out.positions = read_genesis_initial_conf(gro_name, file_path);
    if (out.positions.size() != n)
        throw std::runtime_error("[error] " + gro_name + " holds " + std::to_string(out.positions.size()) +
                                 " atoms but the topology has " + std::to_string(n) + ".");

    const auto& output = inp_section(inp, "OUTPUT");
    const std::string& dcd = entry(output, "dcdfile", "OUTPUT");
    out.file_prefix = dcd.substr(0, std::min(dcd.rfind('.'), dcd.size()));
    if (const auto it = output.find("pdbfile"); it != output.end()) out.pdbfile = it->second;

    const auto& dynamics = inp_section(inp, "DYNAMICS");
    const int nsteps = to_int(entry(dynamics, "nsteps", "DYNAMICS"), "nsteps");
    const int period = to_int(entry(dynamics, "crdout_period", "DYNAMICS"), "crdout_period");
    if (nsteps < 0 || period <= 0)
        throw std::runtime_error("[error] nsteps must be >= 0 and crdout_period > 0 in the [DYNAMICS] section.");
    out.nsteps = static_cast<std::size_t>(nsteps);
    out.crdout_period = static_cast<std::size_t>(period);
    out.timestep = to_double(entry(dynamics, "timestep", "DYNAMICS"), "timestep");
    out.integrator = read_genesis_integrator_gen(dynamics, inp_section(inp, "ENSEMBLE"));
    if (const auto it = inp.find("PLATFORM"); it != inp.end()) out.platform = entry(it->second, "name", "PLATFORM");
    return out;
}

void write_genesis_initial_pdb(const GenesisInput& in) {
    std::cerr << "    output initial state file : " << in.pdbfile << std::endl;
    std::ofstream ofs(in.pdbfile, std::ios::trunc);
    if (!ofs.good()) throw std::runtime_error("[error] file open error ->" + in.pdbfile);
    ofs << "MODEL     0\n" << std::fixed << std::setprecision(3);
    for (std::size_t i = 0; i < in.positions.size(); ++i) {
        ofs << "ATOM  " << std::setw(5) << i + 1 << " " << in.atom_names[i] << " " << in.res_names[i] << "  "
            << "   1    ";
        for (int c = 0; c < 3; ++c) ofs << std::setw(8) << in.positions[i][c] * constant::angstrom_per_nm;
        ofs << "  1.00  0.00\n";
    }
    ofs << "ENDMDL" << std::endl;
}

}  // namespace ocmhost
