// demhip_solve -- DEMPort.solve (src/executor.jl:174-394) as a native host program over the C ABI
// of libdemhip (include/demhip.h).
//
//   demhip_solve CONFIG.toml [--no-bond] [--no-save] [--precision fp64|mixed] [--out DIR]
//                            [--max-steps N]
//       Reads the configuration (from_toml, executor.jl:182) and the first frame of its particle
//       file (executor.jl:217-220), creates the scene (the sizing and tables of executor.jl:
//       221-339 happen in dem_create), runs initialize! unless --no-bond (executor.jl:346-349),
//       then steps one save interval at a time and writes output.p4p / output.p4c after each
//       (save_single, executor.jl:351-358, 363-386; utils.jl:30-78).
//   demhip_solve --print-config CONFIG.toml
//       The parsed configuration, one "key = value" a line, for checking the loader.
//   demhip_solve --format X [X ...]
//       The number format of the output files, for checking it.
//
// Numbers are written as Julia writes a Float64 (save_single's string interpolation,
// utils.jl:43,67), as this package's Python writer does, so both drivers produce the reference's
// text and the same bytes.
#include <zlib.h>

#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/demhip.h"

namespace {

// ---- configuration: DEMConfig (src/config.jl:3-43) and its TOML loader ----------------------

struct Config {
    dem_config_t c{};
    std::string init_particles = "../Resources/bunny.p4p";
};

// config.jl:4-42 defaults
Config default_config() {
    Config k;
    dem_config_t& c = k.c;
    const double dmin[3] = {-200.0, -200.0, -30.0}, dmax[3] = {200.0, 200.0, 90.0};
    for (int a = 0; a < 3; ++a) { c.domain_min[a] = dmin[a]; c.domain_max[a] = dmax[a]; }
    c.particle_contact_radius_multiplier = 1.1;
    c.neighboring_search_safety_factor = 1.01;
    c.particle_elastic_modulus = 7e10;
    c.particle_poisson_ratio = 0.25;
    c.wall_normal[0] = 0.0; c.wall_normal[1] = 0.0; c.wall_normal[2] = -1.0;
    c.wall_distance = 25.0;
    c.wall_density = 7800.0;
    c.wall_elastic_modulus = 2e11;
    c.wall_poisson_ratio = 0.25;
    c.bond_radius_ratio = 0.5;
    c.bond_elastic_modulus = 28e9;
    c.bond_poisson_ratio = 0.2;
    c.bond_compressive_strength = 3e8;
    c.bond_tensile_strength = 6e7;
    c.bond_shear_strength = 6e7;
    c.pp_coefficient_friction = 0.3;
    c.pp_coefficient_restitution = 0.9;
    c.pp_coefficient_rolling_resistance = 0.01;
    c.pw_coefficient_friction = 0.35;
    c.pw_coefficient_restitution = 0.7;
    c.pw_coefficient_rolling_resistance = 0.01;
    c.max_coordinate_number = 64;
    c.collision_pair_init_capacity_factor = 128;
    c.tolerance = 1e-12;
    c.max_particle_count = 1000000000;
    c.gravity[0] = 0.0; c.gravity[1] = 0.0; c.gravity[2] = -9.81;
    c.global_damping = 0.0;
    c.dt = 2.63e-5;
    c.target_time = 0.5;
    c.saving_interval_time = 0.05;
    return k;
}

enum class Kind { Real, Int, Vec3, Str };
struct Field {
    Kind kind;
    void* (*at)(Config&);
};

#define REAL(f) {#f, {Kind::Real, [](Config& k) -> void* { return &k.c.f; }}}
#define INT(f) {#f, {Kind::Int, [](Config& k) -> void* { return &k.c.f; }}}
#define VEC(f) {#f, {Kind::Vec3, [](Config& k) -> void* { return k.c.f; }}}
// field order of config.jl (and of --print-config)
const std::vector<std::pair<std::string, Field>>& fields() {
    static const std::vector<std::pair<std::string, Field>> f = {
        VEC(domain_min), VEC(domain_max),
        {"init_particles", {Kind::Str, [](Config& k) -> void* { return &k.init_particles; }}},
        REAL(particle_contact_radius_multiplier), REAL(neighboring_search_safety_factor),
        REAL(particle_elastic_modulus), REAL(particle_poisson_ratio),
        VEC(wall_normal), REAL(wall_distance), REAL(wall_density), REAL(wall_elastic_modulus),
        REAL(wall_poisson_ratio),
        REAL(bond_radius_ratio), REAL(bond_elastic_modulus), REAL(bond_poisson_ratio),
        REAL(bond_compressive_strength), REAL(bond_tensile_strength), REAL(bond_shear_strength),
        REAL(pp_coefficient_friction), REAL(pp_coefficient_restitution),
        REAL(pp_coefficient_rolling_resistance),
        REAL(pw_coefficient_friction), REAL(pw_coefficient_restitution),
        REAL(pw_coefficient_rolling_resistance),
        INT(max_coordinate_number), INT(collision_pair_init_capacity_factor),
        REAL(tolerance), INT(max_particle_count),
        VEC(gravity), REAL(global_damping), REAL(dt), REAL(target_time), REAL(saving_interval_time)};
    return f;
}
#undef REAL
#undef INT
#undef VEC

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// A TOML number: integer or float, with optional sign, exponent and '_' digit separators.
double toml_number(const std::string& raw, const std::string& key) {
    std::string s;
    for (char ch : raw)
        if (ch != '_') s += ch;
    if (s == "inf" || s == "+inf") return INFINITY;
    if (s == "-inf") return -INFINITY;
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (s.empty() || *end != '\0') throw std::runtime_error("bad number for " + key + ": " + raw);
    return v;
}

// The subset of TOML the reference's configurations use: one "key = value" a line, values that
// are numbers, basic strings or arrays of numbers, and '#' comments.  Unknown keys are an error.
Config load_toml(const std::string& path) {
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) throw std::runtime_error("cannot open " + path);
    Config k = default_config();
    std::map<std::string, Field> table(fields().begin(), fields().end());
    char buf[4096];
    int lineno = 0;
    while (std::fgets(buf, sizeof buf, fp)) {
        ++lineno;
        std::string line = buf;
        bool in_str = false;
        for (size_t i = 0; i < line.size(); ++i) {  // drop a comment outside strings
            if (line[i] == '"' && (i == 0 || line[i - 1] != '\\')) in_str = !in_str;
            if (line[i] == '#' && !in_str) { line.resize(i); break; }
        }
        line = trim(line);
        if (line.empty()) continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) {
            std::fclose(fp);
            throw std::runtime_error(path + ":" + std::to_string(lineno) + ": expected key = value");
        }
        const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        auto it = table.find(key);
        if (it == table.end()) {
            std::fclose(fp);
            throw std::runtime_error("unknown DEMConfig key '" + key + "'");
        }
        const Field& f = it->second;
        switch (f.kind) {
            case Kind::Str: {
                if (val.size() < 2 || val.front() != '"' || val.back() != '"') {
                    std::fclose(fp);
                    throw std::runtime_error(key + " must be a string");
                }
                std::string s;
                for (size_t i = 1; i + 1 < val.size(); ++i) {
                    if (val[i] == '\\' && i + 2 < val.size()) ++i;
                    s += val[i];
                }
                *static_cast<std::string*>(f.at(k)) = s;
                break;
            }
            case Kind::Vec3: {
                if (val.size() < 2 || val.front() != '[' || val.back() != ']') {
                    std::fclose(fp);
                    throw std::runtime_error(key + " must be an array");
                }
                std::vector<double> xs;
                std::string item;
                const std::string body = val.substr(1, val.size() - 2) + ",";
                for (char ch : body) {
                    if (ch == ',') {
                        item = trim(item);
                        if (!item.empty()) xs.push_back(toml_number(item, key));
                        item.clear();
                    } else {
                        item += ch;
                    }
                }
                if (xs.size() != 3) {
                    std::fclose(fp);
                    throw std::runtime_error(key + " must have 3 components");
                }
                double* d = static_cast<double*>(f.at(k));
                for (int a = 0; a < 3; ++a) d[a] = xs[a];
                break;
            }
            case Kind::Real:
                *static_cast<double*>(f.at(k)) = toml_number(val, key);
                break;
            case Kind::Int:
                *static_cast<int32_t*>(f.at(k)) = (int32_t)toml_number(val, key);
                break;
        }
    }
    std::fclose(fp);
    return k;
}

// ---- numbers and files -------------------------------------------------------------------

// The text Julia writes for a Float64 in string interpolation, as save_single does
// (utils.jl:43,67): the shortest digits that read back as x, in decimal notation when the decimal
// point falls at -4 < pt <= 6 (x = 0.d1d2... * 10^pt and, for integral values, the digits dropped
// to the left of the point are zeros), else d.ddd e N with an unpadded exponent; a whole number
// keeps ".0" (1.0, 100000.0, 1.0e6, 1.0e-5, 0.0001, -0.0, NaN, Inf).
std::string num(double x) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "Inf" : "-Inf";
    if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::scientific);
    const std::string s(buf, r.ptr);  // [-]d[.ddd]e(+|-)XX, shortest digits
    const bool neg = s[0] == '-';
    const size_t e = s.find('e');
    std::string digits;
    for (size_t i = neg ? 1 : 0; i < e; ++i)
        if (s[i] != '.') digits += s[i];
    const int olength = (int)digits.size();
    const int pt = std::atoi(s.c_str() + e + 1) + 1;
    bool plain = pt > -4 && pt <= 6;
    if (plain && pt >= olength) {  // Julia's mod(x + 0.05, 10^(pt - olength)) test
        const double step = std::pow(10.0, pt - olength);
        double m = std::fmod(x + 0.05, step);
        if (m != 0.0 && std::signbit(m) != std::signbit(step)) m += step;
        plain = !(std::fabs(m - 0.05) > 0.05);
    }
    std::string out = neg ? "-" : "";
    if (!plain) {
        out += digits[0];
        out += ".";
        out += olength > 1 ? digits.substr(1) : "0";
        out += "e" + std::to_string(pt - 1);
    } else if (pt <= 0) {
        out += "0." + std::string(-pt, '0') + digits;
    } else if (olength <= pt) {
        out += digits + std::string(pt - olength, '0') + ".0";
    } else {
        out += digits.substr(0, pt) + "." + digits.substr(pt);
    }
    return out;
}

// First frame of a p4p file (plain or gzip), in file order (executor.jl:217-220).
std::vector<dem_grain_in_t> read_p4p_rows(const std::string& path) {
    gzFile fp = gzopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open " + path);
    char buf[4096];
    auto next = [&](const char* what) {
        if (!gzgets(fp, buf, sizeof buf)) {
            gzclose(fp);
            throw std::runtime_error(path + ": missing " + std::string(what));
        }
        return std::string(buf);
    };
    next("TIMESTEP line");
    long long n = 0;
    double t = 0;
    if (std::sscanf(next("frame size").c_str(), "%lf %lld", &t, &n) != 2 || n <= 0) {
        gzclose(fp);
        throw std::runtime_error(path + ": bad frame size line");
    }
    next("column header");
    std::vector<dem_grain_in_t> rows((size_t)n);
    for (long long i = 0; i < n; ++i) {
        double v[10];
        const std::string line = next("particle row");
        const char* p = line.c_str();
        for (int c = 0; c < 10; ++c) {
            char* end = nullptr;
            v[c] = std::strtod(p, &end);
            if (end == p) {
                gzclose(fp);
                throw std::runtime_error(path + ": short particle row " + std::to_string(i + 1));
            }
            p = end;
        }
        dem_grain_in_t& g = rows[(size_t)i];
        g.id = (int32_t)v[0];
        g.gid = (int32_t)v[1];
        g.V = v[2];
        g.m = v[3];
        for (int a = 0; a < 3; ++a) { g.k[a] = v[4 + a]; g.v[a] = v[7 + a]; }
    }
    gzclose(fp);
    return rows;
}

void check(int32_t rc, dem_scene_t* h, const char* what) {
    if (rc != DEM_OK)
        throw std::runtime_error(std::string(what) + ": " + dem_last_error(h) + " (status " +
                                 std::to_string(rc) + ")");
}

// save_single (utils.jl:30-78): a particle frame and a contact frame at time t.
void save_frames(dem_scene_t* h, int64_t n, double t, FILE* p4p, FILE* p4c) {
    std::vector<dem_grain_t> g((size_t)n);
    check(dem_read_grains(h, g.data(), n), h, "dem_read_grains");
    std::fprintf(p4p, "TIMESTEP PARTICLES\n%s %lld\nID  GROUP  VOLUME  MASS  PX  PY  PZ  VX  VY  VZ\n",
                 num(t).c_str(), (long long)n);
    for (const dem_grain_t& x : g)
        std::fprintf(p4p, "%d %d %s %s %s %s %s %s %s %s\n", x.id, x.gid, num(x.V).c_str(),
                     num(x.m).c_str(), num(x.k[0]).c_str(), num(x.k[1]).c_str(), num(x.k[2]).c_str(),
                     num(x.v[0]).c_str(), num(x.v[1]).c_str(), num(x.v[2]).c_str());
    int64_t m = 0;
    check(dem_read_contacts(h, nullptr, nullptr, 0, &m), h, "dem_read_contacts");
    std::vector<dem_contact_t> c((size_t)m);
    std::vector<uint8_t> bonded((size_t)m);
    int64_t m2 = 0;
    if (m > 0) {
        check(dem_read_contacts(h, c.data(), bonded.data(), m, &m2), h, "dem_read_contacts");
        if (m2 != m) throw std::runtime_error("contact count changed between calls");
    }
    std::set<std::pair<int32_t, int32_t>> seen;  // utils.jl:58-60
    for (const dem_contact_t& x : c)
        if (!seen.insert({x.i, x.j}).second) throw std::runtime_error("Duplicate contact pairs!");
    std::fprintf(p4c, "TIMESTEP CONTACTS\n%s %lld\nP1  P2  CX  CY  CZ  FX  FY  FZ  CONTACT_IS_BONDED\n",
                 num(t).c_str(), (long long)m);
    for (size_t k = 0; k < c.size(); ++k) {
        const dem_contact_t& x = c[k];
        std::fprintf(p4c, "%d %d %s %s %s %s %s %s %d\n", x.i, x.j, num(x.k[0]).c_str(),
                     num(x.k[1]).c_str(), num(x.k[2]).c_str(), num(x.F[0]).c_str(),
                     num(x.F[1]).c_str(), num(x.F[2]).c_str(), bonded[k] ? 1 : 0);
    }
}

std::string dirname_of(const std::string& path) {
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

int usage() {
    std::fprintf(stderr,
                 "usage: demhip_solve CONFIG.toml [--no-bond] [--no-save] [--precision fp64|mixed]\n"
                 "                    [--out DIR] [--max-steps N]\n"
                 "       demhip_solve --print-config CONFIG.toml\n"
                 "       demhip_solve --format X [X ...]\n");
    return 2;
}

int print_config(const std::string& path) {
    Config k = load_toml(path);
    for (const auto& [name, f] : fields()) {
        void* p = f.at(k);
        switch (f.kind) {
            case Kind::Str: std::printf("%s = %s\n", name.c_str(), static_cast<std::string*>(p)->c_str()); break;
            case Kind::Real: std::printf("%s = %s\n", name.c_str(), num(*static_cast<double*>(p)).c_str()); break;
            case Kind::Int: std::printf("%s = %d\n", name.c_str(), *static_cast<int32_t*>(p)); break;
            case Kind::Vec3: {
                const double* d = static_cast<double*>(p);
                std::printf("%s = %s %s %s\n", name.c_str(), num(d[0]).c_str(), num(d[1]).c_str(),
                            num(d[2]).c_str());
                break;
            }
        }
    }
    return 0;
}

int solve(const std::string& cfg_path, bool no_bond, bool save, int32_t precision,
          const std::string& out_dir, int64_t max_steps) {
    const auto t0 = std::chrono::steady_clock::now();
    Config k = load_toml(cfg_path);
    const std::vector<dem_grain_in_t> rows =
        read_p4p_rows(dirname_of(cfg_path) + "/" + k.init_particles);
    // executor.jl:313: the configured wall, material 2
    dem_wall_t wall{};
    for (int a = 0; a < 3; ++a) wall.n[a] = k.c.wall_normal[a];
    wall.d = k.c.wall_distance;
    wall.mid = 2;
    dem_options_t opt{precision, 0, 0, 0};
    dem_scene_t* h = nullptr;
    check(dem_create(&k.c, &wall, 1, rows.data(), (int64_t)rows.size(), &opt, &h), nullptr, "dem_create");
    int rc = 0;
    FILE *p4p = nullptr, *p4c = nullptr;
    try {
        dem_derived_t der{};
        check(dem_get_derived(h, &der), h, "dem_get_derived");
        std::fprintf(stderr, "Setting: hash_table_size=%lld cell_size=%g\n",
                     (long long)der.hash_table_size, der.cell_size);
        const int64_t total = max_steps >= 0 ? std::min<int64_t>(der.total_steps, max_steps) : der.total_steps;
        const int64_t per_save = std::max<int64_t>(der.save_per_steps, 1);
        if (!no_bond) check(dem_initialize(h), h, "dem_initialize");
        const int64_t n = (int64_t)rows.size();
        if (save) {
            std::error_code ec;
            std::filesystem::create_directories(out_dir, ec);
            p4p = std::fopen((out_dir + "/output.p4p").c_str(), "w");
            p4c = std::fopen((out_dir + "/output.p4c").c_str(), "w");
            if (!p4p || !p4c) throw std::runtime_error("cannot write output files in " + out_dir);
            save_frames(h, n, 0.0, p4p, p4c);
        }
        int64_t step = 0;
        while (step < total) {
            const auto t1 = std::chrono::steady_clock::now();
            check(dem_step(h, per_save), h, "dem_step");
            step += per_save;
            std::fprintf(stderr, "Solving... step=%lld dt=%.3fs\n", (long long)step, seconds_since(t1));
            if (save) save_frames(h, n, (double)step * k.c.dt, p4p, p4c);
        }
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "demhip_solve: %s\n", ex.what());
        rc = 1;
    }
    if (p4p) std::fclose(p4p);
    if (p4c) std::fclose(p4c);
    dem_destroy(h);
    std::fprintf(stderr, "Total time: %.3f s\n", seconds_since(t0));
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return usage();
    try {
        const std::string first = argv[1];
        if (first == "--format") {
            for (int i = 2; i < argc; ++i) std::printf("%s\n", num(std::strtod(argv[i], nullptr)).c_str());
            return 0;
        }
        if (first == "--print-config") return argc == 3 ? print_config(argv[2]) : usage();
        if (first == "-h" || first == "--help") { usage(); return 0; }
        bool no_bond = false, save = true;
        int32_t precision = DEM_PRECISION_FP64;
        std::string out_dir = ".";
        int64_t max_steps = -1;
        for (int i = 2; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--no-bond") no_bond = true;
            else if (a == "--no-save") save = false;
            else if (a == "--precision" && i + 1 < argc) {
                const std::string p = argv[++i];
                if (p == "fp64") precision = DEM_PRECISION_FP64;
                else if (p == "mixed") precision = DEM_PRECISION_MIXED;
                else return usage();
            } else if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
            else if (a == "--max-steps" && i + 1 < argc) max_steps = std::atoll(argv[++i]);
            else return usage();
        }
        return solve(first, no_bond, save, precision, out_dir, max_steps);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "demhip_solve: %s\n", ex.what());
        return 1;
    }
}
