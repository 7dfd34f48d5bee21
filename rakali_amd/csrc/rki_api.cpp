// rki_api.cpp -- librakali.so: the reference's handle API (rki_*) for the
// structured barotropic path, over the MI355X step in librakali_amd.so.
//
// What lives here (host code only):
//   * the strict namelist reader for the keys of config_t
//     (rki_config.F90:37-470, schema rki_config.F90:1234-1810, syntax
//     rki_nml_schema.F90:1160-1600);
//   * the handle, its magic and the two-phase lifecycle
//     (rki_handle.F90:38-64, rki_api.F90:157-345);
//   * Fortran-ordered host arrays with ghost cells and the
//     (on_device, host_is_current) state machine of rki_api.F90:54-137;
//   * the initial conditions of init_from_config (rki_init.F90:26-212) and
//     fill_bathymetry_ghosts_array (rki_bathymetry.F90:165-196);
//   * the ghost rings the reference leaves in h, hu, hv and b after a step
//     (apply_boundary_conditions on U* at t + dt, rki_solver.F90:790-795),
//     recomputed on the host when the state is copied back.
// The step itself runs on the GPU through rkg_step / rkg_advance.
#include "nc_classic.hpp"
#include "rakali.h"
#include "rakali_amd.h"
#include "crmath.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

namespace {

constexpr uint32_t HANDLE_MAGIC = 0x72616B53u;   // 'rakS', rki_handle.F90:38
constexpr uint32_t HANDLE_DEAD = 0xDEADBEEFu;
constexpr int MAX_TC = 10;                       // MAX_TIDAL_CONSTITUENTS
constexpr double GRAVITY = 9.80665;              // rki_constants.F90

thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// ------------------------------------------------------------------------
// Configuration (the subset of config_t the structured 2-D path reads,
// plus every other schema key so that a reference namelist parses).
// ------------------------------------------------------------------------
struct Config {
    std::string sim_type = "coastal", grid_type = "structured";
    int nx = 200, ny = 1, nghost = 3;
    double dx = 1.0, dy = 1.0;
    double t_end = 1.0, cfl = 0.45, dt_max = 1.0e10, dt_fixed = 0.0;
    int cfl_interval = 1;
    std::string time_scheme = "rk2", time_unit = "s";
    double manning_n = 0.0, dry_threshold = 0.001, wind_stress_x = 0.0, wind_stress_y = 0.0;
    double coriolis_f = 0.0, bottom_drag_r = 0.0;
    bool use_hllc = false, use_linear_bottom_drag = false, wind_surface_layer = false;
    std::string bathymetry_file;
    std::string bc[4] = {"wall", "wall", "wall", "wall"};
    double tidal_amplitude = 0.0, tidal_period = 44712.0;
    int n_tidal_constituents = 0;
    double tidal_amp[MAX_TC] = {}, tidal_phase[MAX_TC] = {}, tidal_omega[MAX_TC] = {};
    double inflow_velocity = 0.0, inflow_salinity = -1.0, inflow_temperature = -999.0;
    double discharge_flux = 0.0, clamped_elevation = 0.0;
    int sponge_width = 0;
    double sponge_strength = 0.0, sponge_h_ref = 0.0;
    std::string nesting_file;
    bool use_nonhydrostatic = false, use_multilayer = false;
    int nz_layers = 2;
    bool use_ppm_tracer = true;
    std::string init_condition = "dam_break";
    double eta_initial = 0.0, h0 = 1.0, hump_amplitude = 0.05, hump_sigma = 5.0;
    double hump_x_offset = 0.0, hump_y_offset = 0.0, init_u_shear = 0.0;
    double status_interval = 0.0;
    // keys that are parsed and kept but have no effect on this path
    std::vector<std::string> inert_set;
};

enum class Kind { REAL, INT, LOGICAL, STRING, ENUM, REAL_ARRAY, IGNORED };

struct KeyDef {
    const char* group;
    const char* name;
    Kind kind;
    size_t off;                 // offset into Config (unused for IGNORED)
    std::vector<std::string> allowed;
    bool has_min = false, has_max = false;
    double vmin = 0.0, vmax = 0.0;
    Kind ignored_as = Kind::REAL;   // value syntax checked for IGNORED keys
};

#define OFF(field) offsetof(Config, field)

std::vector<KeyDef> build_schema() {
    std::vector<KeyDef> k;
    auto add = [&](const char* gname, const char* name, Kind kind, size_t off) {
        KeyDef d{gname, name, kind, off, {}};
        k.push_back(d);
        return &k.back();
    };
    auto ign = [&](const char* gname, const char* name, Kind as) {
        KeyDef d{gname, name, Kind::IGNORED, 0, {}};
        d.ignored_as = as;
        k.push_back(d);
    };
    auto en = [&](const char* gname, const char* name, size_t off, std::vector<std::string> allowed) {
        KeyDef d{gname, name, Kind::ENUM, off, std::move(allowed)};
        k.push_back(d);
    };
    auto ign_enum = [&](const char* gname, const char* name, std::vector<std::string> allowed) {
        KeyDef d{gname, name, Kind::IGNORED, 0, std::move(allowed)};
        d.ignored_as = Kind::ENUM;
        k.push_back(d);
    };
    const std::vector<std::string> bcs = {"wall", "open", "tidal", "nested", "inflow",
                                          "discharge", "clamped", "sponge", "chapman"};
    // &sim_nml, &grid_nml, &time_nml (rki_config.F90:1256-1331)
    en("sim", "sim_type", OFF(sim_type), {"coastal"});
    en("grid", "grid_type", OFF(grid_type), {"structured", "unstructured"});
    { auto* d = add("grid", "nx", Kind::INT, OFF(nx)); d->has_min = true; d->vmin = 1; }
    { auto* d = add("grid", "ny", Kind::INT, OFF(ny)); d->has_min = true; d->vmin = 1; }
    add("grid", "dx", Kind::REAL, OFF(dx));
    add("grid", "dy", Kind::REAL, OFF(dy));
    { auto* d = add("grid", "nghost", Kind::INT, OFF(nghost)); d->has_min = true; d->vmin = 1; }
    add("time", "t_end", Kind::REAL, OFF(t_end));
    add("time", "cfl", Kind::REAL, OFF(cfl));
    add("time", "dt_max", Kind::REAL, OFF(dt_max));
    add("time", "dt_fixed", Kind::REAL, OFF(dt_fixed));
    { auto* d = add("time", "cfl_interval", Kind::INT, OFF(cfl_interval)); d->has_min = true; d->vmin = 1; }
    en("time", "time_scheme", OFF(time_scheme), {"euler", "rk2", "rk3", "rk43", "semi_implicit"});
    add("time", "time_unit", Kind::STRING, OFF(time_unit));
    // &si_nml (unstructured semi-implicit only)
    ign_enum("si", "precond", {"none", "jacobi", "mg"});
    ign("si", "rtol", Kind::REAL);
    ign("si", "wetdry", Kind::LOGICAL);
    ign("si", "wetdry_nsub", Kind::INT);
    ign_enum("si", "wetdry_bsub_fill", {"flat", "linear"});
    ign("si", "newton_max", Kind::INT);
    ign("si", "newton_tol", Kind::REAL);
    // &metadata_nml, &logging_nml
    ign("metadata", "reference_time", Kind::STRING);
    ign("metadata", "institution", Kind::STRING);
    ign_enum("logging", "log_level", {"debug", "verbose", "info", "performance", "warning", "error"});
    add("logging", "status_interval", Kind::REAL, OFF(status_interval));
    // &vcoord_nml (multilayer only)
    ign_enum("vcoord", "vcoord_type", {"sigma", "zsigma", "zstar", "zstar_full", "zstar_sigma", "z",
                                       "eulerian_z", "isopycnal", "lagrangian", "gprime", "z_fixed",
                                       "rho", "hycom"});
    ign_enum("vcoord", "remap_method", {"pcm", "plm", "ppm", "ppm_h4", "pqm"});
    for (const char* n : {"vcoord_depth_transition", "vcoord_blend_width", "zstar_h_surf_target",
                          "zstar_h_min", "rho_ref_pressure", "rho_target_light", "rho_target_dense"})
        ign("vcoord", n, Kind::REAL);
    ign_enum("vcoord", "zstar_stretching", {"log", "uniform"});
    ign("vcoord", "zstar_n_surf", Kind::INT);
    // &physics_nml (rki_config.F90:1476-1505)
    add("physics", "manning_n", Kind::REAL, OFF(manning_n));
    add("physics", "dry_threshold", Kind::REAL, OFF(dry_threshold));
    add("physics", "wind_stress_x", Kind::REAL, OFF(wind_stress_x));
    add("physics", "wind_stress_y", Kind::REAL, OFF(wind_stress_y));
    add("physics", "coriolis_f", Kind::REAL, OFF(coriolis_f));
    add("physics", "use_hllc", Kind::LOGICAL, OFF(use_hllc));
    add("physics", "use_linear_bottom_drag", Kind::LOGICAL, OFF(use_linear_bottom_drag));
    add("physics", "bottom_drag_r", Kind::REAL, OFF(bottom_drag_r));
    add("physics", "wind_surface_layer", Kind::LOGICAL, OFF(wind_surface_layer));
    // &output_nml: read by the driver executable, not by this API
    for (const char* n : {"output_to_file", "async_output", "compress_output", "out_h_layer",
                          "out_velocity", "out_salinity", "out_temperature", "out_density", "out_w"})
        ign("output", n, Kind::LOGICAL);
    for (const char* n : {"output_dir", "restart_file", "forcing_file", "gauge_file"})
        ign("output", n, Kind::STRING);
    for (const char* n : {"output_interval", "restart_interval", "forcing_x_origin", "forcing_y_origin",
                          "gauge_output_interval"})
        ign("output", n, Kind::REAL);
    ign_enum("output", "output_format", {"ascii", "netcdf"});
    ign("output", "async_buffer_slots", Kind::INT);
    ign("output", "compress_level", Kind::INT);
    add("output", "bathymetry_file", Kind::STRING, OFF(bathymetry_file));
    // &boundary_nml (rki_config.F90:1571-1628)
    en("boundary", "bc_west", OFF(bc[0]), bcs);
    en("boundary", "bc_east", OFF(bc[1]), bcs);
    en("boundary", "bc_south", OFF(bc[2]), bcs);
    en("boundary", "bc_north", OFF(bc[3]), bcs);
    add("boundary", "tidal_amplitude", Kind::REAL, OFF(tidal_amplitude));
    add("boundary", "tidal_period", Kind::REAL, OFF(tidal_period));
    {
        auto* d = add("boundary", "n_tidal_constituents", Kind::INT, OFF(n_tidal_constituents));
        d->has_min = true; d->vmin = 0; d->has_max = true; d->vmax = MAX_TC;
    }
    add("boundary", "tidal_amp", Kind::REAL_ARRAY, OFF(tidal_amp));
    add("boundary", "tidal_phase", Kind::REAL_ARRAY, OFF(tidal_phase));
    add("boundary", "tidal_omega", Kind::REAL_ARRAY, OFF(tidal_omega));
    add("boundary", "inflow_velocity", Kind::REAL, OFF(inflow_velocity));
    add("boundary", "inflow_salinity", Kind::REAL, OFF(inflow_salinity));
    add("boundary", "inflow_temperature", Kind::REAL, OFF(inflow_temperature));
    add("boundary", "discharge_flux", Kind::REAL, OFF(discharge_flux));
    add("boundary", "clamped_elevation", Kind::REAL, OFF(clamped_elevation));
    { auto* d = add("boundary", "sponge_width", Kind::INT, OFF(sponge_width)); d->has_min = true; d->vmin = 0; }
    add("boundary", "sponge_strength", Kind::REAL, OFF(sponge_strength));
    add("boundary", "sponge_h_ref", Kind::REAL, OFF(sponge_h_ref));
    add("boundary", "nesting_file", Kind::STRING, OFF(nesting_file));
    ign("boundary", "nesting_x_origin", Kind::REAL);
    ign("boundary", "nesting_y_origin", Kind::REAL);
    // &tracer_nml (multilayer only)
    for (const char* n : {"initial_salinity", "S_ref", "beta_S", "S_min", "S_max", "kappa_S_bg",
                          "S_init_surface", "S_init_bottom", "initial_temperature", "T_ref", "alpha_T",
                          "T_min", "T_max", "kappa_T_bg", "T_init_surface", "T_init_bottom",
                          "sediment_w_s", "sediment_initial_conc", "sediment_bed_init",
                          "sediment_tau_crit_erosion", "sediment_erosion_rate", "sediment_porosity"})
        ign("tracer", n, Kind::REAL);
    ign_enum("tracer", "eos_form", {"linear", "wright", "roquet_spv"});
    add("tracer", "use_ppm_tracer", Kind::LOGICAL, OFF(use_ppm_tracer));
    ign_enum("tracer", "tracer_recon", {"ppm", "pqm", "weno5", "weno7", "weno9"});
    ign("tracer", "use_sediment", Kind::LOGICAL);
    // &initial_condition_nml (rki_config.F90:1702-1730)
    en("initial_condition", "init_condition", OFF(init_condition),
       {"dam_break", "gaussian_hump", "castro_dam_break", "still_water"});
    add("initial_condition", "eta_initial", Kind::REAL, OFF(eta_initial));
    add("initial_condition", "h0", Kind::REAL, OFF(h0));
    add("initial_condition", "hump_amplitude", Kind::REAL, OFF(hump_amplitude));
    add("initial_condition", "hump_sigma", Kind::REAL, OFF(hump_sigma));
    add("initial_condition", "hump_x_offset", Kind::REAL, OFF(hump_x_offset));
    add("initial_condition", "hump_y_offset", Kind::REAL, OFF(hump_y_offset));
    add("initial_condition", "init_u_shear", Kind::REAL, OFF(init_u_shear));
    // &nonhydrostatic_nml (rki_config.F90:1735-1810)
    add("nonhydrostatic", "use_nonhydrostatic", Kind::LOGICAL, OFF(use_nonhydrostatic));
    { auto* d = add("nonhydrostatic", "nz_layers", Kind::INT, OFF(nz_layers)); d->has_min = true; d->vmin = 1; }
    add("nonhydrostatic", "use_multilayer", Kind::LOGICAL, OFF(use_multilayer));
    for (const char* n : {"nh_tolerance", "rho_0", "vmix_nu_max", "vmix_nu_bg", "vmix_ri_alpha",
                          "kpp_ri_crit", "kpp_cs_nonlocal", "kpp_c_vt2", "hdiff_kappa", "smagorinsky_cs",
                          "biharmonic_nu4", "ml_min_depth"})
        ign("nonhydrostatic", n, Kind::REAL);
    for (const char* n : {"nh_max_iterations", "nh_solve_interval", "split_ratio", "keps_si_picard"})
        ign("nonhydrostatic", n, Kind::INT);
    for (const char* n : {"nh_preconditioner", "split_averaging", "vmix_use_kpp", "vmix_use_keps",
                          "evap_mass_sink", "evap_latent_cooling"})
        ign("nonhydrostatic", n, Kind::LOGICAL);
    ign_enum("nonhydrostatic", "bpg_method", {"sigma", "jacobian"});
    ign_enum("nonhydrostatic", "keps_stability", {"constant", "galperin", "canuto"});
    return k;
}
#undef OFF

const std::vector<KeyDef>& schema() {
    static const std::vector<KeyDef> s = build_schema();
    return s;
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string strip_nml_suffix(const std::string& name) {
    std::string l = lower(name);
    if (l.size() > 4 && l.compare(l.size() - 4, 4, "_nml") == 0) return l.substr(0, l.size() - 4);
    return l;
}

// Remove a trailing '!' comment outside quotes (rki_nml_schema.F90:1160).
std::string strip_comment(const std::string& line) {
    char q = 0;
    for (size_t i = 0; i < line.size(); ++i) {
        const char c = line[i];
        if (q) {
            if (c == q) q = 0;
        } else if (c == '"' || c == '\'') {
            q = c;
        } else if (c == '!') {
            return line.substr(0, i);
        }
    }
    return line;
}

struct Token {
    std::string text;
    bool quoted;
    int line;
};

bool is_ident_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }

bool parse_real(const std::string& s, double* out) {
    if (s.empty()) return false;
    std::string t = s;
    for (char& c : t)
        if (c == 'd' || c == 'D') c = 'e';
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') return false;
    if (!std::isfinite(v)) return false;
    *out = v;
    return true;
}

bool parse_int(const std::string& s, int* out) {
    if (s.empty()) return false;
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (*end != '\0' || v < -2147483647L || v > 2147483647L) return false;
    *out = (int)v;
    return true;
}

bool parse_logical(const std::string& s, bool* out) {
    const std::string l = lower(s);
    if (l == ".true." || l == "t" || l == "true" || l == ".t.") { *out = true; return true; }
    if (l == ".false." || l == "f" || l == "false" || l == ".f.") { *out = false; return true; }
    return false;
}

bool is_repeat_count(const std::string& w) {
    const size_t star = w.find('*');
    if (star == std::string::npos || star == 0) return false;
    for (size_t i = 0; i < star; ++i)
        if (!std::isdigit((unsigned char)w[i])) return false;
    return true;
}

struct Parser {
    Config* cfg;
    std::vector<std::string> errors;
    std::string label;

    void err(int line, const std::string& msg) {
        errors.push_back(label + ":" + std::to_string(line) + ": " + msg);
    }

    const KeyDef* find_key(const std::string& group, const std::string& name) {
        const std::string ln = lower(name);
        for (const KeyDef& d : schema())
            if (group == d.group && ln == lower(d.name)) return &d;
        return nullptr;
    }

    bool known_group(const std::string& g) {
        for (const KeyDef& d : schema())
            if (g == d.group) return true;
        return false;
    }

    void apply(const KeyDef& d, const std::vector<Token>& vals, int line) {
        char* base = reinterpret_cast<char*>(cfg);
        const Kind kind = (d.kind == Kind::IGNORED) ? d.ignored_as : d.kind;
        if (kind != Kind::REAL_ARRAY && vals.size() != 1) {
            err(line, std::string("key '") + d.name + "' expects one value");
            return;
        }
        switch (kind) {
        case Kind::REAL: {
            double v;
            if (vals[0].quoted || !parse_real(vals[0].text, &v)) {
                err(line, std::string("key '") + d.name + "': invalid real '" + vals[0].text + "'");
                return;
            }
            if (d.kind != Kind::IGNORED) *reinterpret_cast<double*>(base + d.off) = v;
            return;
        }
        case Kind::INT: {
            int v;
            if (vals[0].quoted || !parse_int(vals[0].text, &v)) {
                err(line, std::string("key '") + d.name + "': invalid integer '" + vals[0].text + "'");
                return;
            }
            if ((d.has_min && v < d.vmin) || (d.has_max && v > d.vmax)) {
                err(line, std::string("key '") + d.name + "' = " + vals[0].text + " is out of range");
                return;
            }
            if (d.kind != Kind::IGNORED) *reinterpret_cast<int*>(base + d.off) = v;
            return;
        }
        case Kind::LOGICAL: {
            bool v;
            if (vals[0].quoted || !parse_logical(vals[0].text, &v)) {
                err(line, std::string("key '") + d.name + "': invalid logical '" + vals[0].text + "'");
                return;
            }
            if (d.kind != Kind::IGNORED) *reinterpret_cast<bool*>(base + d.off) = v;
            return;
        }
        case Kind::STRING:
            if (d.kind != Kind::IGNORED) *reinterpret_cast<std::string*>(base + d.off) = vals[0].text;
            return;
        case Kind::ENUM: {
            const std::string l = lower(vals[0].text);
            for (const std::string& a : d.allowed)
                if (l == a) {
                    if (d.kind != Kind::IGNORED) *reinterpret_cast<std::string*>(base + d.off) = a;
                    return;
                }
            err(line, std::string("key '") + d.name + "': '" + vals[0].text + "' is not an allowed value");
            return;
        }
        case Kind::REAL_ARRAY: {
            if (vals.size() > (size_t)MAX_TC) {
                err(line, std::string("key '") + d.name + "' takes at most 10 values");
                return;
            }
            double* arr = reinterpret_cast<double*>(base + d.off);
            for (size_t i = 0; i < vals.size(); ++i) {
                double v;
                if (vals[i].quoted || !parse_real(vals[i].text, &v)) {
                    err(line, std::string("key '") + d.name + "': invalid real '" + vals[i].text + "'");
                    return;
                }
                arr[i] = v;
            }
            return;
        }
        default:
            return;
        }
    }

    // Collect the tokens of one group body up to its closing '/'
    // (rki_nml_schema.F90:1470-1560).
    bool collect(const std::vector<std::string>& lines, size_t& li, size_t& col, std::vector<Token>& toks) {
        while (li < lines.size()) {
            const std::string s = strip_comment(lines[li]);
            size_t i = col;
            while (i < s.size()) {
                const char c = s[i];
                if (c == '/') { col = i + 1; return true; }
                if (c == ' ' || c == '\t' || c == '\r') { ++i; continue; }
                if (c == '=' || c == ',') { toks.push_back({std::string(1, c), false, (int)li + 1}); ++i; continue; }
                if (c == '"' || c == '\'') {
                    size_t j = i + 1;
                    while (j < s.size() && s[j] != c) ++j;
                    if (j >= s.size()) {
                        err((int)li + 1, "unterminated string literal");
                        toks.push_back({s.substr(i + 1), true, (int)li + 1});
                        i = s.size();
                    } else {
                        toks.push_back({s.substr(i + 1, j - i - 1), true, (int)li + 1});
                        i = j + 1;
                    }
                    continue;
                }
                size_t j = i;
                while (j < s.size() && s[j] != ' ' && s[j] != '\t' && s[j] != '=' && s[j] != ',' &&
                       s[j] != '/' && s[j] != '\r')
                    ++j;
                const std::string w = s.substr(i, j - i);
                if (w.find('%') != std::string::npos)
                    err((int)li + 1, "derived-type reference '" + w + "' is not supported");
                else if (w.find('(') != std::string::npos)
                    err((int)li + 1, "indexed assignment '" + w + "' is not supported");
                else if (is_repeat_count(w))
                    err((int)li + 1, "repeat-count syntax '" + w + "' is not supported");
                toks.push_back({w, false, (int)li + 1});
                i = j;
            }
            ++li;
            col = 0;
        }
        err((int)lines.size(), "group not terminated with '/' before end of input");
        return false;
    }

    void group_body(const std::string& group, const std::vector<Token>& toks) {
        size_t i = 0;
        std::set<std::string> seen;
        while (i < toks.size()) {
            const Token& key = toks[i];
            if (key.text == ",") { ++i; continue; }
            if (key.quoted || key.text == "=" || i + 1 >= toks.size() || toks[i + 1].text != "=" ||
                toks[i + 1].quoted) {
                err(key.line, "expected 'key = value', found '" + key.text + "'");
                ++i;
                continue;
            }
            i += 2;
            std::vector<Token> vals;
            while (i < toks.size()) {
                const Token& t = toks[i];
                if (!t.quoted && t.text == ",") { ++i; continue; }
                if (!t.quoted && t.text == "=") break;
                if (i + 1 < toks.size() && !toks[i + 1].quoted && toks[i + 1].text == "=") break;
                vals.push_back(t);
                ++i;
            }
            if (vals.empty()) {
                err(key.line, "key '" + key.text + "' has no value");
                continue;
            }
            const KeyDef* d = find_key(group, key.text);
            if (!d) {
                err(key.line, "unknown key '" + key.text + "' in group '" + group + "'");
                continue;
            }
            if (!seen.insert(lower(key.text)).second) {
                err(key.line, "duplicate key '" + key.text + "'");
                continue;
            }
            apply(*d, vals, key.line);
        }
    }

    void run(const std::string& text) {
        std::vector<std::string> lines;
        size_t start = 0;
        while (start <= text.size()) {
            size_t nl = text.find('\n', start);
            if (nl == std::string::npos) { lines.push_back(text.substr(start)); break; }
            lines.push_back(text.substr(start, nl - start));
            start = nl + 1;
        }
        std::set<std::string> groups_seen;
        size_t li = 0, col = 0;
        while (li < lines.size()) {
            const std::string s = strip_comment(lines[li]);
            const size_t amp = s.find('&', col);
            if (amp == std::string::npos) { ++li; col = 0; continue; }
            size_t j = amp + 1;
            while (j < s.size() && (s[j] == ' ' || s[j] == '\t')) ++j;
            size_t k = j;
            while (k < s.size() && is_ident_char(s[k])) ++k;
            const std::string gname = s.substr(j, k - j);
            const int gline = (int)li + 1;
            col = k;
            std::vector<Token> toks;
            const bool closed = collect(lines, li, col, toks);
            if (gname.empty()) { err(gline, "malformed group header (expected name after '&')"); }
            else {
                const std::string g = strip_nml_suffix(gname);
                if (!known_group(g)) err(gline, "unknown group '" + gname + "'");
                else if (!groups_seen.insert(g).second) err(gline, "duplicate occurrence of group '" + gname + "'");
                else group_body(g, toks);
            }
            if (!closed) break;
        }
    }
};

// apply_time_unit_cascade, rki_config.F90:769-801.
bool time_unit_cascade(Config& c, std::string& msg) {
    const std::string u = lower(c.time_unit);
    double f;
    if (u == "s" || u == "sec" || u == "second" || u == "seconds") f = 1.0;
    else if (u == "min" || u == "minute" || u == "minutes") f = 60.0;
    else if (u == "hr" || u == "hour" || u == "hours" || u == "h") f = 3600.0;
    else if (u == "day" || u == "days" || u == "d") f = 86400.0;
    else if (u == "year" || u == "years" || u == "yr") f = 365.25 * 86400.0;
    else { msg = "invalid time_unit '" + c.time_unit + "': must be one of s/min/hr/day/year"; return false; }
    c.t_end = c.t_end * f;
    c.status_interval = c.status_interval * f;
    return true;
}

// validate_config, rki_config.F90:803-860 (the checks that apply here).
bool validate(const Config& c, std::string& msg) {
    if (c.nx < 1) { msg = "invalid nx: must be >= 1"; return false; }
    if (c.ny < 1) { msg = "invalid ny: must be >= 1"; return false; }
    if (c.nghost < 1) { msg = "invalid nghost: must be >= 1"; return false; }
    if (!(c.dx > 0.0)) { msg = "invalid dx: must be > 0"; return false; }
    if (!(c.dy > 0.0)) { msg = "invalid dy: must be > 0"; return false; }
    if (!(c.cfl > 0.0) || c.cfl > 1.0) { msg = "invalid cfl: must be in (0, 1]"; return false; }
    if (!(c.t_end > 0.0)) { msg = "invalid t_end: must be > 0"; return false; }
    if (c.use_multilayer && c.use_nonhydrostatic) {
        msg = "use_multilayer and use_nonhydrostatic are mutually exclusive";
        return false;
    }
    return true;
}

// Features of config_t this build does not run; -5 with a message.
bool supported(const Config& c, std::string& msg) {
    if (c.grid_type != "structured") { msg = "grid_type 'unstructured' is not implemented here"; return false; }
    if (c.time_scheme == "semi_implicit") { msg = "time_scheme 'semi_implicit' is not implemented here"; return false; }
    if (c.use_multilayer) { msg = "use_multilayer is not implemented here"; return false; }
    if (c.use_nonhydrostatic) { msg = "use_nonhydrostatic is not implemented here"; return false; }
    // use_linear_bottom_drag / bottom_drag_r: read only by the multilayer
    // step (rki_solver.F90:1393,1522,1684); the 2-D step ignores them
    for (int s = 0; s < 4; ++s) {
        if (c.bc[s] == "nested") { msg = "boundary '" + c.bc[s] + "' is not implemented here"; return false; }
    }
    if (c.nghost < 2) { msg = "nghost must be >= 2 (the stencil reaches two cells)"; return false; }
    if (c.nx < 2 || c.ny < 2) { msg = "nx and ny must be >= 2 on this path"; return false; }
    return true;
}

int time_scheme_tag(const std::string& s) {   // parse_time_scheme, rki_timestepper.F90:102-120
    if (s == "euler") return RKG_TIME_EULER;
    if (s == "rk3") return RKG_TIME_RK3;
    if (s == "rk43") return RKG_TIME_RK43;
    return RKG_TIME_RK2;
}

int bc_tag(const std::string& s) {   // bc_type_from_string, rki_boundary_types.F90:97-125
    if (s == "wall") return RKG_BC_WALL;
    if (s == "open") return RKG_BC_OPEN;
    if (s == "tidal") return RKG_BC_TIDAL;
    if (s == "inflow") return RKG_BC_INFLOW;
    if (s == "discharge") return RKG_BC_DISCHARGE;
    if (s == "clamped") return RKG_BC_CLAMPED;
    if (s == "sponge") return RKG_BC_SPONGE;
    if (s == "chapman") return RKG_BC_CHAPMAN;
    return RKG_BC_WALL;
}

// ------------------------------------------------------------------------
// Handle
// ------------------------------------------------------------------------
struct BcParams {   // bc_params_t fields the step reads, rki_boundary_types.F90:37-93
    int tag[4];
    double tidal_amplitude, tidal_period;
    int n_tc;
    double amp[MAX_TC], phase[MAX_TC], omega[MAX_TC];
    double inflow_velocity, inflow_salinity, inflow_temperature, discharge_flux, clamped_elevation;
};

struct Sim {
    uint32_t magic = HANDLE_MAGIC;
    Config cfg;
    int nx = 0, ny = 0, ng = 0, nxt = 0, nyt = 0;
    double dx = 0.0, dy = 0.0;
    std::vector<double> h, hu, hv, b, h0, hu0, hv0;
    // state%forcing (rki_surface_forcing.F90), interior cells only, nx by ny:
    // the kernels read wind and precipitation at the cell and the pressure
    // of the four neighbours, whose ghost values no setter writes (0).
    std::vector<double> wind_x, wind_y, atm_pressure, precipitation, heat_flux, evap_minus_precip;
    bool forcing_active = false;
    BcParams bc{};
    double manning_n = 0.0, coriolis_f = 0.0;
    bool use_hllc = false, use_ppm_tracer = true;
    rkg_grid* dev = nullptr;
    bool is_initialised = false, is_finalised = false;
    bool user_set_bathymetry = false, user_set_h = false;
    bool on_device = false, host_is_current = false;
    bool stepped = false;      // device stepped since the host ghosts were last valid
    double t_current = 0.0;

    size_t ix(int i, int j) const { return (size_t)(i - 1) + (size_t)(j - 1) * (size_t)nxt; }  // 1-based
    double* interior(std::vector<double>& f) { return f.data() + ix(ng + 1, ng + 1); }
};

std::mutex g_registry_mu;
std::set<const void*> g_registry;

Sim* resolve(void* handle) {
    if (!handle) return nullptr;
    std::lock_guard<std::mutex> lk(g_registry_mu);
    if (!g_registry.count(handle)) return nullptr;
    Sim* s = static_cast<Sim*>(handle);
    return s->magic == HANDLE_MAGIC ? s : nullptr;
}

int rkg_rc(int rc, const char* what) {
    if (rc != RKG_OK) set_error("%s failed (%d)", what, rc);
    return rc;
}

rkg_config device_config(const Sim& s) {
    rkg_config c;
    rkg_config_default(&c);
    c.nx = s.nx;
    c.ny = s.ny;
    c.dx = s.dx;
    c.dy = s.dy;
    c.g = GRAVITY;
    c.cfl = s.cfg.cfl;
    c.manning_n = s.manning_n;
    c.coriolis_f = s.coriolis_f;
    c.dt_fixed = s.cfg.dt_fixed;
    c.t_end = s.cfg.t_end;
    c.cfl_interval = s.cfg.cfl_interval;
    c.use_hllc = s.use_hllc ? 1 : 0;
    for (int k = 0; k < 4; ++k) c.bc[k] = s.bc.tag[k];
    c.tidal_amplitude = s.bc.tidal_amplitude;
    c.tidal_period = s.bc.tidal_period;
    c.n_tidal_constituents = s.bc.n_tc;
    for (int k = 0; k < MAX_TC; ++k) {
        c.tidal_amp[k] = s.bc.amp[k];
        c.tidal_phase[k] = s.bc.phase[k];
        c.tidal_omega[k] = s.bc.omega[k];
    }
    c.inflow_velocity = s.bc.inflow_velocity;
    c.discharge_flux = s.bc.discharge_flux;
    c.clamped_elevation = s.bc.clamped_elevation;
    return c;
}

// Scalars changed on the host (BC setters, use_hllc) reach the device
// before the next step, as the reference's bc_params_t does.
int push_config(Sim& s) {
    if (!s.dev) return RKG_OK;
    const rkg_config c = device_config(s);
    return rkg_rc(rkg_set_config(s.dev, &c), "rkg_set_config");
}

void bc_init_from_config(Sim& s) {   // rki_boundary_types.F90:127-153
    const Config& c = s.cfg;
    for (int k = 0; k < 4; ++k) s.bc.tag[k] = bc_tag(c.bc[k]);
    s.bc.tidal_amplitude = c.tidal_amplitude;
    s.bc.tidal_period = c.tidal_period;
    s.bc.n_tc = c.n_tidal_constituents;
    for (int k = 0; k < MAX_TC; ++k) {
        s.bc.amp[k] = c.tidal_amp[k];
        s.bc.phase[k] = c.tidal_phase[k];
        s.bc.omega[k] = c.tidal_omega[k];
    }
    s.bc.inflow_velocity = c.inflow_velocity;
    s.bc.inflow_salinity = c.inflow_salinity;
    s.bc.inflow_temperature = c.inflow_temperature;
    s.bc.discharge_flux = c.discharge_flux;
    s.bc.clamped_elevation = c.clamped_elevation;
}

// fill_bathymetry_ghosts_array, rki_bathymetry.F90:165-196.
void fill_bathymetry_ghosts(Sim& s) {
    for (int j = 1; j <= s.nyt; ++j)
        for (int i = 1; i <= s.ng; ++i) {
            s.b[s.ix(i, j)] = s.b[s.ix(s.ng + 1, j)];
            s.b[s.ix(s.ng + s.nx + i, j)] = s.b[s.ix(s.ng + s.nx, j)];
        }
    for (int j = 1; j <= s.ng; ++j)
        for (int i = 1; i <= s.nxt; ++i) {
            s.b[s.ix(i, j)] = s.b[s.ix(i, s.ng + 1)];
            s.b[s.ix(i, s.ng + s.ny + j)] = s.b[s.ix(i, s.ng + s.ny)];
        }
}

inline double max0(double x) { return x > 0.0 ? x : 0.0; }   // Fortran max(x, 0.0)

// init_from_config and the four initial conditions, rki_init.F90:26-212.
void init_from_config(Sim& s) {
    const Config& c = s.cfg;
    const size_t n = s.h.size();
    bool has_bathymetry = false;
    for (size_t k = 0; k < n; ++k)
        if (s.b[k] != 0.0) { has_bathymetry = true; break; }
    if (c.init_condition == "still_water") {
        for (size_t k = 0; k < n; ++k) {
            s.h[k] = max0(c.eta_initial - s.b[k]);
            s.hu[k] = 0.0;
            s.hv[k] = 0.0;
        }
    } else if (c.init_condition == "gaussian_hump") {
        const double xc = (double)s.nx * s.dx * 0.5 + c.hump_x_offset;
        const double yc = (double)s.ny * s.dy * 0.5 + c.hump_y_offset;
        for (int j = 1; j <= s.nyt; ++j)
            for (int i = 1; i <= s.nxt; ++i) {
                const double x = ((double)(i - s.ng) - 0.5) * s.dx;
                const double y = ((double)(j - s.ng) - 0.5) * s.dy;
                const double r2 = (x - xc) * (x - xc) + (y - yc) * (y - yc);
                const size_t k = s.ix(i, j);
                if (!has_bathymetry) s.b[k] = 0.0;
                s.h[k] = c.h0 + c.hump_amplitude * std::exp(-r2 / (2.0 * (c.hump_sigma * c.hump_sigma)));
                s.hu[k] = 0.0;
                s.hv[k] = 0.0;
            }
    } else if (c.init_condition == "castro_dam_break") {
        const double x0 = -0.5 * (double)s.nx * s.dx;
        const double y0 = -0.5 * (double)s.ny * s.dy;
        for (int j = 1; j <= s.nyt; ++j)
            for (int i = 1; i <= s.nxt; ++i) {
                const double x = x0 + ((double)(i - s.ng) - 0.5) * s.dx;
                const double y = y0 + ((double)(j - s.ng) - 0.5) * s.dy;
                const double depth = 1.0 - 0.4 * std::exp(-(x * x) - y * y);
                const size_t k = s.ix(i, j);
                s.b[k] = -depth;
                s.h[k] = (x * x + y * y <= 0.36) ? 3.0 - s.b[k] : 1.0 - s.b[k];
                s.hu[k] = 0.0;
                s.hv[k] = 0.0;
            }
    } else {   // dam_break (default)
        const int mid_x = s.ng + s.nx / 2;
        for (int j = 1; j <= s.nyt; ++j)
            for (int i = 1; i <= s.nxt; ++i) {
                const size_t k = s.ix(i, j);
                if (!has_bathymetry) s.b[k] = 0.0;
                s.hu[k] = 0.0;
                s.hv[k] = 0.0;
                s.h[k] = (i <= mid_x) ? 2.0 : 1.0;
            }
    }
}

// Tidal surface at `time`, rki_boundary.F90:92-101, with the same sin/cos
// the device uses.
double tidal_eta(const BcParams& bc, double time) {
    double e = 0.0;
    if (bc.n_tc > 0) {
        for (int n = 0; n < bc.n_tc; ++n) e = e + bc.amp[n] * rk::crm::cos(bc.omega[n] * time - bc.phase[n]);
    } else if (bc.tidal_period > 0.0) {
        e = bc.tidal_amplitude * rk::crm::sin(6.283185307179586 * time / bc.tidal_period);
    }
    return e;
}

// apply_boundary_conditions on full ghosted arrays, rki_boundary.F90:40-466:
// edges W, E, S, N in turn, W/E spanning all rows and S/N all columns.
// Chapman edges (:471-552) radiate on the adjacent layer of interior rows
// / columns against eta_old = prev_h + b, the previous boundary call's
// state, when prev_h is given and bc_dt > 0; elsewhere (the outer layer,
// and the corner blocks, whose eta_old the library does not keep) the ghost
// surface is the interior eta.
void apply_bc(Sim& s, double time, double* h, double* hu, double* hv, double* b, double bc_dt = 0.0,
              const double* prev_h = nullptr) {
    const int ng = s.ng;
    const double eta_tide = tidal_eta(s.bc, time);
    for (int side = 0; side < 4; ++side) {
        const bool xedge = side < 2;
        const bool lo = (side % 2) == 0;
        const int n_tan = xedge ? s.nyt : s.nxt;
        const int tag = s.bc.tag[side];
        const double sgn = lo ? -1.0 : 1.0;
        for (int t = 1; t <= n_tan; ++t) {
            for (int kk = 1; kk <= ng; ++kk) {
                // ghost (ig, jg), mirror (im, jm), first interior (ii, ji)
                int ig, jg, im, jm, ii, ji;
                if (xedge) {
                    jg = jm = ji = t;
                    ig = lo ? ng + 1 - kk : ng + s.nx + kk;
                    im = lo ? ng + kk : ng + s.nx + 1 - kk;
                    ii = lo ? ng + 1 : ng + s.nx;
                } else {
                    ig = im = ii = t;
                    jg = lo ? ng + 1 - kk : ng + s.ny + kk;
                    jm = lo ? ng + kk : ng + s.ny + 1 - kk;
                    ji = lo ? ng + 1 : ng + s.ny;
                }
                const size_t G = s.ix(ig, jg), M = s.ix(im, jm), I = s.ix(ii, ji);
                double* nrm = xedge ? hu : hv;
                double* tan = xedge ? hv : hu;
                switch (tag) {
                case RKG_BC_WALL:
                case RKG_BC_SPONGE:
                    b[G] = b[M];
                    h[G] = max0(h[M] + b[M] - b[G]);
                    nrm[G] = -nrm[M];
                    tan[G] = tan[M];
                    break;
                case RKG_BC_OPEN:
                case RKG_BC_TIDAL: {
                    const bool prescribe = tag == RKG_BC_TIDAL;
                    const double eta_ext = prescribe ? eta_tide : 0.0;
                    const double h_int = h[I];
                    const double eta_int = h_int + b[I];
                    const double hmax = h_int > 1.0e-6 ? h_int : 1.0e-6;
                    const double speed = std::sqrt(GRAVITY * hmax);
                    const double hg = prescribe ? max0(eta_ext - b[G]) : h_int;
                    h[G] = hg;
                    nrm[G] = hg * (nrm[I] / hmax + sgn * speed * (eta_int - eta_ext) / hmax);
                    tan[G] = tan[I];
                    break;
                }
                case RKG_BC_CLAMPED:
                    h[G] = max0(s.bc.clamped_elevation - b[G]);
                    hu[G] = hu[I];
                    hv[G] = hv[I];
                    break;
                case RKG_BC_CHAPMAN: {
                    const double h_int = h[I];
                    const double eta_int = h_int + b[I];
                    const double hmax = h_int > 1.0e-6 ? h_int : 1.0e-6;
                    const double speed = std::sqrt(GRAVITY * hmax);
                    double eta_g = eta_int;
                    const bool interior_t = t > ng && t <= n_tan - ng;
                    if (kk == 1 && interior_t && prev_h && bc_dt > 0.0) {
                        const size_t I2 = xedge ? s.ix(lo ? ng + 2 : ng + s.nx - 1, t)
                                                : s.ix(t, lo ? ng + 2 : ng + s.ny - 1);
                        const double dd = xedge ? s.dx : s.dy;
                        const double eta_old = prev_h[I] + b[I];
                        const double deta_dt = (eta_int - eta_old) / bc_dt;
                        const double deta_dx = (h[I2] + b[I2] - eta_int) / dd;
                        double cp = std::fabs(deta_dx) > 1.0e-10 ? -deta_dt / deta_dx : 0.0;
                        cp = std::fmin(cp, speed);
                        cp = 0.0 > cp ? 0.0 : cp;
                        eta_g = eta_old - cp * bc_dt / dd * (eta_old - eta_int);
                    }
                    const double hg = max0(eta_g - b[G]);
                    h[G] = hg;
                    nrm[G] = hg * (nrm[I] / hmax + sgn * speed * (eta_int - eta_g) / hmax);
                    tan[G] = tan[I];
                    break;
                }
                default: {   // inflow, discharge
                    const double in_sgn = -sgn;
                    const double h_int = h[I];
                    const double q = (tag == RKG_BC_INFLOW) ? in_sgn * h_int * s.bc.inflow_velocity
                                                            : in_sgn * s.bc.discharge_flux;
                    h[G] = h_int;
                    nrm[G] = q;
                    tan[G] = tan[I];
                    break;
                }
                }
            }
        }
    }
}

int download_interior(Sim& s, int field, std::vector<double>& dst) {
    return rkg_rc(rkg_download(s.dev, field, s.interior(dst), s.nxt), "rkg_download");
}

// Device -> host copy of the state.  After a step the reference's ghost
// cells hold the step's last boundary call: on U* at the new time for
// SSP-RK2 (rki_solver.F90:790-795), on the last stage's input at
// t + c_last dt for the generic schemes (:895-897); they are rebuilt here
// from that state.
int copy_back(Sim& s) {
    int rc;
    if ((rc = download_interior(s, RKG_FIELD_H, s.h))) return rc;
    if ((rc = download_interior(s, RKG_FIELD_HU, s.hu))) return rc;
    if ((rc = download_interior(s, RKG_FIELD_HV, s.hv))) return rc;
    if ((rc = download_interior(s, RKG_FIELD_H0, s.h0))) return rc;
    if ((rc = download_interior(s, RKG_FIELD_HU0, s.hu0))) return rc;
    if ((rc = download_interior(s, RKG_FIELD_HV0, s.hv0))) return rc;
    if (s.stepped) {
        const size_t n = s.h.size();
        std::vector<double> uh(n, 0.0), uhu(n, 0.0), uhv(n, 0.0);
        std::vector<double> prev(n, 0.0);
        if ((rc = download_interior(s, RKG_FIELD_BC_H, uh))) return rc;
        if ((rc = download_interior(s, RKG_FIELD_BC_HU, uhu))) return rc;
        if ((rc = download_interior(s, RKG_FIELD_BC_HV, uhv))) return rc;
        if ((rc = download_interior(s, RKG_FIELD_BC_PREV_H, prev))) return rc;
        // ghosts of U* from the current ghosts of h, hu, hv (corner
        // blocks read them), then the boundary call at the new time
        for (int j = 1; j <= s.nyt; ++j)
            for (int i = 1; i <= s.nxt; ++i) {
                const bool in = i > s.ng && i <= s.ng + s.nx && j > s.ng && j <= s.ng + s.ny;
                if (in) continue;
                const size_t k = s.ix(i, j);
                uh[k] = s.h[k];
                uhu[k] = s.hu[k];
                uhv[k] = s.hv[k];
            }
        double t_bc, dt_last;
        if ((rc = rkg_rc(rkg_get_last_bc(s.dev, &t_bc, &dt_last), "rkg_get_last_bc"))) return rc;
        // the last boundary call of the step: its state, bc%time and bc%dt,
        // and the previous call's state (Chapman eta_old)
        apply_bc(s, t_bc, uh.data(), uhu.data(), uhv.data(), s.b.data(), dt_last, prev.data());
        for (int j = 1; j <= s.nyt; ++j)
            for (int i = 1; i <= s.nxt; ++i) {
                const bool in = i > s.ng && i <= s.ng + s.nx && j > s.ng && j <= s.ng + s.ny;
                if (in) continue;
                const size_t k = s.ix(i, j);
                s.h[k] = uh[k];
                s.hu[k] = uhu[k];
                s.hv[k] = uhv[k];
            }
        s.stepped = false;
    }
    return RKG_OK;
}

int upload_all(Sim& s) {
    int rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_B, s.interior(s.b), s.nxt), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_H, s.interior(s.h), s.nxt), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_HU, s.interior(s.hu), s.nxt), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_HV, s.interior(s.hv), s.nxt), "rkg_upload");
    return rc;
}

// rki_api.F90:104-137.
int ensure_host(Sim& s) {
    if (s.host_is_current) return 0;
    if (s.on_device) {
        const int rc = copy_back(s);
        if (rc) return rc;
        s.on_device = false;
        s.host_is_current = true;
    }
    return 0;
}

int ensure_device(Sim& s) {
    if (!s.on_device) {
        const int rc = upload_all(s);
        if (rc) return rc;
        s.on_device = true;
    }
    s.host_is_current = false;
    return 0;
}

int refresh_host(Sim& s) {
    if (!s.is_initialised || s.host_is_current) return 0;
    if (s.on_device) {
        const int rc = copy_back(s);
        if (rc) return rc;
        s.host_is_current = true;
    }
    return 0;
}

int push_state(Sim& s) {
    if (!s.on_device) return 0;
    int rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_H, s.interior(s.h), s.nxt), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_HU, s.interior(s.hu), s.nxt), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_HV, s.interior(s.hv), s.nxt), "rkg_upload");
    if (rc) return rc;
    s.host_is_current = false;
    return 0;
}

// solver_update_forcing: the forcing planes of the device grid, once it
// exists and the forcing is on.
int push_forcing(Sim& s) {
    if (!s.dev || !s.forcing_active) return 0;
    int rc = rkg_rc(rkg_set_forcing(s.dev, 1), "rkg_set_forcing");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_WIND_X, s.wind_x.data(), s.nx), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_WIND_Y, s.wind_y.data(), s.nx), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_PRESSURE, s.atm_pressure.data(), s.nx), "rkg_upload");
    if (!rc) rc = rkg_rc(rkg_upload(s.dev, RKG_FIELD_PRECIP, s.precipitation.data(), s.nx), "rkg_upload");
    return rc;
}

int read_clock(Sim& s) {
    double t, dt;
    long long n;
    const int rc = rkg_rc(rkg_get_clock(s.dev, &t, &dt, &n), "rkg_get_clock");
    if (!rc) s.t_current = t;
    return rc;
}

// Phase 1 common to rki_create and rki_create_from_string.
int create_common(const std::string& text, const std::string& label, void** handle) {
    if (!handle) return -100;
    *handle = nullptr;
    g_last_error.clear();
    Sim* s = new Sim();
    Parser p{&s->cfg, {}, label};
    p.run(text);
    std::string msg;
    int rc = 0;
    if (!p.errors.empty()) {
        std::string all;
        for (const std::string& e : p.errors) all += (all.empty() ? "" : "; ") + e;
        set_error("namelist validation failed: %s", all.c_str());
        rc = -2;
    } else if (!time_unit_cascade(s->cfg, msg) || !validate(s->cfg, msg)) {
        set_error("%s", msg.c_str());
        rc = -2;
    } else if (!supported(s->cfg, msg)) {
        set_error("%s", msg.c_str());
        rc = RKG_ERR_UNSUPPORTED;
    }
    if (rc) {
        delete s;
        return rc;
    }
    const Config& c = s->cfg;
    s->nx = c.nx;
    s->ny = c.ny;
    s->ng = c.nghost;
    s->nxt = c.nx + 2 * c.nghost;
    s->nyt = c.ny + 2 * c.nghost;
    s->dx = c.dx;
    s->dy = c.dy;
    const size_t n = (size_t)s->nxt * (size_t)s->nyt;
    for (std::vector<double>* f : {&s->h, &s->hu, &s->hv, &s->b, &s->h0, &s->hu0, &s->hv0}) f->assign(n, 0.0);
    const size_t ni = (size_t)c.nx * (size_t)c.ny;
    s->wind_x.assign(ni, c.wind_stress_x);   // uniform fallback, rki_state.F90:272-274
    s->wind_y.assign(ni, c.wind_stress_y);
    for (std::vector<double>* f : {&s->atm_pressure, &s->precipitation, &s->heat_flux, &s->evap_minus_precip})
        f->assign(ni, 0.0);
    s->manning_n = c.manning_n;
    s->coriolis_f = c.coriolis_f;
    s->use_hllc = c.use_hllc;
    s->use_ppm_tracer = c.use_ppm_tracer;
    bc_init_from_config(*s);
    s->t_current = 0.0;
    s->is_initialised = true;
    {
        std::lock_guard<std::mutex> lk(g_registry_mu);
        g_registry.insert(s);
    }
    *handle = s;
    return 0;
}

bool read_file(const std::string& path, std::string& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[65536];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, r);
    std::fclose(f);
    return true;
}

int set_interior(Sim& s, std::vector<double>& f, const double* src, int nx, int ny) {
    if (!src) return -2;
    for (int j = 0; j < ny; ++j)
        std::memcpy(&f[s.ix(s.ng + 1, s.ng + 1 + j)], src + (size_t)j * nx, (size_t)nx * sizeof(double));
    return 0;
}

int field_ptr(void* handle, std::vector<double> Sim::*field, double** ptr, int* nxt, int* nyt) {
    if (ptr) *ptr = nullptr;
    if (nxt) *nxt = 0;
    if (nyt) *nyt = 0;
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    if (!ptr || !nxt || !nyt) return -2;
    const int rc = ensure_host(*s);
    if (rc) return rc;
    *ptr = (s->*field).data();
    *nxt = s->nxt;
    *nyt = s->nyt;
    return 0;
}

int field_check(Sim* s, const void* data, int nx, int ny) {
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    if (nx != s->nx || ny != s->ny) return -2;
    if (!data) return -2;
    return 0;
}

int bc_scalar(void* handle, double BcParams::*field, double v) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    s->bc.*field = v;
    return push_config(*s);
}

constexpr char RESTART_MAGIC[4] = {'R', 'K', 'R', 'S'};
constexpr int32_t RESTART_VERSION = 1;

}  // namespace

// ==========================================================================
// Exported entry points
// ==========================================================================
extern "C" {

int rki_working_precision(void) { return (int)sizeof(double); }

const char* rki_last_error(void) { return g_last_error.c_str(); }

int rki_create(const char* filename, int len, void** handle) {
    if (!handle) return -100;
    *handle = nullptr;
    if (!filename || len < 0) return -2;
    const std::string path(filename, (size_t)len);
    std::string text;
    // read_config keeps the defaults when the file does not exist
    // (rki_config.F90:484-491)
    read_file(path, text);
    return create_common(text, path, handle);
}

int rki_create_from_string(const char* text, int len, void** handle) {
    if (!handle) return -100;
    *handle = nullptr;
    if (!text || len < 0) return -2;
    return create_common(std::string(text, (size_t)len), "<config-string>", handle);
}

int rki_finalize(void* handle) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    if (s->is_finalised) return -2;
    if (!s->user_set_bathymetry && !s->cfg.bathymetry_file.empty()) {
        set_error("bathymetry_file '%s': NetCDF input is not implemented here; use rki_set_bathymetry",
                  s->cfg.bathymetry_file.c_str());
        return RKG_ERR_UNSUPPORTED;
    }
    fill_bathymetry_ghosts(*s);
    if (!s->user_set_h) init_from_config(*s);
    bc_init_from_config(*s);
    const rkg_config c = device_config(*s);
    int rc = rkg_rc(rkg_create(&c, nullptr, &s->dev), "rkg_create");
    if (!rc) rc = rkg_rc(rkg_set_rk_save(s->dev, 1), "rkg_set_rk_save");
    if (!rc) rc = rkg_rc(rkg_set_time_scheme(s->dev, time_scheme_tag(s->cfg.time_scheme)), "rkg_set_time_scheme");
    if (!rc) rc = rkg_rc(rkg_set_clock(s->dev, s->t_current, 0), "rkg_set_clock");
    if (!rc) rc = upload_all(*s);
    if (!rc) rc = push_forcing(*s);
    if (!rc)   // bc_params_t sponge fields (rki_boundary_types.F90:147-149)
        rc = rkg_rc(rkg_set_sponge(s->dev, s->cfg.sponge_width, s->cfg.sponge_strength, s->cfg.sponge_h_ref),
                    "rkg_set_sponge");
    if (rc) {
        if (s->dev) rkg_destroy(&s->dev);
        return rc;
    }
    s->on_device = true;
    s->host_is_current = false;
    s->stepped = false;
    s->is_finalised = true;
    return 0;
}

int rki_create_from_namelist(const char* filename, int len, void** handle) {
    int rc = rki_create(filename, len, handle);
    if (rc) return rc;
    rc = rki_finalize(*handle);
    if (rc) rki_destroy(handle);
    return rc;
}

int rki_create_from_namelist_string(const char* text, int len, void** handle) {
    int rc = rki_create_from_string(text, len, handle);
    if (rc) return rc;
    rc = rki_finalize(*handle);
    if (rc) rki_destroy(handle);
    return rc;
}

void rki_destroy(void** handle) {
    if (!handle) return;
    Sim* s = resolve(*handle);
    if (s) {
        {
            std::lock_guard<std::mutex> lk(g_registry_mu);
            g_registry.erase(s);
        }
        if (s->dev) rkg_destroy(&s->dev);
        s->magic = HANDLE_DEAD;
        delete s;
    }
    *handle = nullptr;
}

int rki_step(void* handle, int n_steps) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_finalised) return -1;
    int rc = ensure_device(*s);
    if (rc) return rc;
    s->host_is_current = false;
    if (n_steps <= 0) return 0;
    rc = rkg_rc(rkg_step(s->dev, n_steps), "rkg_step");
    if (!rc) rc = read_clock(*s);
    s->stepped = true;
    return rc;
}

int rki_advance(void* handle, double target_time) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_finalised) return -1;
    int rc = ensure_device(*s);
    if (rc) return rc;
    s->host_is_current = false;
    long long taken = 0;
    rc = rkg_rc(rkg_advance(s->dev, target_time, &taken), "rkg_advance");
    if (!rc) rc = read_clock(*s);
    if (taken > 0) s->stepped = true;
    return rc;
}

double rki_get_time(void* handle) {
    Sim* s = resolve(handle);
    return s ? s->t_current : 0.0;
}

double rki_get_dt(void* handle) {
    Sim* s = resolve(handle);
    if (!s || !s->is_initialised || !s->dev) return 0.0;
    double t, dt;
    long long n;
    return rkg_get_clock(s->dev, &t, &dt, &n) == RKG_OK ? dt : 0.0;
}

void rki_get_grid_info(void* handle, int* nx, int* ny, int* nghost, double* dx, double* dy) {
    Sim* s = resolve(handle);
    const bool ok = s && s->is_initialised;
    if (nx) *nx = ok ? s->nx : 0;
    if (ny) *ny = ok ? s->ny : 0;
    if (nghost) *nghost = ok ? s->ng : 0;
    if (dx) *dx = ok ? s->dx : 0.0;
    if (dy) *dy = ok ? s->dy : 0.0;
}

int rki_get_step_count(void* handle) {
    Sim* s = resolve(handle);
    if (!s || !s->is_initialised || !s->dev) return 0;
    double t, dt;
    long long n;
    return rkg_get_clock(s->dev, &t, &dt, &n) == RKG_OK ? (int)n : 0;
}

int rki_get_nz_ml(void* handle) {
    (void)handle;
    return 0;
}

int rki_get_h_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::h, p, nx, ny); }
int rki_get_hu_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::hu, p, nx, ny); }
int rki_get_hv_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::hv, p, nx, ny); }
int rki_get_b_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::b, p, nx, ny); }
int rki_get_h0_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::h0, p, nx, ny); }
int rki_get_hu0_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::hu0, p, nx, ny); }
int rki_get_hv0_ptr(void* h, double** p, int* nx, int* ny) { return field_ptr(h, &Sim::hv0, p, nx, ny); }

static int layer_ptr(void* handle, double** ptr, int* nxt, int* nyt, int* nz) {
    if (ptr) *ptr = nullptr;
    if (nxt) *nxt = 0;
    if (nyt) *nyt = 0;
    if (nz) *nz = 0;
    Sim* s = resolve(handle);
    if (!s) return -100;
    return -1;   // not initialised or multilayer inactive
}
int rki_get_h_layer_ptr(void* h, double** p, int* a, int* b, int* c) { return layer_ptr(h, p, a, b, c); }
int rki_get_hu_layer_ptr(void* h, double** p, int* a, int* b, int* c) { return layer_ptr(h, p, a, b, c); }
int rki_get_hv_layer_ptr(void* h, double** p, int* a, int* b, int* c) { return layer_ptr(h, p, a, b, c); }
int rki_get_salinity_ptr(void* h, double** p, int* a, int* b, int* c) { return layer_ptr(h, p, a, b, c); }
int rki_get_temperature_ptr(void* h, double** p, int* a, int* b, int* c) { return layer_ptr(h, p, a, b, c); }

int rki_set_bathymetry(void* handle, const double* b, int nx, int ny) {
    Sim* s = resolve(handle);
    int rc = field_check(s, b, nx, ny);
    if (rc) return rc;
    if (s->on_device && (rc = ensure_host(*s))) return rc;
    set_interior(*s, s->b, b, nx, ny);
    s->user_set_bathymetry = true;
    if (s->is_finalised) fill_bathymetry_ghosts(*s);
    return 0;
}

int rki_set_initial_h(void* handle, const double* h, int nx, int ny) {
    Sim* s = resolve(handle);
    int rc = field_check(s, h, nx, ny);
    if (rc) return rc;
    if (s->on_device && (rc = ensure_host(*s))) return rc;
    set_interior(*s, s->h, h, nx, ny);
    s->user_set_h = true;
    return 0;
}

static int set_state_field(void* handle, std::vector<double> Sim::*field, const double* src, int nx, int ny,
                           bool is_h) {
    Sim* s = resolve(handle);
    int rc = field_check(s, src, nx, ny);
    if (rc) return rc;
    if ((rc = refresh_host(*s))) return rc;
    set_interior(*s, s->*field, src, nx, ny);
    if (is_h) s->user_set_h = true;
    return push_state(*s);
}

int rki_set_h(void* h, const double* d, int nx, int ny) { return set_state_field(h, &Sim::h, d, nx, ny, true); }
int rki_set_hu(void* h, const double* d, int nx, int ny) { return set_state_field(h, &Sim::hu, d, nx, ny, false); }
int rki_set_hv(void* h, const double* d, int nx, int ny) { return set_state_field(h, &Sim::hv, d, nx, ny, false); }

// Stores one forcing field (rki_api.F90:896-1034).  `activate` is the
// setters of wind, pressure and precipitation, which set
// state%forcing%active; heat flux and E-P only feed the multilayer path
// and leave the 2-D step unchanged.
static int set_forcing_field(void* handle, std::vector<double> Sim::*field, const double* src, int nx, int ny,
                             bool activate) {
    Sim* s = resolve(handle);
    int rc = field_check(s, src, nx, ny);
    if (rc) return rc;
    std::memcpy((s->*field).data(), src, (size_t)nx * (size_t)ny * sizeof(double));
    if (activate) s->forcing_active = true;
    return push_forcing(*s);
}

int rki_set_wind(void* h, const double* wx, const double* wy, int nx, int ny) {
    Sim* s = resolve(h);
    int rc = field_check(s, wx, nx, ny);
    if (!rc) rc = field_check(s, wy, nx, ny);
    if (rc) return rc;
    std::memcpy(s->wind_x.data(), wx, (size_t)nx * (size_t)ny * sizeof(double));
    return set_forcing_field(h, &Sim::wind_y, wy, nx, ny, true);
}
int rki_set_atm_pressure(void* h, const double* p, int nx, int ny) {
    return set_forcing_field(h, &Sim::atm_pressure, p, nx, ny, true);
}
int rki_set_precipitation(void* h, const double* p, int nx, int ny) {
    return set_forcing_field(h, &Sim::precipitation, p, nx, ny, true);
}
int rki_set_heat_flux(void* h, const double* q, int nx, int ny) {
    return set_forcing_field(h, &Sim::heat_flux, q, nx, ny, false);
}
int rki_set_evap_minus_precip(void* h, const double* e, int nx, int ny) {
    return set_forcing_field(h, &Sim::evap_minus_precip, e, nx, ny, false);
}

static int tracer_set(void* handle) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    return -3;   // multilayer inactive (rki_api.F90:1138-1141)
}
int rki_set_salinity(void* h, const double*, int, int, int) { return tracer_set(h); }
int rki_set_temperature(void* h, const double*, int, int, int) { return tracer_set(h); }

int rki_set_inflow_velocity(void* h, double v) { return bc_scalar(h, &BcParams::inflow_velocity, v); }
int rki_set_inflow_salinity(void* h, double v) { return bc_scalar(h, &BcParams::inflow_salinity, v); }
int rki_set_inflow_temperature(void* h, double v) { return bc_scalar(h, &BcParams::inflow_temperature, v); }
int rki_set_discharge_flux(void* h, double v) { return bc_scalar(h, &BcParams::discharge_flux, v); }
int rki_set_clamped_elevation(void* h, double v) { return bc_scalar(h, &BcParams::clamped_elevation, v); }

double rki_get_clamped_elevation(void* handle) {
    Sim* s = resolve(handle);
    return (s && s->is_initialised) ? s->bc.clamped_elevation : 0.0;
}
double rki_get_manning_n(void* handle) {
    Sim* s = resolve(handle);
    return (s && s->is_initialised) ? s->manning_n : 0.0;
}
double rki_get_coriolis_f(void* handle) {
    Sim* s = resolve(handle);
    return (s && s->is_initialised) ? s->coriolis_f : 0.0;
}

int rki_set_tidal_constituents(void* handle, const double* amp, const double* phase, const double* omega, int n) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    if (n < 0 || n > MAX_TC) return -2;
    if (n > 0 && (!amp || !phase || !omega)) return -2;
    for (int k = 0; k < MAX_TC; ++k) s->bc.amp[k] = s->bc.phase[k] = s->bc.omega[k] = 0.0;
    for (int k = 0; k < n; ++k) {
        s->bc.amp[k] = amp[k];
        s->bc.phase[k] = phase[k];
        s->bc.omega[k] = omega[k];
    }
    s->bc.n_tc = n;
    return push_config(*s);
}

void rki_sync_host(void* handle) {
    Sim* s = resolve(handle);
    if (!s) return;
    if (s->is_initialised && s->on_device && !s->host_is_current) {
        if (copy_back(*s) == 0) s->host_is_current = true;
    }
}

void rki_sync_from_device(void* handle) {
    Sim* s = resolve(handle);
    if (!s) return;
    if (s->is_initialised && s->on_device) {
        if (!s->host_is_current && copy_back(*s) != 0) return;
        s->on_device = false;
        s->host_is_current = true;
    }
}

int rki_save_restart(void* handle, const char* filename, int len) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    if (!s->is_finalised) return -2;
    if (!filename || len < 0) return -2;
    int rc = refresh_host(*s);
    if (rc) return rc;
    const std::string path(filename, (size_t)len);
    // restart_write, rki_output.F90:740-848: interior h, hu, hv, b and time
    std::vector<rk::nc::Field2D> f = {
        {"h", "m", s->interior(s->h), s->nxt},
        {"hu", "m2 s-1", s->interior(s->hu), s->nxt},
        {"hv", "m2 s-1", s->interior(s->hv), s->nxt},
        {"b", "m", s->interior(s->b), s->nxt},
    };
    char dx[64], dy[64];
    std::snprintf(dx, sizeof dx, "%.17g", s->dx);
    std::snprintf(dy, sizeof dy, "%.17g", s->dy);
    if (!rk::nc::write_restart(path, s->nx, s->ny, f, s->t_current, dx, dy, std::to_string(s->ng))) {
        set_error("cannot write restart file '%s'", path.c_str());
        return -3;
    }
    return 0;
}

int rki_load_restart(void* handle, const char* filename, int len) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    if (s->is_finalised) return -2;
    if (!filename || len < 0) return -2;
    const std::string path(filename, (size_t)len);
    {   // classic NetCDF (restart_read, rki_output.F90:850-988)
        rk::nc::Reader r;
        if (rk::nc::load_file(path, r.buf) && r.buf.size() >= 3 && std::memcmp(r.buf.data(), "CDF", 3) == 0) {
            if (!r.parse()) {
                set_error("restart file '%s' is not a readable classic NetCDF file", path.c_str());
                return -3;
            }
            if ((int)r.dim_len("x") != s->nx || (int)r.dim_len("y") != s->ny) {
                set_error("restart file '%s' holds a %ux%u grid, the configuration %dx%d", path.c_str(),
                          r.dim_len("x"), r.dim_len("y"), s->nx, s->ny);
                return -3;
            }
            const size_t n = (size_t)s->nx * (size_t)s->ny;
            std::vector<double> buf(4 * n);
            const char* names[4] = {"h", "hu", "hv", "b"};
            double t = 0.0;
            for (int k = 0; k < 4; ++k) {
                const rk::nc::Var* v = r.find(names[k]);
                if (!v || v->dims.size() != 2 || (int)v->dims[0] != r.dim_id("y") ||
                    (int)v->dims[1] != r.dim_id("x") || !r.read_doubles(*v, n, buf.data() + k * n)) {
                    set_error("restart file '%s': variable %s missing or not FP64 (y, x)", path.c_str(), names[k]);
                    return -3;
                }
            }
            const rk::nc::Var* vt = r.find("time");
            if (!vt || !r.read_doubles(*vt, 1, &t)) {
                set_error("restart file '%s': no FP64 time", path.c_str());
                return -3;
            }
            set_interior(*s, s->h, buf.data(), s->nx, s->ny);
            set_interior(*s, s->hu, buf.data() + n, s->nx, s->ny);
            set_interior(*s, s->hv, buf.data() + 2 * n, s->nx, s->ny);
            set_interior(*s, s->b, buf.data() + 3 * n, s->nx, s->ny);
            s->t_current = t;
            s->user_set_h = true;
            s->user_set_bathymetry = true;
            return 0;
        }
    }
    // the binary format of earlier versions of this library
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        set_error("cannot open restart file '%s'", path.c_str());
        return -3;
    }
    char magic[4];
    int32_t hdr[4];
    double times[2];
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, RESTART_MAGIC, 4) == 0 &&
              std::fread(hdr, sizeof hdr, 1, f) == 1 && std::fread(times, sizeof times, 1, f) == 1 &&
              hdr[0] == RESTART_VERSION && hdr[1] == s->nx && hdr[2] == s->ny;
    if (!ok) {
        std::fclose(f);
        set_error("restart file '%s' is not a %dx%d restart of this format", path.c_str(), s->nx, s->ny);
        return -3;
    }
    const size_t n = (size_t)s->nx * (size_t)s->ny;
    std::vector<double> buf(4 * n);
    ok = std::fread(buf.data(), sizeof(double), buf.size(), f) == buf.size();
    std::fclose(f);
    if (!ok) {
        set_error("restart file '%s' is truncated", path.c_str());
        return -3;
    }
    set_interior(*s, s->h, buf.data(), s->nx, s->ny);
    set_interior(*s, s->hu, buf.data() + n, s->nx, s->ny);
    set_interior(*s, s->hv, buf.data() + 2 * n, s->nx, s->ny);
    set_interior(*s, s->b, buf.data() + 3 * n, s->nx, s->ny);
    s->t_current = times[0];
    s->user_set_h = true;
    s->user_set_bathymetry = true;
    return 0;
}

int rki_set_use_ppm_tracer(void* handle, int flag) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    s->use_ppm_tracer = flag != 0;
    return 0;
}

int rki_set_use_hllc(void* handle, int flag) {
    Sim* s = resolve(handle);
    if (!s) return -100;
    if (!s->is_initialised) return -1;
    s->use_hllc = flag != 0;
    return push_config(*s);
}

void* rki_get_device_grid(void* handle) {
    Sim* s = resolve(handle);
    return s ? s->dev : nullptr;
}

}  // extern "C"
