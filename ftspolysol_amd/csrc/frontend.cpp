// frontend.cpp -- the reference's input decks and field files on the host:
//   parameter file   INI-like sections, multi-line bracket lists, the reference's
//                    keys, aliases and defaults (param.cu:165-436, parse.cu:48-144)
//   sequence file    [polymer.i] monomers / charges / bonds blocks (polymer.cu:1-273)
//   .rf files        "m1 m2 m3" and "L1 L2 L3" headers, then one row per grid point
//                    of "re  im" pairs, one pair per field (common.h:229-471,
//                    fieldIO.cu:213-726)
//   observables.dat  the header, row and ensemble-average formats of
//                    complexLangevin.cu:13-152, 370-470
// Plain C++ (no device code); the C ABI at the end is what the CLI (fts_run.cpp),
// tests and other callers use.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fts.h"
#include "engine.h"

namespace {

using fts_host::fail;

// ---- text helpers (parse.cu:48-144) -------------------------------------------
std::string trim(const std::string& t)
{
    size_t a = 0, b = t.size();
    while (a < b && std::isspace(static_cast<unsigned char>(t[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(t[b - 1]))) --b;
    return t.substr(a, b - a);
}

// '#' starts a comment outside quotes
std::string strip_comment(const std::string& t)
{
    bool quoted = false;
    char q = 0;
    for (size_t i = 0; i < t.size(); ++i) {
        const char c = t[i];
        if ((c == '"' || c == '\'') && (i == 0 || t[i - 1] != '\\')) {
            if (!quoted) {
                quoted = true;
                q = c;
            } else if (c == q) {
                quoted = false;
            }
        }
        if (c == '#' && !quoted) return t.substr(0, i);
    }
    return t;
}

std::string unquote(std::string v)
{
    v = trim(v);
    if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') || (v.front() == '\'' && v.back() == '\''))) v = v.substr(1, v.size() - 2);
    return v;
}

bool to_bool(const std::string& v, bool& out)
{
    std::string t = unquote(v);
    for (char& c : t) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    if (t == "true" || t == "1" || t == "yes" || t == "on") return out = true, true;
    if (t == "false" || t == "0" || t == "no" || t == "off") return out = false, true;
    return false;
}

bool to_real(const std::string& v, double& out)
{
    const std::string t = unquote(v);
    char* end = nullptr;
    out = std::strtod(t.c_str(), &end);
    return end != t.c_str() && *end == '\0';
}

bool to_i64(const std::string& v, long long& out)
{
    const std::string t = unquote(v);
    char* end = nullptr;
    out = std::strtoll(t.c_str(), &end, 10);
    return end != t.c_str() && *end == '\0';
}

bool to_int(const std::string& v, int& out)
{
    long long x = 0;
    if (!to_i64(v, x)) return false;
    out = static_cast<int>(x);
    return static_cast<long long>(out) == x;
}

bool to_u64(const std::string& v, unsigned long long& out)
{
    const std::string t = unquote(v);
    char* end = nullptr;
    out = std::strtoull(t.c_str(), &end, 0);
    return end != t.c_str() && *end == '\0';
}

// net '[' minus ']' outside quotes: a value continues on the next lines while positive
int bracket_delta(const std::string& t)
{
    int d = 0;
    bool quoted = false;
    char q = 0;
    for (size_t i = 0; i < t.size(); ++i) {
        const char c = t[i];
        if ((c == '"' || c == '\'') && (i == 0 || t[i - 1] != '\\')) {
            if (!quoted) {
                quoted = true;
                q = c;
            } else if (q == c) {
                quoted = false;
            }
            continue;
        }
        if (quoted) continue;
        d += c == '[' ? 1 : c == ']' ? -1 : 0;
    }
    return d;
}

// tokens of a list value: brackets, commas and semicolons separate like blanks
std::vector<std::string> list_tokens(std::string v)
{
    for (char& c : v)
        if (c == '[' || c == ']' || c == ',' || c == ';') c = ' ';
    std::istringstream in(v);
    std::vector<std::string> out;
    std::string tok;
    while (in >> tok) out.push_back(unquote(tok));
    return out;
}

bool to_real_list(const std::string& v, std::vector<double>& out)
{
    out.clear();
    for (const std::string& t : list_tokens(v)) {
        char* end = nullptr;
        const double x = std::strtod(t.c_str(), &end);
        if (end == t.c_str() || *end != '\0' || !std::isfinite(x)) return false;
        out.push_back(x);
    }
    return !out.empty();
}

bool polymer_section(const std::string& s, int& idx)
{
    const std::string pre = "polymer.";
    if (s.size() <= pre.size() || s.compare(0, pre.size(), pre) != 0) return false;
    const std::string t = s.substr(pre.size());
    char* end = nullptr;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (end == t.c_str() || *end != '\0' || v < 0 || v > std::numeric_limits<int>::max()) return false;
    idx = static_cast<int>(v);
    return true;
}

std::string clean_name(std::string n)
{
    for (char& c : n)
        if (!std::isalnum(static_cast<unsigned char>(c)) && c != '_' && c != '-') c = '_';
    return n.empty() ? "ion" : n;
}

constexpr double UNSET = -1.0e300;

} // namespace

// ---- the deck -------------------------------------------------------------------
struct fts_deck {
    struct Species {
        std::string name;
        long long num_chains = 0;
        int N = 0;
        bool seen = false, name_set = false, N_set = false, n_set = false;
        std::vector<int> types;
        std::vector<double> charges, bonds;
    };
    // [grid]
    int m1 = 0, m2 = 0, m3 = 0;
    double a = 0.0, dL = 0.0;
    // [monomers]
    int M = 2;
    bool M_set = false, names_set = false, chi_set = false;
    std::vector<std::string> names;
    // [polymer]
    int species_count = 1;
    bool species_count_set = false;
    std::string sequence_file;
    std::vector<Species> species;
    // [interactions]
    std::vector<double> chi;
    double xi = 0.0, lB = 0.0, eigen_tol_rel = 1.0e-12;
    // [dynamics] (and the scheme / precision / seed keys this engine adds)
    double dt = 0.0, lambda_psi = 0.0;
    std::vector<double> lambda_omega;
    int Nt = 0, Nss = 0, Ns = 0, Ntmp = 0, Nso = -1, N_stat = -1, max_nan_restarts = 20;
    bool noise = true;
    int scheme = FTS_SCHEME_PC_REFERENCE, precision = FTS_PREC_FP64;
    unsigned long long seed = 0x5EEDull;
    // [ions]
    bool ions_enabled = false;
    double default_smearing = UNSET;
    bool ci_enabled = false, ci_auto = true;
    long long ci_number = 0;
    double ci_valence = 0.0, ci_smearing = UNSET;
    std::string ci_name = "counterion";
    bool salt_enabled = false, salt_combine = true;
    long long num_salt = 0;
    double cat_valence = 0.0, an_valence = 0.0, salt_smearing = UNSET;
    int cat_stoich = 1, an_stoich = 1;
    std::string cat_name = "salt_cation", an_name = "salt_anion";
    // [observables]
    bool stress = false, comp[6] = {false, false, false, false, false, false}, p_iso = false;
    // [output]
    std::map<std::string, std::string> files;

    // derived
    std::vector<fts_species_desc> sp_desc;
    fts_run_settings settings{};
    std::vector<std::string> ion_names; // merged ion species, in the engine's order (ion.cu:117-134)

    bool legacy_ab() const { return M == 2 && !M_set && !names_set && !chi_set; }
    int apply(const std::string& sec, const std::string& key, const std::string& val);
    int read_param(const char* path);
    int read_sequence(const std::string& path);
    int derive_and_validate();
};

int fts_deck::apply(const std::string& sec, const std::string& key, const std::string& v)
{
    bool ok = true, sink_b = false;
    int sink_i = 0, pidx = -1;
    double sink_r = 0.0;
    std::string sink_s;
    auto unknown = [&]() { return fail(FTS_INVALID, "unknown parameter [" + sec + "] " + key); };
    if (sec == "grid") {
        if (key == "m1") ok = to_int(v, m1);
        else if (key == "m2") ok = to_int(v, m2);
        else if (key == "m3") ok = to_int(v, m3);
        else if (key == "a") ok = to_real(v, a);
        else if (key == "dL") ok = to_real(v, dL);
        else return unknown();
    } else if (sec == "monomers") {
        if (key == "M" || key == "m" || key == "count") M_set = ok = to_int(v, M);
        else if (key == "names" || key == "types") {
            names = list_tokens(v);
            names_set = ok = !names.empty();
        } else return unknown();
    } else if (sec == "polymer") {
        if (key == "species_count") species_count_set = ok = to_int(v, species_count);
        else if (key == "sequence_file") sequence_file = unquote(v);
        else return unknown();
    } else if (polymer_section(sec, pidx)) {
        if (species.size() <= static_cast<size_t>(pidx)) species.resize(static_cast<size_t>(pidx) + 1);
        Species& s = species[static_cast<size_t>(pidx)];
        s.seen = true;
        if (key == "name") {
            s.name = unquote(v);
            s.name_set = true;
        } else if (key == "num_chains") s.n_set = ok = to_i64(v, s.num_chains);
        else if (key == "N") s.N_set = ok = to_int(v, s.N);
        else return unknown();
    } else if (sec == "interactions") {
        if (key == "chi" || key == "chi_matrix") chi_set = ok = to_real_list(v, chi);
        else if (key == "xi" || key == "v") ok = to_real(v, xi);
        else if (key == "lB") ok = to_real(v, lB);
        else if (key == "eigen_tol_rel" || key == "interaction_eigen_tol_rel") ok = to_real(v, eigen_tol_rel);
        else return unknown();
    } else if (sec == "dynamics") {
        if (key == "dt") ok = to_real(v, dt);
        else if (key == "lambda_omega" || key == "lambda_w" || key == "lambda") ok = to_real_list(v, lambda_omega);
        else if (key == "lambdaPsi") ok = to_real(v, lambda_psi);
        else if (key == "Nt") ok = to_int(v, Nt);
        else if (key == "Nss") ok = to_int(v, Nss);
        else if (key == "Ns") ok = to_int(v, Ns);
        else if (key == "Ntmp" || key == "ntmp") ok = to_int(v, Ntmp);
        else if (key == "Nso" || key == "nso") ok = to_int(v, Nso);
        else if (key == "N_stat" || key == "n_stat" || key == "Nstat" || key == "nstat") ok = to_int(v, N_stat);
        else if (key == "max_nan_restarts") ok = to_int(v, max_nan_restarts);
        else if (key == "noise") ok = to_bool(v, noise);
        else if (key == "scheme") {
            std::string t = unquote(v);
            for (char& c : t) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
            if (t == "pc" || t == "predictor_corrector") scheme = FTS_SCHEME_PC_REFERENCE;
            else if (t == "em" || t == "euler_maruyama") scheme = FTS_SCHEME_EM;
            else if (t == "si" || t == "semi_implicit") scheme = FTS_SCHEME_SEMI_IMPLICIT;
            else if (t == "si_learned" || t == "learned") scheme = FTS_SCHEME_SI_LEARNED;
            else if (t == "sipc" || t == "si_pc") scheme = FTS_SCHEME_SIPC;
            else if (t == "sipc_learned") scheme = FTS_SCHEME_SIPC_LEARNED;
            else ok = false;
        } else if (key == "precision") {
            std::string t = unquote(v);
            for (char& c : t) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
            if (t == "fp64" || t == "double") precision = FTS_PREC_FP64;
            else if (t == "fp32" || t == "single") precision = FTS_PREC_FP32;
            else ok = false;
        } else if (key == "seed") ok = to_u64(v, seed);
        else return unknown();
    } else if (sec == "ions") {
        if (key == "enabled") ok = to_bool(v, ions_enabled);
        else if (key == "default_smearing_length") ok = to_real(v, default_smearing);
        else return unknown();
    } else if (sec == "ions.counterions") {
        if (key == "enabled") ok = to_bool(v, ci_enabled);
        else if (key == "number") {
            const std::string t = unquote(v);
            if (t == "auto_neutralize") ci_auto = true;
            else {
                ok = to_i64(t, ci_number);
                ci_auto = false;
            }
        } else if (key == "valence") ok = to_real(v, ci_valence);
        else if (key == "name") ci_name = unquote(v);
        else if (key == "smearing_length") ok = to_real(v, ci_smearing);
        else return unknown();
    } else if (sec == "ions.salt") {
        if (key == "enabled") ok = to_bool(v, salt_enabled);
        else if (key == "num_salt") ok = to_i64(v, num_salt);
        else if (key == "cation_name") cat_name = unquote(v);
        else if (key == "cation_valence") ok = to_real(v, cat_valence);
        else if (key == "cation_stoich") ok = to_int(v, cat_stoich);
        else if (key == "anion_name") an_name = unquote(v);
        else if (key == "anion_valence") ok = to_real(v, an_valence);
        else if (key == "anion_stoich") ok = to_int(v, an_stoich);
        else if (key == "smearing_length") ok = to_real(v, salt_smearing);
        else if (key == "combine_identical_species") ok = to_bool(v, salt_combine);
        else return unknown();
    } else if (sec == "structure_factor") {
        // accepted and ignored, as the reference does (param.cu:271-286)
        if (key == "enabled" || key == "desmear" || key == "particle_center" || key == "skip_zero") ok = to_bool(v, sink_b);
        else if (key == "file" || key == "output_file") sink_s = unquote(v);
        else if (key == "interval") ok = to_int(v, sink_i);
        else if (key == "kmax" || key == "shell_width" || key == "bin_width" || key == "dk" || key == "shell_tolerance" ||
                 key == "k_tolerance")
            ok = to_real(v, sink_r);
        else return unknown();
    } else if (sec == "observables") {
        static const char* names6[6] = {"component_xx", "component_yy", "component_zz", "component_xy", "component_xz", "component_yz"};
        bool done = false;
        if (key == "stress") ok = to_bool(v, stress), done = true;
        for (int c = 0; c < 6 && !done; ++c)
            if (key == names6[c]) ok = to_bool(v, comp[c]), done = true;
        if (!done && (key == "isotropic_pressure" || key == "pressure" || key == "pressure_iso" || key == "p_iso" || key == "pressure_trace" ||
                      key == "p_trace"))
            ok = to_bool(v, p_iso), done = true;
        if (!done) return unknown();
    } else if (sec == "output") {
        static const std::map<std::string, std::string> alias = {
            {"omega_file", "omega_file"}, {"w_file", "omega_file"}, {"psi_file", "psi_file"}, {"tmp_omega_file", "tmp_omega_file"},
            {"tmp_w_file", "tmp_omega_file"}, {"tmp_psi_file", "tmp_psi_file"}, {"rho_type_file", "rho_type_file"},
            {"rhop_file", "rhop_file"}, {"rhoc_file", "rhoc_file"}, {"rhoc_poly_file", "rhoc_poly_file"},
            {"rhoc_ion_file", "rhoc_ion_file"}, {"ion_density_prefix", "ion_density_prefix"},
            {"step_omega_file", "step_omega_file"}, {"step_w_file", "step_omega_file"}, {"step_psi_file", "step_psi_file"},
            {"step_rho_type_file", "step_rho_type_file"}, {"step_rhop_file", "step_rhop_file"}, {"step_rhoc_file", "step_rhoc_file"},
            {"step_rhoc_poly_file", "step_rhoc_poly_file"}, {"step_rhoc_ion_file", "step_rhoc_ion_file"},
            {"step_ion_density_prefix", "step_ion_density_prefix"}, {"observable_file", "observable_file"},
            {"drift_max_file", "drift_max_file"}};
        const auto it = alias.find(key);
        if (it != alias.end()) files[it->second] = unquote(v);
        else if (key == "Nso" || key == "nso") ok = to_int(v, Nso);
        else if (key == "N_stat" || key == "n_stat" || key == "Nstat" || key == "nstat") ok = to_int(v, N_stat);
        else if (key == "order_param" || key == "order_param_file") sink_s = unquote(v);
        else if (key == "order_param_kc" || key == "order_kc") ok = to_real(v, sink_r);
        else if (key == "order_param_enabled" || key == "order_parameter_enabled") ok = to_bool(v, sink_b);
        else return unknown();
    } else {
        return unknown();
    }
    return ok ? FTS_OK : fail(FTS_INVALID, "cannot parse [" + sec + "] " + key);
}

// param::read (param.cu:165-201): sections, key = value, bracket lists over several lines
int fts_deck::read_param(const char* path)
{
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(FTS_INVALID, std::string("cannot open ") + path);
    std::string sec;
    char buf[4096];
    int lineno = 0, rc = FTS_OK;
    while (rc == FTS_OK && std::fgets(buf, sizeof(buf), fp)) {
        ++lineno;
        const std::string t = trim(strip_comment(buf));
        if (t.empty()) continue;
        if (t.front() == '[' && t.back() == ']') {
            sec = trim(t.substr(1, t.size() - 2));
            continue;
        }
        const size_t eq = t.find('=');
        if (eq == std::string::npos) {
            rc = fail(FTS_INVALID, std::string(path) + ": cannot parse line " + std::to_string(lineno));
            break;
        }
        const std::string key = trim(t.substr(0, eq));
        std::string val = trim(t.substr(eq + 1));
        int bal = bracket_delta(val);
        while (bal > 0 && std::fgets(buf, sizeof(buf), fp)) {
            ++lineno;
            const std::string extra = trim(strip_comment(buf));
            if (!extra.empty()) {
                val += " " + extra;
                bal += bracket_delta(extra);
            }
        }
        if (bal != 0) {
            rc = fail(FTS_INVALID, std::string(path) + ": unbalanced brackets while reading [" + sec + "] " + key);
            break;
        }
        rc = apply(sec, key, val);
    }
    std::fclose(fp);
    return rc;
}

// polymerSpecies::read_block_sequence_file (polymer.cu:219-273) and the value grammar of
// polymer.cu:30-111: monomer labels (or one compact word of one-letter labels), charges
// with bare signs ("+" = +1, "-" = -1) and optional bonds.
int fts_deck::read_sequence(const std::string& path)
{
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return fail(FTS_INVALID, "cannot open sequence file " + path);
    struct Block {
        bool present = false;
        std::string monomers, charges, bonds;
    };
    std::vector<Block> blocks(static_cast<size_t>(std::max(1, species_count)));
    int cur = -1, lineno = 0;
    char buf[4096];
    int rc = FTS_OK;
    while (rc == FTS_OK && std::fgets(buf, sizeof(buf), fp)) {
        ++lineno;
        const std::string t = trim(strip_comment(buf));
        if (t.empty()) continue;
        if (t.front() == '[' && t.back() == ']') {
            int idx = -1;
            if (!polymer_section(trim(t.substr(1, t.size() - 2)), idx)) rc = fail(FTS_INVALID, "unknown sequence section " + t);
            else if (idx >= species_count) rc = fail(FTS_INVALID, "sequence polymer species index");
            else blocks[static_cast<size_t>(cur = idx)].present = true;
            continue;
        }
        const size_t eq = t.find('=');
        if (cur < 0 || eq == std::string::npos) {
            rc = fail(FTS_INVALID, path + ": cannot parse line " + std::to_string(lineno));
            break;
        }
        std::string key = trim(t.substr(0, eq));
        const std::string val = trim(t.substr(eq + 1));
        for (char& c : key) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        std::string* dst = key == "monomers" ? &blocks[static_cast<size_t>(cur)].monomers
                           : key == "charges" ? &blocks[static_cast<size_t>(cur)].charges
                           : key == "bonds"   ? &blocks[static_cast<size_t>(cur)].bonds
                                              : nullptr;
        if (!dst) rc = fail(FTS_INVALID, "unknown sequence key " + key);
        else {
            if (!dst->empty()) *dst += ' ';
            *dst += val;
        }
    }
    std::fclose(fp);
    if (rc) return rc;
    auto is_sep = [](char c) { return std::isspace(static_cast<unsigned char>(c)) || c == ',' || c == ';' || c == '[' || c == ']'; };
    auto reals = [&](const std::string& text, std::vector<double>& out) {
        out.clear();
        size_t i = 0;
        while (true) {
            while (i < text.size() && is_sep(text[i])) ++i;
            if (i >= text.size()) return true;
            const char c = text[i];
            double sign = 1.0;
            if (c == '+' || c == '-') {
                sign = c == '+' ? 1.0 : -1.0;
                ++i;
                if (i >= text.size() || is_sep(text[i]) || text[i] == '+' || text[i] == '-') {
                    out.push_back(sign);
                    continue;
                }
            }
            const char* b = text.c_str() + i;
            char* e = nullptr;
            const double x = std::strtod(b, &e);
            if (e == b) return false;
            out.push_back(sign * x);
            i = static_cast<size_t>(e - text.c_str());
        }
    };
    for (int p = 0; p < species_count; ++p) {
        const Block& blk = blocks[static_cast<size_t>(p)];
        Species& s = species[static_cast<size_t>(p)];
        const std::string tag = "polymer." + std::to_string(p);
        if (!blk.present) return fail(FTS_INVALID, "missing [" + tag + "] sequence block");
        if (blk.monomers.empty()) return fail(FTS_INVALID, "missing monomers for " + tag);
        if (blk.charges.empty()) return fail(FTS_INVALID, "missing charges for " + tag);
        std::string clean = blk.monomers;
        for (char& c : clean)
            if (is_sep(c)) c = ' ';
        std::vector<std::string> toks;
        {
            std::istringstream in(clean);
            std::string tok;
            while (in >> tok) toks.push_back(unquote(tok));
        }
        bool single = true;
        for (const std::string& n : names) single = single && n.size() == 1;
        if (toks.size() == 1 && toks[0].size() > 1 && single) {
            const std::string word = toks[0];
            toks.clear();
            for (char c : word) toks.emplace_back(1, c);
        }
        s.types.clear();
        for (const std::string& tok : toks) {
            const auto it = std::find(names.begin(), names.end(), tok);
            if (it == names.end()) return fail(FTS_INVALID, "unknown monomer label in sequence: " + tok);
            s.types.push_back(static_cast<int>(it - names.begin()));
        }
        if (!reals(blk.charges, s.charges)) return fail(FTS_INVALID, "cannot parse sequence charges of " + tag);
        s.bonds.clear();
        if (!blk.bonds.empty() && !reals(blk.bonds, s.bonds)) return fail(FTS_INVALID, "cannot parse sequence bonds of " + tag);
        if (s.name.empty()) return fail(FTS_INVALID, "polymer species name is required before reading sequence");
        if (s.num_chains <= 0) return fail(FTS_INVALID, "num_chains must be positive for " + s.name);
        if (s.N <= 0) return fail(FTS_INVALID, "N must be positive for " + s.name);
        if (static_cast<int>(s.types.size()) != s.N) return fail(FTS_INVALID, "sequence length mismatch for " + s.name);
        if (static_cast<int>(s.charges.size()) != s.N) return fail(FTS_INVALID, "charge sequence length mismatch for " + s.name);
        if (!s.bonds.empty() && static_cast<int>(s.bonds.size()) != s.N - 1) return fail(FTS_INVALID, "bond sequence length must be N-1 for " + s.name);
        if (s.bonds.empty()) s.bonds.assign(static_cast<size_t>(std::max(0, s.N - 1)), 1.0);
        for (double b : s.bonds)
            if (!(b > 0.0)) return fail(FTS_INVALID, "sequence bond length");
    }
    return FTS_OK;
}

// param::derive / finalize_interactions / validate (param.cu:342-501), the parts that
// do not need the engine (modes, ions and mobilities are checked by fts_create)
int fts_deck::derive_and_validate()
{
    if (species_count > 0 && species.size() < static_cast<size_t>(species_count)) species.resize(static_cast<size_t>(species_count));
    if (Ns <= 0) Ns = Nt;
    if (default_smearing == UNSET) default_smearing = a;
    if (ci_smearing == UNSET) ci_smearing = default_smearing;
    if (salt_smearing == UNSET) salt_smearing = default_smearing;
    if (!M_set && names_set) M = static_cast<int>(names.size());
    else if (!M_set && !names_set && chi_set) {
        const int root = static_cast<int>(std::sqrt(static_cast<double>(chi.size())) + 0.5);
        if (root > 0 && static_cast<size_t>(root * root) == chi.size()) M = root;
    } else if (!M_set && !names_set) M = 2; // legacy A/B decks (param.cu:370-372)
    if (M <= 0 || M > FTS_MAX_TYPES) return fail(FTS_INVALID, "[monomers] M");
    if (names_set && names.size() != static_cast<size_t>(M)) return fail(FTS_INVALID, "[monomers] names length must equal M");
    if (!names_set) {
        names.clear();
        for (int i = 0; i < M; ++i) names.push_back(i < 26 ? std::string(1, static_cast<char>('A' + i)) : "T" + std::to_string(i));
    }
    if (!chi_set) chi.assign(static_cast<size_t>(M) * M, 0.0);
    if (chi.size() != static_cast<size_t>(M) * M) return fail(FTS_INVALID, "[interactions] chi matrix size must be M*M");
    if (!species_count_set) return fail(FTS_INVALID, "[polymer] species_count is required");
    if (species_count <= 0 || species_count > FTS_MAX_SPECIES) return fail(FTS_INVALID, "polymer species_count");
    if (species.size() > static_cast<size_t>(species_count)) return fail(FTS_INVALID, "polymer species section index exceeds species_count");
    for (int p = 0; p < species_count; ++p) {
        const Species& s = species[static_cast<size_t>(p)];
        const std::string pre = "[polymer." + std::to_string(p) + "]";
        if (!s.seen) return fail(FTS_INVALID, pre + " section is required");
        if (!s.name_set) return fail(FTS_INVALID, pre + " name is required");
        if (!s.N_set) return fail(FTS_INVALID, pre + " N is required");
        if (!s.n_set) return fail(FTS_INVALID, pre + " num_chains is required");
    }
    if (sequence_file.empty()) return fail(FTS_INVALID, "[polymer] sequence_file is required");
    if (m1 <= 0 || m2 <= 0 || m3 <= 0) return fail(FTS_INVALID, "mesh");
    if (a <= 0.0 || dL <= 0.0 || lB < 0.0) return fail(FTS_INVALID, "a, dL, or lB");
    if (dt < 0.0 || lambda_psi < 0.0) return fail(FTS_INVALID, "dt or lambdaPsi");
    if (Nss < 0 || Ns < 0 || Ntmp < 0 || Nso < -1 || N_stat < -1) return fail(FTS_INVALID, "sampling/checkpoint/step-output/drift-stat intervals");
    const bool any_comp = comp[0] || comp[1] || comp[2] || comp[3] || comp[4] || comp[5];
    if (stress && !any_comp && !p_iso) return fail(FTS_INVALID, "stress observable requires at least one component_* = true or isotropic_pressure = true");
    if (!stress && (any_comp || p_iso)) return fail(FTS_INVALID, "stress components or isotropic_pressure require [observables] stress = true");
    if (max_nan_restarts < 0) return fail(FTS_INVALID, "max_nan_restarts");
    if (xi < 0.0) return fail(FTS_INVALID, "xi");
    if (default_smearing <= 0.0) return fail(FTS_INVALID, "default_smearing_length");
    if (ci_smearing <= 0.0) return fail(FTS_INVALID, "counterion smearing_length");
    if (salt_smearing <= 0.0) return fail(FTS_INVALID, "salt smearing_length");
    static const std::pair<const char*, const char*> defaults[] = {
        {"omega_file", "omega.rf"}, {"psi_file", "psi.rf"}, {"tmp_omega_file", "tmp_omega.rf"}, {"tmp_psi_file", "tmp_psi.rf"},
        {"rho_type_file", "rho_type.rf"}, {"rhop_file", "rhop.rf"}, {"rhoc_file", "rhoc.rf"}, {"rhoc_poly_file", "rhoc_poly.rf"},
        {"rhoc_ion_file", "rhoc_ion.rf"}, {"ion_density_prefix", "rho_ion"}, {"observable_file", "observables.dat"},
        {"drift_max_file", "drift_max.dat"}};
    for (const auto& d : defaults)
        if (!files.count(d.first)) files[d.first] = d.second;
    // ion species after merging, for output labels (ion.cu:3-46, 117-134)
    ion_names.clear();
    std::vector<std::pair<double, double>> zs;
    double polymer_charge = 0.0;
    for (int p = 0; p < species_count; ++p) {
        const Species& s = species[static_cast<size_t>(p)];
        double q = 0.0;
        for (double z : s.charges) q += z;
        polymer_charge += static_cast<double>(s.num_chains) * q;
    }
    auto add = [&](const std::string& name, double z, long long count, double sm) {
        if (count == 0) return;
        if (salt_combine)
            for (const auto& e : zs)
                if (std::fabs(e.first - z) <= 1e-12 * std::max(1.0, std::max(std::fabs(e.first), std::fabs(z))) &&
                    std::fabs(e.second - sm) <= 1e-12 * std::max(1.0, std::max(std::fabs(e.second), std::fabs(sm))))
                    return;
        zs.emplace_back(z, sm);
        ion_names.push_back(clean_name(name));
    };
    if (ions_enabled && ci_enabled) {
        long long n = ci_number;
        if (ci_auto) n = (lB > 0.0 && ci_valence != 0.0) ? std::llround(-polymer_charge / ci_valence) : 0;
        add(ci_name, ci_valence, n, ci_smearing);
    }
    if (ions_enabled && salt_enabled) {
        add(cat_name, cat_valence, num_salt * cat_stoich, salt_smearing);
        add(an_name, an_valence, num_salt * an_stoich, salt_smearing);
    }
    sp_desc.assign(static_cast<size_t>(species_count), fts_species_desc{});
    for (int p = 0; p < species_count; ++p) {
        Species& s = species[static_cast<size_t>(p)];
        fts_species_desc& d = sp_desc[static_cast<size_t>(p)];
        d.num_chains = s.num_chains;
        d.N = s.N;
        d.types = s.types.data();
        d.charges = s.charges.data();
        d.bonds = s.bonds.empty() ? nullptr : s.bonds.data();
    }
    settings.Nt = Nt;
    settings.Nss = Nss;
    settings.Ns = Ns;
    settings.Ntmp = Ntmp;
    settings.Nso = Nso;
    settings.N_stat = N_stat;
    settings.max_nan_restarts = max_nan_restarts;
    settings.legacy_ab = legacy_ab() ? 1 : 0;
    settings.n_ion_names = static_cast<int>(ion_names.size());
    return FTS_OK;
}

// ---- .rf tables ------------------------------------------------------------------
namespace {

struct RfTable {
    int m[3] = {0, 0, 0};
    bool header = false;
    int nf = 0;
    std::vector<double> rows; // row-major: [row][2 nf]
};

int read_rf(const char* path, RfTable& t)
{
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(FTS_INVALID, std::string("cannot open ") + path);
    bool first = true, maybe_box = false;
    char buf[8192];
    size_t lineno = 0;
    std::vector<double> vals;
    while (std::fgets(buf, sizeof(buf), fp)) {
        ++lineno;
        char* hash = std::strchr(buf, '#');
        if (hash) *hash = '\0';
        char* p = buf;
        while (std::isspace(static_cast<unsigned char>(*p))) ++p;
        if (!*p) continue;
        if (first) {
            first = false;
            int a = 0, b = 0, c = 0;
            char extra = 0;
            if (std::sscanf(p, "%d %d %d %c", &a, &b, &c, &extra) == 3 && a > 0 && b > 0 && c > 0) {
                t.m[0] = a;
                t.m[1] = b;
                t.m[2] = c;
                t.header = true;
                maybe_box = true;
                continue;
            }
        }
        if (maybe_box) {
            maybe_box = false;
            double L1 = 0, L2 = 0, L3 = 0;
            char extra = 0;
            if (std::sscanf(p, "%lf %lf %lf %c", &L1, &L2, &L3, &extra) == 3 && L1 > 0 && L2 > 0 && L3 > 0) continue;
        }
        vals.clear();
        while (*p) {
            while (std::isspace(static_cast<unsigned char>(*p))) ++p;
            if (!*p) break;
            char* e = nullptr;
            const double x = std::strtod(p, &e);
            if (e == p) {
                std::fclose(fp);
                return fail(FTS_INVALID, std::string(path) + ": cannot parse line " + std::to_string(lineno));
            }
            vals.push_back(x);
            p = e;
        }
        if (vals.empty() || vals.size() % 2) {
            std::fclose(fp);
            return fail(FTS_INVALID, std::string(path) + ": expected an even number of numeric columns per row");
        }
        const int nf = static_cast<int>(vals.size() / 2);
        if (t.nf == 0) t.nf = nf;
        if (nf != t.nf) {
            std::fclose(fp);
            return fail(FTS_INVALID, std::string(path) + ": inconsistent number of field columns");
        }
        t.rows.insert(t.rows.end(), vals.begin(), vals.end());
    }
    std::fclose(fp);
    return FTS_OK;
}

} // namespace

extern "C" {

int fts_deck_read(const char* param_file, const char* sequence_file, fts_deck** out)
{
    if (!param_file || !out) return fail(FTS_INVALID, "deck_read: null argument");
    *out = nullptr;
    fts_deck* d = new fts_deck();
    int rc = d->read_param(param_file);
    if (!rc && sequence_file) d->sequence_file = sequence_file;
    if (!rc) rc = d->derive_and_validate();
    if (!rc) rc = d->read_sequence(d->sequence_file);
    if (!rc) rc = d->derive_and_validate(); // sequence pointers and ion labels
    if (rc) {
        delete d;
        return rc;
    }
    *out = d;
    return FTS_OK;
}

void fts_deck_destroy(fts_deck* d) { delete d; }

int fts_deck_system(fts_deck* d, fts_system* s)
{
    if (!d || !s) return fail(FTS_INVALID, "deck_system: null argument");
    std::memset(s, 0, sizeof(*s));
    s->m1 = d->m1;
    s->m2 = d->m2;
    s->m3 = d->m3;
    s->dL = d->dL;
    s->a = d->a;
    s->monomer_types = d->M;
    s->xi = d->xi;
    s->chi = d->chi.data();
    s->eigen_tol_rel = d->eigen_tol_rel;
    s->lB = d->lB;
    s->n_species = d->species_count;
    s->species = d->sp_desc.data();
    s->ions_enabled = d->ions_enabled ? 1 : 0;
    s->default_smearing_length = d->default_smearing;
    s->counterions.enabled = d->ci_enabled ? 1 : 0;
    s->counterions.auto_number = d->ci_auto ? 1 : 0;
    s->counterions.number = d->ci_number;
    s->counterions.valence = d->ci_valence;
    s->counterions.smearing_length = d->ci_smearing;
    s->salt.enabled = d->salt_enabled ? 1 : 0;
    s->salt.num_salt = d->num_salt;
    s->salt.cation_valence = d->cat_valence;
    s->salt.anion_valence = d->an_valence;
    s->salt.cation_stoich = d->cat_stoich;
    s->salt.anion_stoich = d->an_stoich;
    s->salt.smearing_length = d->salt_smearing;
    s->salt.combine_identical_species = d->salt_combine ? 1 : 0;
    s->dt = d->dt;
    s->n_lambda_omega = static_cast<int>(d->lambda_omega.size());
    s->lambda_omega = d->lambda_omega.empty() ? nullptr : d->lambda_omega.data();
    s->lambda_psi = d->lambda_psi;
    s->noise = d->noise ? 1 : 0;
    s->scheme = d->scheme;
    s->precision = d->precision;
    s->seed = d->seed;
    int mask = 0;
    if (d->stress) {
        for (int c = 0; c < 6; ++c)
            if (d->comp[c]) mask |= 1 << c;
        if (d->p_iso) mask |= FTS_STRESS_PISO;
    }
    s->stress_mask = mask;
    s->fft_schedule = FTS_SCHED_AUTO;
    s->chain_symmetry = FTS_SYMMETRY_AUTO;
    return FTS_OK;
}

int fts_deck_settings(fts_deck* d, fts_run_settings* out)
{
    if (!d || !out) return fail(FTS_INVALID, "deck_settings: null argument");
    *out = d->settings;
    return FTS_OK;
}

const char* fts_deck_string(fts_deck* d, const char* key)
{
    if (!d || !key) return nullptr;
    static thread_local std::string scratch;
    const std::string k(key);
    if (k == "sequence_file") return d->sequence_file.c_str();
    const auto it = d->files.find(k);
    if (it != d->files.end()) return it->second.c_str();
    if (k.compare(0, 8, "species.") == 0) {
        const int p = std::atoi(k.c_str() + 8);
        if (p >= 0 && p < d->species_count) return d->species[static_cast<size_t>(p)].name.c_str();
    }
    if (k.compare(0, 4, "ion.") == 0) {
        const int j = std::atoi(k.c_str() + 4);
        if (j >= 0 && j < static_cast<int>(d->ion_names.size())) return d->ion_names[static_cast<size_t>(j)].c_str();
    }
    if (k.compare(0, 8, "monomer.") == 0) {
        const int j = std::atoi(k.c_str() + 8);
        if (j >= 0 && j < d->M) return d->names[static_cast<size_t>(j)].c_str();
    }
    (void)scratch;
    return nullptr;
}

int fts_rf_read(const char* path, int* m, int* n_fields, double* data, size_t capacity_doubles)
{
    if (!path || !m || !n_fields) return fail(FTS_INVALID, "rf_read: null argument");
    RfTable t;
    int rc = read_rf(path, t);
    if (rc) return rc;
    if (!t.header) {
        t.m[0] = m[0];
        t.m[1] = m[1];
        t.m[2] = m[2];
    }
    const size_t rows = t.nf > 0 ? t.rows.size() / (2 * static_cast<size_t>(t.nf)) : 0;
    const size_t M = static_cast<size_t>(t.m[0]) * t.m[1] * t.m[2];
    if (rows != M) return fail(FTS_INVALID, std::string(path) + ": " + std::to_string(rows) + " rows, expected " + std::to_string(M));
    m[0] = t.m[0];
    m[1] = t.m[1];
    m[2] = t.m[2];
    *n_fields = t.nf;
    if (!data) return FTS_OK;
    if (capacity_doubles < 2 * M * static_cast<size_t>(t.nf)) return fail(FTS_INVALID, "rf_read: output buffer too small");
    for (int f = 0; f < t.nf; ++f)
        for (size_t i = 0; i < M; ++i) {
            data[(static_cast<size_t>(f) * M + i) * 2] = t.rows[i * 2 * t.nf + 2 * f];
            data[(static_cast<size_t>(f) * M + i) * 2 + 1] = t.rows[i * 2 * t.nf + 2 * f + 1];
        }
    return FTS_OK;
}

// write_multi_complex_fields_file (fieldIO.cu:476-509): header, then "%.10f  %.10f" per field
// Periodic trilinear interpolation of `count` fields from an n-mesh onto an m-mesh:
// target point x samples the source at x n / m (trilinear::resize, trilinear.cu:11-64).
int fts_field_resize(const double* in, const int* n, int count, double* out, const int* m)
{
    if (!in || !n || !out || !m || count < 0) return fail(FTS_INVALID, "field_resize: bad argument");
    for (int d = 0; d < 3; ++d)
        if (n[d] <= 0 || m[d] <= 0) return fail(FTS_INVALID, "field_resize: mesh");
    const size_t Ns = static_cast<size_t>(n[0]) * n[1] * n[2], Mt = static_cast<size_t>(m[0]) * m[1] * m[2];
    auto axis = [](int i, int src, int dst, int& i0, int& i1, double& t) {
        const double s = static_cast<double>(i) * src / dst;
        i0 = static_cast<int>(std::floor(s));
        i1 = src == 1 ? 0 : (i0 + 1) % src;
        t = src == 1 ? 0.0 : s - i0;
        i0 %= src;
    };
    for (int f = 0; f < count; ++f) {
        const double* src = in + 2 * Ns * static_cast<size_t>(f);
        double* dst = out + 2 * Mt * static_cast<size_t>(f);
        for (int x = 0; x < m[0]; ++x) {
            int x0, x1;
            double tx;
            axis(x, n[0], m[0], x0, x1, tx);
            for (int y = 0; y < m[1]; ++y) {
                int y0, y1;
                double ty;
                axis(y, n[1], m[1], y0, y1, ty);
                for (int z = 0; z < m[2]; ++z) {
                    int z0, z1;
                    double tz;
                    axis(z, n[2], m[2], z0, z1, tz);
                    for (int c = 0; c < 2; ++c) {
                        auto at = [&](int a, int b, int e) { return src[2 * ((static_cast<size_t>(a) * n[1] + b) * n[2] + e) + c]; };
                        auto lerp = [](double u, double v, double t) { return u + t * (v - u); };
                        const double c00 = lerp(at(x0, y0, z0), at(x0, y0, z1), tz), c01 = lerp(at(x0, y1, z0), at(x0, y1, z1), tz);
                        const double c10 = lerp(at(x1, y0, z0), at(x1, y0, z1), tz), c11 = lerp(at(x1, y1, z0), at(x1, y1, z1), tz);
                        dst[2 * ((static_cast<size_t>(x) * m[1] + y) * m[2] + z) + c] = lerp(lerp(c00, c01, ty), lerp(c10, c11, ty), tx);
                    }
                }
            }
        }
    }
    return FTS_OK;
}

int fts_rf_write(const char* path, const int* m, const double* L, int n_fields, const double* fields)
{
    if (!path || !m || !L || n_fields < 0 || (n_fields > 0 && !fields)) return fail(FTS_INVALID, "rf_write: bad argument");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(FTS_INVALID, std::string("cannot open ") + path);
    bool ok = std::fprintf(fp, "%d %d %d\n", m[0], m[1], m[2]) > 0 && std::fprintf(fp, "%.17g %.17g %.17g\n", L[0], L[1], L[2]) > 0;
    const size_t M = static_cast<size_t>(m[0]) * m[1] * m[2];
    for (size_t i = 0; ok && i < M; ++i) {
        for (int f = 0; ok && f < n_fields; ++f) {
            const double* z = fields + (static_cast<size_t>(f) * M + i) * 2;
            ok = std::fprintf(fp, "%s%.10f  %.10f", f == 0 ? "" : "  ", z[0], z[1]) >= 0;
        }
        ok = ok && std::fprintf(fp, "\n") >= 0;
    }
    if (std::fclose(fp) != 0) ok = false;
    return ok ? FTS_OK : fail(FTS_INVALID, std::string("cannot write ") + path);
}

} // extern "C"
