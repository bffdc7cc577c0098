// fts_run.cpp -- command-line driver of the engine: the counterpart of the
// reference's main program (main.cu:9-93, parse.cu:3-43) over the C ABI.
//
//   fts_run -p deck.param [--omega w.rf] [--psi psi.rf] [--seq seq.txt]
//           [-i run_id [-s parent_dir]] [--lrk params.npy] [--resume state] [--screenprint]
//
// Reads the deck and the starting fields (missing default omega.rf -> zero fields;
// meshes that differ are resized trilinearly; legacy A/B columns are mapped to the
// active modes by eigenvalue sign, fieldIO.cu:440-470), runs Nt CL steps, writes
// observables.dat (a row every Nss steps, the ensemble average over steps > Nss, at
// most Ns records: complexLangevin.cu:370-470), the Ntmp checkpoint files through
// ".new" + rename (complexLangevin.cu:583-612), the step outputs from step Nso on
// (complexLangevin.cu:523-580), and the final fields (fieldIO.cu:716-726).
// Non-finite steps restart from the device checkpoint with a new noise key; records
// taken before the checkpoint are kept (the reference truncates observables.dat).
// Every Ntmp checkpoint also writes checkpoint.state (fts_save_state: exact fields,
// step counter and noise key); --resume continues such a state to step Nt, on the
// trajectory the uninterrupted run takes.  drift_max.dat has a row per step after
// N_stat (complexLangevin.cu:466-497).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fts.h"

namespace {

bool err(const std::string& m)
{
    std::fprintf(stderr, "error: %s\n", m.c_str());
    return false;
}

bool ok_or_report(int rc, const char* what)
{
    if (rc == FTS_OK) return true;
    return err(std::string(what) + ": " + fts_last_error());
}

std::string join(const std::string& a, const std::string& b)
{
    if (a.empty() || a == ".") return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

std::string base(const std::string& p)
{
    const size_t k = p.find_last_of("/\\");
    return k == std::string::npos ? p : p.substr(k + 1);
}

bool mkdirs(const std::string& path)
{
    std::string cur;
    size_t i = 0;
    if (!path.empty() && path[0] == '/') cur = "/", i = 1;
    while (i <= path.size()) {
        const size_t nx = path.find('/', i);
        const std::string part = path.substr(i, nx == std::string::npos ? std::string::npos : nx - i);
        if (!part.empty() && part != ".") {
            if (!cur.empty() && cur.back() != '/') cur += '/';
            cur += part;
            struct stat st;
            if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0775) != 0) return err("cannot create " + cur);
        }
        if (nx == std::string::npos) break;
        i = nx + 1;
    }
    return true;
}

bool copy_file(const std::string& from, const std::string& to)
{
    FILE* in = std::fopen(from.c_str(), "rb");
    if (!in) return err("cannot open " + from);
    FILE* out = std::fopen(to.c_str(), "wb");
    if (!out) {
        std::fclose(in);
        return err("cannot open " + to);
    }
    char buf[65536];
    size_t n;
    bool ok = true;
    while ((n = std::fread(buf, 1, sizeof(buf), in)) > 0) ok = ok && std::fwrite(buf, 1, n, out) == n;
    std::fclose(in);
    return (std::fclose(out) == 0 && ok) || err("cannot write " + to);
}

// parameters of the learned kernel from an .npy file of FTS_LRK_NPARAMS little-endian float32
bool read_npy_f32(const std::string& path, std::vector<float>& out)
{
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return err("cannot open " + path);
    unsigned char head[10];
    bool ok = std::fread(head, 1, 10, fp) == 10 && std::memcmp(head, "\x93NUMPY", 6) == 0;
    size_t hlen = 0;
    if (ok && head[6] == 1) hlen = head[8] | (head[9] << 8);
    else if (ok) {
        unsigned char extra[2];
        ok = std::fread(extra, 1, 2, fp) == 2;
        hlen = head[8] | (head[9] << 8) | (static_cast<size_t>(extra[0]) << 16) | (static_cast<size_t>(extra[1]) << 24);
    }
    std::string dict(hlen, ' ');
    ok = ok && std::fread(&dict[0], 1, hlen, fp) == hlen && dict.find("'<f4'") != std::string::npos && dict.find("False") != std::string::npos;
    out.assign(FTS_LRK_NPARAMS, 0.0f);
    ok = ok && std::fread(out.data(), sizeof(float), out.size(), fp) == out.size();
    std::fclose(fp);
    return ok || err(path + ": expected an .npy array of " + std::to_string(FTS_LRK_NPARAMS) + " little-endian float32");
}

struct Out {
    FILE* fp = nullptr;
    std::vector<std::string> labels;
    std::vector<double> sum;
    long long count = 0;
};

// column labels and values in the reference's order (complexLangevin.cu:38-85)
std::vector<std::string> labels_of(fts_deck* deck, const fts_info& info, int stress_mask)
{
    std::vector<std::string> L;
    const int ns = info.n_species;
    auto poly = [&](const char* sfx) {
        for (int p = 0; p < ns; ++p) {
            if (ns == 1) {
                L.push_back(std::string("mu_poly_") + sfx);
                continue;
            }
            std::string name = fts_deck_string(deck, ("species." + std::to_string(p)).c_str());
            for (char& c : name)
                if (!std::isalnum(static_cast<unsigned char>(c)) && c != '_' && c != '-') c = '_';
            L.push_back("mu_poly" + std::to_string(p) + "_" + name + "_" + sfx);
        }
    };
    auto ion = [&](const char* sfx) {
        for (int j = 0; j < info.n_ions; ++j) {
            const char* nm = fts_deck_string(deck, ("ion." + std::to_string(j)).c_str());
            L.push_back("mu_ion" + std::to_string(j) + "_" + (nm ? nm : "ion") + "_" + sfx);
        }
    };
    poly("ex");
    poly("id");
    poly("self");
    ion("ex");
    ion("id");
    ion("self");
    L.push_back("Hmf_density_real");
    L.push_back("Hmf_density_imag");
    static const char* st[6] = {"stress_xx", "stress_yy", "stress_zz", "stress_xy", "stress_xz", "stress_yz"};
    for (int c = 0; c < 6; ++c)
        if (stress_mask & (1 << c)) L.push_back(st[c]);
    if (stress_mask & FTS_STRESS_PISO) L.push_back("pressure_iso");
    return L;
}

std::vector<double> values_of(const fts_observables& r, int stress_mask)
{
    std::vector<double> v;
    for (int p = 0; p < r.n_species; ++p) v.push_back(r.mu_ex_poly[p][0]);
    for (int p = 0; p < r.n_species; ++p) v.push_back(r.mu_id_poly[p]);
    for (int p = 0; p < r.n_species; ++p) v.push_back(r.mu_self_poly[p]);
    for (int j = 0; j < r.n_ions; ++j) v.push_back(r.mu_ex_ion[j][0]);
    for (int j = 0; j < r.n_ions; ++j) v.push_back(r.mu_id_ion[j]);
    for (int j = 0; j < r.n_ions; ++j) v.push_back(r.mu_self_ion[j]);
    v.push_back(r.H[0]);
    v.push_back(r.H[1]);
    for (int c = 0; c < 6; ++c)
        if (stress_mask & (1 << c)) v.push_back(r.stress[c][0]);
    if (stress_mask & FTS_STRESS_PISO) v.push_back(r.p_iso[0]);
    return v;
}

void write_labels(FILE* fp, const std::vector<std::string>& L, const char* first)
{
    std::fprintf(fp, "#%*s", 9, first);
    for (const std::string& l : L) std::fprintf(fp, " %*s", std::max(22, static_cast<int>(l.size())), l.c_str());
    std::fprintf(fp, "\n");
}

void write_row(FILE* fp, const std::vector<std::string>& L, const std::string& first, const std::vector<double>& v)
{
    std::fprintf(fp, "%10s", first.c_str());
    for (size_t i = 0; i < v.size(); ++i) std::fprintf(fp, " % *.*e", i < L.size() ? std::max(22, static_cast<int>(L[i].size())) : 22, 10, v[i]);
    std::fprintf(fp, "\n");
}

struct Fields {
    std::vector<double> omega, psi;
};

bool write_fields(fts_engine* e, const fts_info& info, fts_deck* deck, const int* m, const double* L, bool legacy,
                  const std::string& omega_file, const std::string& psi_file, const std::string& rho_type_file, const std::string& rhop_file,
                  const std::string& rhoc_file, const std::string& rhoc_poly_file, const std::string& rhoc_ion_file,
                  const std::string& ion_prefix)
{
    const size_t M = static_cast<size_t>(info.grid_size);
    auto get = [&](int id, int count, std::vector<double>& v) {
        v.assign(2 * M * static_cast<size_t>(std::max(0, count)), 0.0);
        return count <= 0 || ok_or_report(fts_get_field(e, id, v.data(), v.size()), "get_field");
    };
    std::vector<double> v, w;
    bool ok = true;
    if (!omega_file.empty() && (ok = get(FTS_FIELD_OMEGA, info.n_omega, v))) {
        if (legacy) { // active modes -> legacy A/B columns (fieldIO.cu:459-474)
            w.assign(4 * M, 0.0);
            for (int l = 0; l < info.n_omega; ++l) {
                const int col = info.mode_eigenvalue[l] < 0.0 ? 0 : 1;
                std::copy(v.begin() + static_cast<long>(2 * M * l), v.begin() + static_cast<long>(2 * M * (l + 1)), w.begin() + static_cast<long>(2 * M * col));
            }
            ok = ok_or_report(fts_rf_write(omega_file.c_str(), m, L, 2, w.data()), "write omega");
        } else {
            ok = ok_or_report(fts_rf_write(omega_file.c_str(), m, L, info.n_omega, v.data()), "write omega");
        }
    }
    auto one = [&](const std::string& f, int id, int count) {
        if (f.empty() || !ok) return;
        ok = get(id, count, v) && ok_or_report(fts_rf_write(f.c_str(), m, L, count, v.data()), "write field");
    };
    one(psi_file, FTS_FIELD_PSI, 1);
    one(rho_type_file, FTS_FIELD_RHO_TYPE, info.monomer_types);
    one(rhop_file, FTS_FIELD_RHO_P, 1);
    one(rhoc_file, FTS_FIELD_RHO_C, 1);
    one(rhoc_poly_file, FTS_FIELD_RHO_C_POLY, 1);
    one(rhoc_ion_file, FTS_FIELD_RHO_C_ION, 1);
    if (ok && !ion_prefix.empty() && info.n_ions > 0 && (ok = get(FTS_FIELD_ION_RHO, info.n_ions, v)))
        for (int j = 0; ok && j < info.n_ions; ++j) {
            const char* nm = fts_deck_string(deck, ("ion." + std::to_string(j)).c_str());
            const std::string f = ion_prefix + "_" + std::to_string(j) + "_" + (nm ? nm : "ion") + ".rf";
            ok = ok_or_report(fts_rf_write(f.c_str(), m, L, 1, v.data() + 2 * M * static_cast<size_t>(j)), "write ion density");
        }
    return ok;
}

std::string step_name(const std::string& pattern, int index)
{
    const size_t k = pattern.find("%d");
    return k == std::string::npos ? pattern : pattern.substr(0, k) + std::to_string(index) + pattern.substr(k + 2);
}

// a step template "dir/name_%d.rf" writes into dir/name/ (fieldIO.cu:207-230)
std::string step_template(const std::string& pattern)
{
    if (pattern.empty() || pattern.find("%d") == std::string::npos) return pattern;
    const size_t k = pattern.find_last_of('/');
    const std::string parent = k == std::string::npos ? "" : pattern.substr(0, k), b = base(pattern);
    std::string folder = b.substr(0, b.find("%d"));
    while (!folder.empty() && (folder.back() == '_' || folder.back() == '-' || folder.back() == '.')) folder.pop_back();
    if (folder.empty()) folder = "step_fields";
    const std::string full = (!parent.empty() && base(parent) == folder) ? parent : join(parent, folder);
    mkdirs(full);
    return join(full, b);
}

} // namespace

int main(int argc, char** argv)
{
    const auto t_start = std::chrono::steady_clock::now();
    std::string param_file, omega_in, psi_in, seq, run_id, run_parent, lrk_file, resume;
    bool omega_override = false, psi_override = false, screenprint = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](std::string& dst) {
            if (++i == argc) return err("missing value for " + a);
            dst = argv[i];
            return true;
        };
        bool ok = true;
        if (a == "-p") ok = next(param_file);
        else if (a == "--omega") ok = next(omega_in), omega_override = true;
        else if (a == "-psi" || a == "--psi") ok = next(psi_in), psi_override = true;
        else if (a == "-i" || a == "--id") ok = next(run_id);
        else if (a == "-s" || a == "--dir" || a == "--run-dir") ok = next(run_parent);
        else if (a == "--seq") ok = next(seq);
        else if (a == "--lrk") ok = next(lrk_file);
        else if (a == "--resume") ok = next(resume);
        else if (a == "--screenprint") screenprint = true;
        else if (a == "-h" || a == "--help") {
            std::printf("usage: %s -p param_file [--omega file] [--psi file] [--seq file] [-i run_id [-s dir]] [--lrk params.npy] [--resume state] [--screenprint]\n", argv[0]);
            return EXIT_SUCCESS;
        } else ok = err("unknown option " + a);
        if (!ok) return EXIT_FAILURE;
    }
    if (param_file.empty()) return err("-p param_file is required"), EXIT_FAILURE;
    if (!run_parent.empty() && run_id.empty()) return err("-s requires -i"), EXIT_FAILURE;

    fts_deck* deck = nullptr;
    if (!ok_or_report(fts_deck_read(param_file.c_str(), seq.empty() ? nullptr : seq.c_str(), &deck), "deck")) return EXIT_FAILURE;
    fts_run_settings rs;
    fts_deck_settings(deck, &rs);
    auto file = [&](const char* k) { return std::string(fts_deck_string(deck, k)); };
    std::string run_dir;
    if (!run_id.empty()) {
        if (run_id == "." || run_id == ".." || run_id.find('/') != std::string::npos) return err("invalid run id " + run_id), EXIT_FAILURE;
        run_dir = join(run_parent, run_id);
        if (!mkdirs(run_dir) || !copy_file(param_file, join(run_dir, run_id + ".param")) ||
            !copy_file(file("sequence_file"), join(run_dir, run_id + ".sequence")))
            return EXIT_FAILURE;
    }
    auto out_path = [&](const std::string& f) { return run_dir.empty() ? f : join(run_dir, base(f)); };
    if (!omega_override) omega_in = file("omega_file");
    if (!psi_override) psi_in = file("psi_file");

    fts_system sys;
    fts_deck_system(deck, &sys);
    fts_engine* e = nullptr;
    if (!ok_or_report(fts_create(&sys, &e), "engine")) return EXIT_FAILURE;
    fts_info info;
    fts_get_info(e, &info);
    const int m[3] = {sys.m1, sys.m2, sys.m3};
    const double L[3] = {sys.dL * sys.m1, sys.dL * sys.m2, sys.dL * sys.m3};
    const size_t M = static_cast<size_t>(info.grid_size);
    if (!lrk_file.empty()) {
        std::vector<float> params;
        if (!read_npy_f32(lrk_file, params) || !ok_or_report(fts_lrk_set_network(e, params.data(), 1e-3), "learned kernel")) return EXIT_FAILURE;
    }

    // starting fields (fieldIO.cu:515-612)
    auto read_table = [&](const std::string& path, int want_fields, std::vector<double>& out, int& got_fields) {
        int fm[3] = {m[0], m[1], m[2]};
        if (!ok_or_report(fts_rf_read(path.c_str(), fm, &got_fields, nullptr, 0), "read fields")) return false;
        if (want_fields > 0 && got_fields != want_fields) return err(path + ": expected " + std::to_string(2 * want_fields) + " numeric columns per row");
        const size_t Mf = static_cast<size_t>(fm[0]) * fm[1] * fm[2];
        std::vector<double> raw(2 * Mf * static_cast<size_t>(got_fields));
        if (!ok_or_report(fts_rf_read(path.c_str(), fm, &got_fields, raw.data(), raw.size()), "read fields")) return false;
        out.assign(2 * M * static_cast<size_t>(got_fields), 0.0);
        if (fm[0] == m[0] && fm[1] == m[1] && fm[2] == m[2]) out = raw;
        else {
            if (screenprint) std::printf("resizing %s from %d %d %d to %d %d %d\n", path.c_str(), fm[0], fm[1], fm[2], m[0], m[1], m[2]);
            if (!ok_or_report(fts_field_resize(raw.data(), fm, got_fields, out.data(), m), "resize")) return false;
        }
        return true;
    };
    std::vector<double> omega(2 * M * static_cast<size_t>(info.n_omega), 0.0), psi(2 * M, 0.0);
    struct stat st_probe;
    if (info.n_omega > 0) {
        if (stat(omega_in.c_str(), &st_probe) != 0 && !omega_override && omega_in == "omega.rf") {
            if (screenprint) std::printf("%s not found; initializing omega fields to zero.\n", omega_in.c_str());
        } else {
            std::vector<double> t;
            int nf = 0;
            if (!read_table(omega_in, rs.legacy_ab ? 0 : info.n_omega, t, nf)) return EXIT_FAILURE;
            if (rs.legacy_ab) { // column 0: composition mode (d < 0), column 1: density mode (d > 0)
                if (nf != 1 && nf != 2) return err("legacy A/B omega file must contain one or two complex fields"), EXIT_FAILURE;
                for (int l = 0; l < info.n_omega; ++l) {
                    const int col = info.mode_eigenvalue[l] < 0.0 ? 0 : 1;
                    if (col < nf) std::copy(t.begin() + static_cast<long>(2 * M * col), t.begin() + static_cast<long>(2 * M * (col + 1)), omega.begin() + static_cast<long>(2 * M * l));
                }
            } else {
                omega = t;
            }
        }
    }
    if (sys.lB > 0.0) {
        if (stat(psi_in.c_str(), &st_probe) != 0) {
            if (screenprint) std::printf("%s not found; initializing Psi to zero.\n", psi_in.c_str());
        } else {
            int nf = 0;
            if (!read_table(psi_in, 1, psi, nf)) return EXIT_FAILURE;
        }
    }
    for (double x : omega)
        if (!std::isfinite(x)) return err("input omega has non-finite values"), EXIT_FAILURE;
    for (double x : psi)
        if (!std::isfinite(x)) return err("input psi has non-finite values"), EXIT_FAILURE;
    if (!ok_or_report(fts_set_fields(e, omega.data(), psi.data()), "set fields")) return EXIT_FAILURE;
    if (rs.Ntmp > 0 && !ok_or_report(fts_set_restart(e, rs.Ntmp, rs.max_nan_restarts), "restart")) return EXIT_FAILURE;
    if (!resume.empty()) { // after fts_set_restart, which zeroes the restart count the state carries
        if (!ok_or_report(fts_load_state(e, resume.c_str()), "resume")) return EXIT_FAILURE;
        fts_get_info(e, &info);
        if (screenprint) std::printf("resuming %s at step %lld\n", resume.c_str(), info.step);
    }
    // drift_max.dat (complexLangevin.cu:381-408, 466-516): rows for the steps after N_stat
    FILE* drift_fp = nullptr;
    std::vector<std::string> drift_labels;
    for (int a = 0; a < info.n_omega; ++a) drift_labels.push_back("u_omega" + std::to_string(a) + "_max");
    drift_labels.push_back("u_psi_max");
    const std::string drift_file = rs.N_stat >= 0 ? out_path(file("drift_max_file")) : std::string();
    auto open_drift = [&]() {
        if (drift_fp) std::fclose(drift_fp);
        drift_fp = std::fopen(drift_file.c_str(), "w");
        if (!drift_fp) return err("cannot open " + drift_file);
        write_labels(drift_fp, drift_labels, "step");
        std::fflush(drift_fp);
        return true;
    };
    if (!drift_file.empty()) {
        if (!open_drift() || !ok_or_report(fts_set_drift_stats(e, 1), "drift statistics")) return EXIT_FAILURE;
    }

    // observables.dat (complexLangevin.cu:370-470)
    Out obs;
    obs.labels = labels_of(deck, info, sys.stress_mask);
    obs.sum.assign(obs.labels.size(), 0.0);
    const std::string obs_file = out_path(file("observable_file"));
    if (!obs_file.empty()) {
        obs.fp = std::fopen(obs_file.c_str(), "w");
        if (!obs.fp) return err("cannot open " + obs_file), EXIT_FAILURE;
        std::fprintf(obs.fp, "# observables Nss=%d Ns=%d\n", rs.Nss, rs.Ns);
        write_labels(obs.fp, obs.labels, "step");
        std::fflush(obs.fp);
    }
    const std::string tmp_omega = out_path(file("tmp_omega_file")), tmp_psi = out_path(file("tmp_psi_file"));
    const std::string state_file = out_path("checkpoint.state");
    auto checkpoint_files = [&]() {
        std::vector<double> w(2 * M * static_cast<size_t>(info.n_omega)), p(2 * M);
        if (info.n_omega > 0 && !ok_or_report(fts_get_field(e, FTS_FIELD_OMEGA, w.data(), w.size()), "checkpoint")) return false;
        if (!ok_or_report(fts_get_field(e, FTS_FIELD_PSI, p.data(), p.size()), "checkpoint")) return false;
        const std::string nw = tmp_omega + ".new", np = tmp_psi + ".new";
        if (!ok_or_report(fts_rf_write(nw.c_str(), m, L, info.n_omega, w.data()), "checkpoint") ||
            !ok_or_report(fts_rf_write(np.c_str(), m, L, 1, p.data()), "checkpoint"))
            return false;
        if (std::rename(nw.c_str(), tmp_omega.c_str()) != 0 || std::rename(np.c_str(), tmp_psi.c_str()) != 0) return err("cannot replace checkpoint files");
        return ok_or_report(fts_save_state(e, state_file.c_str()), "checkpoint state");
    };
    if (rs.Ntmp > 0 && !checkpoint_files()) return EXIT_FAILURE;
    const char* step_keys[8] = {"step_omega_file", "step_psi_file", "step_rho_type_file", "step_rhop_file",
                                "step_rhoc_file", "step_rhoc_poly_file", "step_rhoc_ion_file", "step_ion_density_prefix"};
    std::string step_tpl[8];
    bool step_out = false;
    for (int k = 0; k < 8; ++k) {
        const char* v = fts_deck_string(deck, step_keys[k]);
        step_tpl[k] = (v && rs.Nso >= 0) ? step_template(out_path(v)) : std::string();
        step_out = step_out || !step_tpl[k].empty();
    }
    const int first_out = rs.Nso > 1 ? rs.Nso : 1;

    const int record_every = std::max(1, rs.Nss);
    std::vector<fts_observables> recs;
    long long step = info.step; // 0, or the resumed state's step
    int restarts_seen = info.restarts;
    while (step < rs.Nt) {
        long long n = std::min<long long>(rs.Nt - step, 1000);
        if (rs.Ntmp > 0) n = std::min<long long>(n, rs.Ntmp - step % rs.Ntmp);
        if (step_out && step + 1 >= first_out) n = 1;
        recs.assign(static_cast<size_t>(n / record_every + 1), fts_observables{});
        int nrec = 0;
        const int rc = fts_step(e, static_cast<int>(n), record_every, recs.data(), static_cast<int>(recs.size()), &nrec);
        fts_get_info(e, &info);
        const bool restarted = info.restarts != restarts_seen;
        if (restarted && obs.fp)
            std::fprintf(obs.fp, "# restarted after a non-finite value from the checkpoint (restart %d); records after the checkpoint discarded\n", info.restarts);
        if (drift_fp) {
            if (restarted) { // the reference starts the file again (reset_drift_statistics_after_restart)
                if (!open_drift()) return EXIT_FAILURE;
                std::fprintf(drift_fp, "# restarted after non-finite value at step %lld from checkpoint step %lld; previous drift statistics discarded\n",
                             info.failed_step, info.checkpoint_step);
            }
            const size_t w = drift_labels.size() + 1; // step, then the values
            const int cap = static_cast<int>(n) + 1;
            std::vector<double> rows(static_cast<size_t>(cap) * w);
            int nr = 0;
            if (!ok_or_report(fts_get_drift_rows(e, rows.data(), cap, &nr), "drift statistics")) return EXIT_FAILURE;
            for (int r = 0; r < nr; ++r) {
                const double* row = rows.data() + static_cast<size_t>(r) * w;
                const long long st = static_cast<long long>(row[0]);
                if (st <= rs.N_stat || (restarted && st <= info.checkpoint_step)) continue;
                for (size_t c = 1; c < w; ++c)
                    if (!std::isfinite(row[c])) return err("non-finite " + drift_labels[c - 1] + " at step " + std::to_string(st)), EXIT_FAILURE;
                write_row(drift_fp, drift_labels, std::to_string(st), std::vector<double>(row + 1, row + w));
            }
            std::fflush(drift_fp);
        }
        restarts_seen = info.restarts;
        for (int k = 0; k < nrec; ++k) {
            const std::vector<double> v = values_of(recs[static_cast<size_t>(k)], sys.stress_mask);
            if (obs.fp) write_row(obs.fp, obs.labels, std::to_string(recs[static_cast<size_t>(k)].step), v);
            if (screenprint) {
                for (double x : v) std::printf("% 13.4e", x);
                std::printf("\n");
            }
            if (recs[static_cast<size_t>(k)].step > rs.Nss && obs.count < rs.Ns) {
                for (size_t c = 0; c < v.size(); ++c) obs.sum[c] += v[c];
                ++obs.count;
            }
        }
        if (obs.fp) std::fflush(obs.fp);
        if (rc != FTS_OK) return ok_or_report(rc, "step"), EXIT_FAILURE;
        step += n;
        if (rs.Ntmp > 0 && step % rs.Ntmp == 0 && !checkpoint_files()) return EXIT_FAILURE;
        if (step_out && step >= first_out) {
            const int idx = static_cast<int>(step - first_out + 1);
            if ((!step_tpl[2].empty() || !step_tpl[3].empty() || !step_tpl[4].empty() || !step_tpl[5].empty() || !step_tpl[6].empty() ||
                 !step_tpl[7].empty()) &&
                !ok_or_report(fts_compute_densities(e, nullptr), "step output densities"))
                return EXIT_FAILURE;
            auto nm = [&](int k) { return step_tpl[k].empty() ? std::string() : step_name(step_tpl[k], idx); };
            if (!write_fields(e, info, deck, m, L, rs.legacy_ab != 0, nm(0), nm(1), nm(2), nm(3), nm(4), nm(5), nm(6), nm(7))) return EXIT_FAILURE;
        }
    }
    if (drift_fp) std::fclose(drift_fp);
    if (obs.fp) { // ensemble average (complexLangevin.cu:443-470)
        std::fprintf(obs.fp, "\n# ensemble_average Nss=%d Ns=%d count=%lld\n", rs.Nss, rs.Ns, obs.count);
        if (obs.count == 0) std::fprintf(obs.fp, "# no_observables\n");
        else {
            std::vector<double> avg(obs.sum.size());
            for (size_t c = 0; c < avg.size(); ++c) avg[c] = obs.sum[c] / static_cast<double>(obs.count);
            write_labels(obs.fp, obs.labels, "stat");
            write_row(obs.fp, obs.labels, "average", avg);
        }
        std::fclose(obs.fp);
    }
    // final densities and fields (main.cu:72-88)
    if (!ok_or_report(fts_compute_densities(e, nullptr), "final densities")) return EXIT_FAILURE;
    const bool ok = write_fields(e, info, deck, m, L, rs.legacy_ab != 0, out_path(file("omega_file")), out_path(file("psi_file")),
                                 out_path(file("rho_type_file")), out_path(file("rhop_file")), out_path(file("rhoc_file")),
                                 out_path(file("rhoc_poly_file")), out_path(file("rhoc_ion_file")), out_path(file("ion_density_prefix")));
    fts_destroy(e);
    fts_deck_destroy(deck);
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::printf("total time = % .4e s\n", seconds);
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}
