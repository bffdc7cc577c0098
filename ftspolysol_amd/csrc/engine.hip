// engine.hip -- host engine and C ABI of the MI355X complex-Langevin engine.
//
// One fts_engine owns every device buffer (the role of deviceState,
// deviceState.cu:8-105) and one HIP stream (default priority).  The hot path
// per species and state is
//   smoothing           3-pass FFT applies (plane FWD, pencil x K/M, plane INV)
//   Boltzmann factors   one e^{-W} field per distinct (type, charge) bead class
//   two sweeps          forward and backward propagators advance together:
//                       pencil(FWD, x phi/M, INV) + plane(INV, x e^{-W}, store q, FWD)
//                       per bead step, both sweeps in each launch
//   Q                   FP64 reduction of qF_{N-1}
//   contraction         one pass over the stored qF, qB -> unsmoothed densities
// followed by Gaussian smoothing, ions, drifts and the fused CL update kernels.
// Reductions stay on the device; the host reads a scalar only where it needs
// one (finiteness checks, records).
#include <mutex>

#include "engine.h"
#include "kernels.h"

void fts_engine::release()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& pr : pending) {
        event_pool.push_back(pr.first);
        event_pool.push_back(pr.second);
    }
    pending.clear();
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    event_pool.clear();
    for (DevBuf<double>* b : {&kvec, &sep_gamma, &sep_bond, &sep_ion, &si_table_omega, &si_table_psi, &norms, &logc, &rfac, &rfac2}) b->release();
    tw64_x.release();
    tw64_yz.release();
    tw32_x.release();
    tw32_yz.release();
    tw64_z.release();
    tw32_z.release();
    for (DevBuf<double2>* b : {&omega, &omega_bar, &omega_s, &drift, &psi, &psi_bar, &psi_s, &dpsi, &dpsi_bar, &W_type_s, &rho_hat,
                               &rho_type, &rho_c_hat, &rho_c, &rho_c_poly, &rho_c_ion, &rho_p, &ion_psi, &ion_b, &ion_rho, &t1, &t2,
                               &t3, &psi_hat, &mode_coef, &partials, &scal, &sipc_d, &sipc_dpsi})
        b->release();
    flag.release();
    ckpt_omega.release();
    ckpt_psi.release();
    stress_acc.release();
    if (pin) (void)hipHostFree(pin);
    pin = nullptr;
    nkey.release();
    if (hkey) (void)hipHostFree(hkey);
    hkey = nullptr;
    for (HostSpecies& sp : species) {
        sp.d_cls_type.release();
        sp.d_cls_start.release();
        sp.d_cls_beads.release();
        sp.d_cls_z.release();
        sp.d_cls_shift.release();
        sp.d_half_start.release();
        sp.d_half_beads.release();
    }
    for (void** p : {&qF, &qB, &E, &tmpF, &tmpB, &kA, &kB}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    for (DevBuf<double>* b : {&si_A_omega, &si_A_psi, &lrk_ad_perm, &lrk_chunk_sums, &lrk_ring}) b->release();
    for (DevBuf<int>* b : {&lrk_perm, &lrk_chunk_start, &lrk_bin_chunk}) b->release();
    lrk_bin.release();
    lrk_u.release();
    for (auto& f : lrk_hat)
        for (auto& k : f)
            for (auto& b : k) b.release();
    lrk_params.release();
    lrk_state.release();
    if (stream && own_stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    stream = nullptr;
}

// ---------------------------------------------------------------------------
// Setup: interaction modes (interactionMode.cu:9-298), ions (ion.cu:3-46, 117-134)
// ---------------------------------------------------------------------------
namespace {

double frob(const std::vector<double>& x)
{
    long double s = 0.0L;
    for (double v : x) s += static_cast<long double>(v) * v;
    return std::sqrt(static_cast<double>(s));
}

bool jacobi_eigen(const std::vector<double>& in, int n, std::vector<double>& ev, std::vector<double>& vecs)
{
    std::vector<double> A = in;
    vecs.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) vecs[static_cast<size_t>(i) * n + i] = 1.0;
    ev.assign(static_cast<size_t>(n), 0.0);
    if (n == 1) {
        ev[0] = A[0];
        return true;
    }
    auto off = [&]() {
        long double s = 0.0L;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j) s += static_cast<long double>(A[static_cast<size_t>(i) * n + j]) * A[static_cast<size_t>(i) * n + j];
        return std::sqrt(static_cast<double>(s));
    };
    const double scale = std::max(1.0, frob(in));
    const int max_sweeps = std::max(50, 100 * n * n);
    bool converged = false;
    for (int sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
        if (off() <= 1.0e-14 * scale) {
            converged = true;
            break;
        }
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[static_cast<size_t>(p) * n + q];
                if (std::fabs(apq) <= 2.220446049250313e-16 * scale) continue;
                const double app = A[static_cast<size_t>(p) * n + p], aqq = A[static_cast<size_t>(q) * n + q];
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau)) : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = A[static_cast<size_t>(k) * n + p], akq = A[static_cast<size_t>(k) * n + q];
                    const double nkp = c * akp - s * akq, nkq = s * akp + c * akq;
                    A[static_cast<size_t>(k) * n + p] = A[static_cast<size_t>(p) * n + k] = nkp;
                    A[static_cast<size_t>(k) * n + q] = A[static_cast<size_t>(q) * n + k] = nkq;
                }
                A[static_cast<size_t>(p) * n + p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
                A[static_cast<size_t>(q) * n + q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
                A[static_cast<size_t>(p) * n + q] = A[static_cast<size_t>(q) * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double vkp = vecs[static_cast<size_t>(k) * n + p], vkq = vecs[static_cast<size_t>(k) * n + q];
                    vecs[static_cast<size_t>(k) * n + p] = c * vkp - s * vkq;
                    vecs[static_cast<size_t>(k) * n + q] = s * vkp + c * vkq;
                }
                rotated = true;
            }
        if (!rotated) converged = true;
    }
    if (!converged && off() > 1.0e-10 * scale) return false;
    for (int i = 0; i < n; ++i) ev[static_cast<size_t>(i)] = A[static_cast<size_t>(i) * n + i];
    return true;
}

inline int spectral_mode(int i, int m) { return i > m / 2 ? i - m : i; }
inline double wave_number(int i, int m, double L) { return 2.0 * Pi * spectral_mode(i, m) / L; }

} // namespace

int fts_engine::build_modes(const fts_system* s)
{
    nT = s->monomer_types;
    if (nT <= 0 || nT > FTS_MAX_TYPES) return fail(FTS_INVALID, "monomer_types out of range");
    xi = s->xi;
    if (!std::isfinite(xi) || xi < 0.0) return fail(FTS_INVALID, "xi");
    const double tol_rel = s->eigen_tol_rel < 0.0 ? 1.0e-12 : s->eigen_tol_rel;
    std::vector<double> chi(static_cast<size_t>(nT) * nT, 0.0);
    if (s->chi) chi.assign(s->chi, s->chi + static_cast<size_t>(nT) * nT);
    double max_chi = 0.0;
    for (double x : chi) {
        if (!std::isfinite(x)) return fail(FTS_INVALID, "chi matrix entry");
        max_chi = std::max(max_chi, std::fabs(x));
    }
    const double sym_tol = 1.0e-12 * std::max(1.0, max_chi);
    for (int al = 0; al < nT; ++al) {
        if (std::fabs(chi[static_cast<size_t>(al) * nT + al]) > sym_tol) return fail(FTS_INVALID, "chi diagonal entries must be zero");
        chi[static_cast<size_t>(al) * nT + al] = 0.0;
        for (int be = al + 1; be < nT; ++be) {
            double& ab = chi[static_cast<size_t>(al) * nT + be];
            double& ba = chi[static_cast<size_t>(be) * nT + al];
            if (std::fabs(ab - ba) > sym_tol) return fail(FTS_INVALID, "chi matrix must be symmetric");
            ab = ba = 0.5 * (ab + ba);
        }
    }
    K.assign(static_cast<size_t>(nT) * nT, 0.0);
    for (int al = 0; al < nT; ++al)
        for (int be = 0; be < nT; ++be) K[static_cast<size_t>(al) * nT + be] = al == be ? xi : xi + chi[static_cast<size_t>(al) * nT + be];
    std::vector<double> rev, rvec;
    if (!jacobi_eigen(K, nT, rev, rvec)) return fail(FTS_INVALID, "Jacobi diagonalization of interaction matrix did not converge");
    std::vector<int> order(static_cast<size_t>(nT));
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        if (rev[static_cast<size_t>(x)] != rev[static_cast<size_t>(y)]) return rev[static_cast<size_t>(x)] > rev[static_cast<size_t>(y)];
        return x < y;
    });
    eig.assign(static_cast<size_t>(nT), 0.0);
    O.assign(static_cast<size_t>(nT) * nT, 0.0);
    for (int l = 0; l < nT; ++l) {
        const int raw = order[static_cast<size_t>(l)];
        eig[static_cast<size_t>(l)] = rev[static_cast<size_t>(raw)];
        double* row = &O[static_cast<size_t>(l) * nT];
        for (int al = 0; al < nT; ++al) row[al] = rvec[static_cast<size_t>(al) * nT + raw];
        int piv = 0;
        for (int al = 1; al < nT; ++al)
            if (std::fabs(row[al]) > std::fabs(row[piv])) piv = al;
        if (row[piv] < 0.0)
            for (int al = 0; al < nT; ++al) row[al] *= -1.0;
    }
    double max_abs = 0.0;
    for (double d : eig) max_abs = std::max(max_abs, std::fabs(d));
    const double tol_abs = tol_rel * std::max(1.0, max_abs);
    active.clear();
    for (int l = 0; l < nT; ++l)
        if (std::fabs(eig[static_cast<size_t>(l)]) > tol_abs) active.push_back(l);
    n_omega = static_cast<int>(active.size());
    std::vector<double> res(static_cast<size_t>(nT) * nT), orth(static_cast<size_t>(nT) * nT);
    for (int al = 0; al < nT; ++al)
        for (int be = 0; be < nT; ++be) {
            double v = 0.0, d = 0.0;
            for (int l = 0; l < nT; ++l) {
                v += O[static_cast<size_t>(l) * nT + al] * eig[static_cast<size_t>(l)] * O[static_cast<size_t>(l) * nT + be];
                d += O[static_cast<size_t>(al) * nT + l] * O[static_cast<size_t>(be) * nT + l];
            }
            res[static_cast<size_t>(al) * nT + be] = K[static_cast<size_t>(al) * nT + be] - v;
            orth[static_cast<size_t>(al) * nT + be] = d - (al == be ? 1.0 : 0.0);
        }
    if (frob(res) / std::max(1.0, frob(K)) > 1.0e-10) return fail(FTS_INVALID, "interaction eigenmode reconstruction check failed");
    if (frob(orth) > 1.0e-10) return fail(FTS_INVALID, "interaction eigenvector orthogonality check failed");
    return FTS_OK;
}

int fts_engine::build_ions(const fts_system* s)
{
    ions.clear();
    double polymer_charge = 0.0;
    for (const HostSpecies& sp : species) polymer_charge += static_cast<double>(sp.n) * sp.net_charge();
    const double def_a = s->default_smearing_length > 0.0 ? s->default_smearing_length : a;
    const bool combine = s->salt.combine_identical_species != 0;
    auto append = [&](double z, long long n, double aj) {
        if (n == 0) return;
        if (combine)
            for (HostIon& ion : ions)
                if (std::fabs(ion.z - z) <= 1.0e-12 && std::fabs(ion.a - aj) <= 1.0e-12) {
                    ion.n += n;
                    return;
                }
        ions.push_back(HostIon{z, aj, n});
    };
    if (s->ions_enabled && s->counterions.enabled) {
        const double z = s->counterions.valence;
        const double aj = s->counterions.smearing_length > 0.0 ? s->counterions.smearing_length : def_a;
        if (electro() && z == 0.0) return fail(FTS_INVALID, "counterion valence");
        long long n = s->counterions.number;
        if (s->counterions.auto_number) {
            if (electro()) {
                const double need = -polymer_charge / z;
                if (need < -1.0e-10) return fail(FTS_INVALID, "counterion sign");
                const double r = std::round(need);
                if (std::fabs(need - r) > 1.0e-8 * std::max(1.0, std::fabs(need))) return fail(FTS_INVALID, "counterion number");
                n = static_cast<long long>(r);
            } else {
                n = 0;
            }
        }
        if (n < 0 || aj <= 0.0) return fail(FTS_INVALID, "counterion");
        append(z, n, aj);
    }
    if (s->ions_enabled && s->salt.enabled) {
        const fts_salt_desc& sl = s->salt;
        const double aj = sl.smearing_length > 0.0 ? sl.smearing_length : def_a;
        if (sl.num_salt < 0 || sl.cation_stoich < 0 || sl.anion_stoich < 0) return fail(FTS_INVALID, "salt counts");
        if (electro()) {
            if (sl.cation_valence == 0.0 || sl.anion_valence == 0.0) return fail(FTS_INVALID, "salt");
            if (std::fabs(sl.cation_stoich * sl.cation_valence + sl.anion_stoich * sl.anion_valence) > 1.0e-10)
                return fail(FTS_INVALID, "salt is not neutral");
        }
        append(sl.cation_valence, sl.num_salt * static_cast<long long>(sl.cation_stoich), aj);
        append(sl.anion_valence, sl.num_salt * static_cast<long long>(sl.anion_stoich), aj);
    }
    double total = polymer_charge, scale = std::fabs(polymer_charge);
    for (const HostIon& ion : ions) {
        if (ion.n < 0 || (electro() && ion.z == 0.0) || ion.a <= 0.0) return fail(FTS_INVALID, "ion species");
        total += ion.z * static_cast<double>(ion.n);
        scale += std::fabs(ion.z * static_cast<double>(ion.n));
    }
    if (electro() && std::fabs(total) > 1.0e-8 * std::max(1.0, scale)) return fail(FTS_INVALID, "electroneutrality failed");
    if (ions.size() > FTS_MAX_IONS) return fail(FTS_INVALID, "too many ion species");
    return FTS_OK;
}

// Device buffers, FFT tables and per-species bead classes.
int fts_engine::setup_device()
{
    if (!stream) {
        HIP_TRY(hipStreamCreate(&stream));
        own_stream = true;
    }
    const size_t Ms = static_cast<size_t>(M);
    // wave numbers and twiddles
    std::vector<double> kv;
    for (int x = 0; x < m1; ++x) kv.push_back(wave_number(x, m1, L1));
    for (int y = 0; y < m2; ++y) kv.push_back(wave_number(y, m2, L2));
    for (int z = 0; z < m3; ++z) kv.push_back(wave_number(z, m3, L3));
    HIP_TRY(kvec.alloc(kv.size()));
    HIP_TRY(hipMemcpy(kvec.p, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice));
    auto twiddles = [](int n, std::vector<double2>& t64, std::vector<float2>& t32) {
        const long double two_pi = 6.283185307179586476925286766559L;
        for (int t = 0; t < n; ++t) {
            const long double ang = -two_pi * t / n;
            double2 v;
            v.x = static_cast<double>(cosl(ang));
            v.y = static_cast<double>(sinl(ang));
            t64.push_back(v);
            float2 w;
            w.x = static_cast<float>(v.x);
            w.y = static_cast<float>(v.y);
            t32.push_back(w);
        }
    };
    std::vector<double2> tx64, tyz64, tz64;
    std::vector<float2> tx32, tyz32, tz32;
    twiddles(m1, tx64, tx32);
    twiddles(m2, tyz64, tyz32);
    twiddles(m3, tz64, tz32);
    HIP_TRY(tw64_z.alloc(m3));
    HIP_TRY(tw32_z.alloc(m3));
    HIP_TRY(hipMemcpy(tw64_z.p, tz64.data(), sizeof(double2) * m3, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tw32_z.p, tz32.data(), sizeof(float2) * m3, hipMemcpyHostToDevice));
    HIP_TRY(tw64_x.alloc(m1));
    HIP_TRY(tw64_yz.alloc(m2));
    HIP_TRY(tw32_x.alloc(m1));
    HIP_TRY(tw32_yz.alloc(m2));
    HIP_TRY(hipMemcpy(tw64_x.p, tx64.data(), sizeof(double2) * m1, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tw64_yz.p, tyz64.data(), sizeof(double2) * m2, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tw32_x.p, tx32.data(), sizeof(float2) * m1, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tw32_yz.p, tyz32.data(), sizeof(float2) * m2, hipMemcpyHostToDevice));
    g64 = FftGeom{m1, m2, m3, kvec.p, tw64_x.p, tw64_yz.p};
    g32 = FftGeom{m1, m2, m3, kvec.p, tw32_x.p, tw32_yz.p};
    // separable kernel tables: factor per dimension of exp(-c k^2) = prod_d exp(-c k_d^2)
    auto sep_table = [&](double c, std::vector<double>& out) {
        for (double k : kv) out.push_back(std::exp(-c * k * k));
    };
    std::vector<double> sg;
    sep_table(0.5 * a * a, sg); // Gamma = exp(-a^2 k^2 / 2)  (common.h:290-293)
    HIP_TRY(sep_gamma.alloc(sg.size()));
    HIP_TRY(hipMemcpy(sep_gamma.p, sg.data(), sizeof(double) * sg.size(), hipMemcpyHostToDevice));
    std::vector<double> sb;
    for (double b : bond_cls_len) sep_table(b * b / 6.0, sb); // phi = exp(-b^2 k^2 / 6)  (common.h:295-298)
    HIP_TRY(sep_bond.alloc(std::max<size_t>(1, sb.size())));
    if (!sb.empty()) HIP_TRY(hipMemcpy(sep_bond.p, sb.data(), sizeof(double) * sb.size(), hipMemcpyHostToDevice));
    std::vector<double> si;
    for (const HostIon& ion : ions) sep_table(0.5 * ion.a * ion.a, si);
    HIP_TRY(sep_ion.alloc(std::max<size_t>(1, si.size())));
    if (!si.empty()) HIP_TRY(hipMemcpy(sep_ion.p, si.data(), sizeof(double) * si.size(), hipMemcpyHostToDevice));
    // stress mode term (stress.cu:106-149)
    {
        double sum_over_k2[6] = {0, 0, 0, 0, 0, 0};
        long long modes = 0;
        for (int x = 0; x < m1; ++x)
            for (int y = 0; y < m2; ++y)
                for (int z = 0; z < m3; ++z) {
                    const double kx = kv[static_cast<size_t>(x)], ky = kv[static_cast<size_t>(m1 + y)], kz = kv[static_cast<size_t>(m1 + m2 + z)];
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    if (k2 > 0.0) {
                        ++modes;
                        const double v[6] = {kx * kx, ky * ky, kz * kz, kx * ky, kx * kz, ky * kz};
                        for (int c = 0; c < 6; ++c) sum_over_k2[c] += v[c] / k2;
                    }
                }
        for (int c = 0; c < 6; ++c)
            mode_term[c] = electro() ? ((1.0 / 3.0) * static_cast<double>(modes) * (c < 3 ? 1.0 : 0.0) - sum_over_k2[c]) / V : 0.0;
    }
    // fields
    const size_t nO = static_cast<size_t>(n_omega), nTs = static_cast<size_t>(nT), nI = std::max<size_t>(1, ions.size());
    for (DevBuf<double2>* b : {&omega, &omega_bar, &omega_s, &drift}) HIP_TRY(b->alloc(std::max<size_t>(1, nO) * Ms));
    for (DevBuf<double2>* b : {&psi, &psi_bar, &psi_s, &dpsi, &dpsi_bar, &rho_c_hat, &rho_c, &rho_c_poly, &rho_c_ion, &rho_p, &t1, &t2, &t3})
        HIP_TRY(b->alloc(Ms));
    if (electro() && plane_ok) HIP_TRY(psi_hat.alloc(Ms));
    for (DevBuf<double2>* b : {&W_type_s, &rho_hat, &rho_type}) HIP_TRY(b->alloc(nTs * Ms));
    for (DevBuf<double2>* b : {&ion_psi, &ion_b, &ion_rho}) HIP_TRY(b->alloc(nI * Ms));
    HIP_TRY(partials.alloc(8 * RED_BLOCKS));
    HIP_TRY(scal.alloc(SC_COUNT));
    HIP_TRY(stress_acc.alloc((stress_chain_base(static_cast<int>(species.size())) + nTs + 1 + nI) * 7));
    pin_bytes = sizeof(double2) * (SC_COUNT + stress_acc.n) + sizeof(double) * species.size() * (static_cast<size_t>(Nmax) + MAX_CLASSES);
    HIP_TRY(hipHostMalloc(&pin, pin_bytes));
    HIP_TRY(nkey.alloc(2));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&hkey), 2 * sizeof(unsigned long long)));
    HIP_TRY(flag.alloc(1));
    HIP_TRY(mode_coef.alloc(std::max<size_t>(1, nO * nTs)));
    mode_coef_host.assign(std::max<size_t>(1, nO * nTs), double2{0.0, 0.0});
    for (int l = 0; l < n_omega; ++l)
        for (int al = 0; al < nT; ++al) {
            // gamma_l * O[l][alpha]  (param.cu:125-133, complex_scale(gamma, O))
            const double o = Ocoef(l, al);
            mode_coef_host[static_cast<size_t>(l) * nT + al] = gamma_imag(l) ? double2{0.0 * o, 1.0 * o} : double2{1.0 * o, 0.0 * o};
        }
    HIP_TRY(hipMemcpy(mode_coef.p, mode_coef_host.data(), sizeof(double2) * mode_coef_host.size(), hipMemcpyHostToDevice));
    for (DevBuf<double2>* b : {&omega, &omega_bar, &psi, &psi_bar, &scal}) HIP_TRY(hipMemset(b->p, 0, sizeof(double2) * b->n));
    // propagators
    const size_t cs = csize();
    // qB is only stored for species without the mirror symmetry
    int NmaxB = 0;
    for (const HostSpecies& sp : species)
        if (!sp.mirror) NmaxB = std::max(NmaxB, sp.N);
    HIP_TRY(hipMalloc(&qF, cs * static_cast<size_t>(Nmax) * Ms));
    if (NmaxB > 0) HIP_TRY(hipMalloc(&qB, cs * static_cast<size_t>(NmaxB) * Ms));
    HIP_TRY(hipMalloc(&E, cs * static_cast<size_t>(max_cls) * Ms));
    for (void** p : {&tmpF, &tmpB, &kA, &kB}) HIP_TRY(hipMalloc(p, cs * Ms));
    // bead classes
    for (HostSpecies& sp : species) {
        const int nc = static_cast<int>(sp.cls_type.size());
        HIP_TRY(sp.d_cls_type.alloc(nc));
        HIP_TRY(sp.d_cls_z.alloc(nc));
        HIP_TRY(sp.d_cls_shift.alloc(nc));
        HIP_TRY(sp.d_cls_start.alloc(nc + 1));
        HIP_TRY(sp.d_cls_beads.alloc(sp.N));
        HIP_TRY(hipMemcpy(sp.d_cls_type.p, sp.cls_type.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(sp.d_cls_z.p, sp.cls_z.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(sp.d_cls_shift.p, 0, sizeof(double) * nc));
        HIP_TRY(hipMemcpy(sp.d_cls_start.p, sp.cls_start.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(sp.d_cls_beads.p, sp.cls_beads.data(), sizeof(int) * sp.N, hipMemcpyHostToDevice));
        if (sp.mirror) {
            HIP_TRY(sp.d_half_start.alloc(nc + 1));
            HIP_TRY(sp.d_half_beads.alloc(std::max<size_t>(1, sp.half_beads.size())));
            HIP_TRY(hipMemcpy(sp.d_half_start.p, sp.half_start.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(sp.d_half_beads.p, sp.half_beads.data(), sizeof(int) * sp.half_beads.size(), hipMemcpyHostToDevice));
        }
        sp.cls_shift.assign(static_cast<size_t>(nc), 0.0);
    }
    q_shift_total.assign(species.size(), 0.0);
    q_log_gain.assign(species.size(), 0.0);
    norm_stride = std::max(static_cast<size_t>(m1), static_cast<size_t>(m1) * std::max(m2, m3) / LINE_TC + 1);
    HIP_TRY(norms.alloc(4 * norm_stride));
    HIP_TRY(logc.alloc(2 * static_cast<size_t>(Nmax)));
    HIP_TRY(rfac.alloc(static_cast<size_t>(Nmax)));
    HIP_TRY(rfac2.alloc(static_cast<size_t>(Nmax)));
    HIP_TRY(hipMemset(logc.p, 0, sizeof(double) * logc.n));
    // mode coefficients for the update kernels
    mc.nO = n_omega;
    mc.nT = nT;
    for (int l = 0; l < n_omega; ++l) {
        mc.c[l] = omega_coefficient(l);
        mc.lam_dt[l] = dt * lambda_omega[static_cast<size_t>(l)];
        mc.sigma[l] = (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_omega[static_cast<size_t>(l)] * dt / dV);
        for (int al = 0; al < nT; ++al) mc.g[l * nT + al] = mode_coef_host[static_cast<size_t>(l) * nT + al];
    }
    HIP_TRY(hipDeviceSynchronize());
    return FTS_OK;
}

// Linear-response kernels of the semi-implicit scheme (same definition as the CPU
// oracle's; no reference counterpart):
//   A_l(k)   = c_l + s_l Gamma^2 sum_p (n_p/V) sum_{s,s'} O_{l a(s)} O_{l a(s')} phi_{ss'}(k)
//   A_psi(k) = k^2/(4 pi lB) + Gamma^2 sum_p (n_p/V) sum_{s,s'} z_s z_s' phi_{ss'}(k) + sum_j z_j^2 (n_j/V) Gamma_j^2
// The tables hold P(k) = 1 / (1 + dt lam A(k)).
int fts_engine::build_si_tables()
{
    std::vector<double> kv(static_cast<size_t>(m1 + m2 + m3));
    HIP_TRY(hipMemcpy(kv.data(), kvec.p, sizeof(double) * kv.size(), hipMemcpyDeviceToHost));
    std::vector<double> Pw(static_cast<size_t>(n_omega) * M), Pp(static_cast<size_t>(M), 1.0);
    std::vector<double> Aw(static_cast<size_t>(n_omega) * M), Ap(static_cast<size_t>(M), 0.0);
    auto pair_sum = [&](const HostSpecies& sp, const std::vector<double>& w, double ksq) {
        double total = 0.0, h = 0.0;
        for (int s = 0; s < sp.N; ++s) {
            total += w[static_cast<size_t>(s)] * w[static_cast<size_t>(s)] + 2.0 * w[static_cast<size_t>(s)] * h;
            if (s < sp.N - 1) {
                const double b = sp.bond[static_cast<size_t>(s)];
                h = (h + w[static_cast<size_t>(s)]) * std::exp(-b * b * ksq / 6.0);
            }
        }
        return total;
    };
    std::vector<std::vector<std::vector<double>>> wmode(static_cast<size_t>(n_omega));
    for (int l = 0; l < n_omega; ++l)
        for (const HostSpecies& sp : species) {
            std::vector<double> w(static_cast<size_t>(sp.N));
            for (int s = 0; s < sp.N; ++s) w[static_cast<size_t>(s)] = Ocoef(l, sp.type[static_cast<size_t>(s)]);
            wmode[static_cast<size_t>(l)].push_back(w);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        const int z = i % m3, y = (i / m3) % m2, x = i / (m2 * m3);
        const double kx = kv[static_cast<size_t>(x)], ky = kv[static_cast<size_t>(m1 + y)], kz = kv[static_cast<size_t>(m1 + m2 + z)];
        const double ksq = kx * kx + ky * ky + kz * kz;
        const double G2 = std::exp(-a * a * ksq);
        for (int l = 0; l < n_omega; ++l) {
            double resp = 0.0;
            for (size_t p = 0; p < species.size(); ++p)
                resp += static_cast<double>(species[p].n) / V * pair_sum(species[p], wmode[static_cast<size_t>(l)][p], ksq);
            const double A = omega_coefficient(l) + (gamma_imag(l) ? 1.0 : -1.0) * G2 * resp;
            Aw[static_cast<size_t>(l) * M + i] = A;
            Pw[static_cast<size_t>(l) * M + i] = 1.0 / (1.0 + dt * lambda_omega[static_cast<size_t>(l)] * A);
        }
        if (electro()) {
            double resp = 0.0;
            for (const HostSpecies& sp : species) resp += static_cast<double>(sp.n) / V * pair_sum(sp, sp.charge, ksq);
            double ion_resp = 0.0;
            for (const HostIon& ion : ions) {
                const double g = std::exp(-0.5 * ion.a * ion.a * ksq);
                ion_resp += ion.z * ion.z * static_cast<double>(ion.n) / V * g * g;
            }
            const double A = ksq / (4.0 * Pi * lB) + G2 * resp + ion_resp;
            Ap[static_cast<size_t>(i)] = A;
            Pp[static_cast<size_t>(i)] = 1.0 / (1.0 + dt * lambda_psi * A);
        }
    }
    for (double v : Pw)
        if (!(v > 0.0) || !std::isfinite(v)) return fail(FTS_INVALID, "semi-implicit kernel 1 + dt lam A must be positive");
    HIP_TRY(si_table_omega.alloc(Pw.size()));
    HIP_TRY(si_table_psi.alloc(Pp.size()));
    HIP_TRY(hipMemcpy(si_table_omega.p, Pw.data(), sizeof(double) * Pw.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(si_table_psi.p, Pp.data(), sizeof(double) * Pp.size(), hipMemcpyHostToDevice));
    HIP_TRY(si_A_omega.alloc(Aw.size()));
    HIP_TRY(si_A_psi.alloc(Ap.size()));
    if (!Aw.empty()) HIP_TRY(hipMemcpy(si_A_omega.p, Aw.data(), sizeof(double) * Aw.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(si_A_psi.p, Ap.data(), sizeof(double) * Ap.size(), hipMemcpyHostToDevice));
    if (scheme_two_stage(scheme)) {
        HIP_TRY(sipc_d.alloc(static_cast<size_t>(n_omega) * M));
        HIP_TRY(sipc_dpsi.alloc(static_cast<size_t>(M)));
    }
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// Learned linear-response kernel (no reference counterpart)
// ---------------------------------------------------------------------------
// Exact |k| bin of the wave vector with spectral indices (nx, ny, nz): the largest b
// with 3 b^2 m1^2 m2^2 m3^2 <= 4 B^2 (nx^2 m2^2 m3^2 + ny^2 m1^2 m3^2 + nz^2 m1^2 m2^2),
// i.e. b = floor(B |k| / k_ref) with k_ref = sqrt(3) pi / dL, clamped to B - 1.
// Integer arithmetic, so no rounding moves a wave vector across a bin edge.
static int lrk_bin_of(long long nx, long long ny, long long nz, long long m1, long long m2, long long m3)
{
    using u128 = unsigned __int128;
    const u128 a1 = static_cast<u128>(m1 * m1), a2 = static_cast<u128>(m2 * m2), a3 = static_cast<u128>(m3 * m3);
    const u128 rhs = static_cast<u128>(4) * LRK_NB * LRK_NB *
                     (static_cast<u128>(nx * nx) * a2 * a3 + static_cast<u128>(ny * ny) * a1 * a3 + static_cast<u128>(nz * nz) * a1 * a2);
    const u128 den = static_cast<u128>(3) * a1 * a2 * a3;
    int lo = 0, hi = LRK_NB - 1; // largest b in [0, B-1] with den b^2 <= rhs
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (den * static_cast<u128>(mid) * static_cast<u128>(mid) <= rhs) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Device buffers of the learned kernel and its judge (needs the A_D tables).
int fts_engine::lrk_setup()
{
    if (lrk_ready) return FTS_OK;
    if (n_omega > 1 || (n_omega == 0 && !electro()))
        return fail(FTS_INVALID, "the learned kernel covers interaction mode 0 and psi: needs n_omega <= 1 and at least one field");
    const size_t Ms = static_cast<size_t>(M);
    auto mode = [](int i, int m) { return i > m / 2 ? i - m : i; };
    std::vector<unsigned char> bin(Ms);
    for (int x = 0; x < m1; ++x)
        for (int y = 0; y < m2; ++y)
            for (int z = 0; z < m3; ++z)
                bin[(static_cast<size_t>(x) * m2 + y) * m3 + z] = static_cast<unsigned char>(lrk_bin_of(mode(x, m1), mode(y, m2), mode(z, m3), m1, m2, m3));
    // wave vectors k != 0 sorted by bin (stable in the grid index), split into chunks
    std::vector<int> count(LRK_NB + 1, 0);
    for (size_t i = 1; i < Ms; ++i) ++count[bin[i] + 1];
    for (int b = 0; b < LRK_NB; ++b) count[b + 1] += count[b];
    lrk_nperm = M - 1;
    std::vector<int> perm(static_cast<size_t>(std::max(1, lrk_nperm))), fill(count.begin(), count.end() - 1);
    for (size_t i = 1; i < Ms; ++i) perm[static_cast<size_t>(fill[bin[i]]++)] = static_cast<int>(i);
    std::vector<int> chunk_start, bin_chunk(LRK_NB + 1, 0);
    for (int b = 0; b < LRK_NB; ++b) {
        bin_chunk[b] = static_cast<int>(chunk_start.size());
        for (int j = count[b]; j < count[b + 1]; j += LRK_CHUNK) chunk_start.push_back(j);
    }
    bin_chunk[LRK_NB] = static_cast<int>(chunk_start.size());
    lrk_nchunks = static_cast<int>(chunk_start.size());
    // chunk c covers [cs[c], cs[c + 1]): after a bin's last chunk comes the first chunk of the
    // next non-empty bin, which starts where the bin ends
    std::vector<int> cs(chunk_start);
    cs.push_back(lrk_nperm);
    std::vector<double> Aw(Ms, 0.0), Ap(Ms, 0.0), ad(2 * static_cast<size_t>(std::max(1, lrk_nperm)), 0.0);
    if (n_omega > 0) HIP_TRY(hipMemcpy(Aw.data(), si_A_omega.p, sizeof(double) * Ms, hipMemcpyDeviceToHost));
    if (electro()) HIP_TRY(hipMemcpy(Ap.data(), si_A_psi.p, sizeof(double) * Ms, hipMemcpyDeviceToHost));
    for (int j = 0; j < lrk_nperm; ++j) {
        ad[static_cast<size_t>(j)] = Aw[static_cast<size_t>(perm[static_cast<size_t>(j)])];
        ad[static_cast<size_t>(lrk_nperm + j)] = Ap[static_cast<size_t>(perm[static_cast<size_t>(j)])];
    }
    HIP_TRY(lrk_bin.alloc(Ms));
    HIP_TRY(lrk_perm.alloc(perm.size()));
    HIP_TRY(lrk_chunk_start.alloc(cs.size()));
    HIP_TRY(lrk_bin_chunk.alloc(bin_chunk.size()));
    HIP_TRY(lrk_ad_perm.alloc(ad.size()));
    HIP_TRY(lrk_chunk_sums.alloc(6 * static_cast<size_t>(std::max(1, lrk_nchunks))));
    HIP_TRY(hipMemcpy(lrk_bin.p, bin.data(), Ms, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(lrk_perm.p, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(lrk_chunk_start.p, cs.data(), sizeof(int) * cs.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(lrk_bin_chunk.p, bin_chunk.data(), sizeof(int) * bin_chunk.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(lrk_ad_perm.p, ad.data(), sizeof(double) * ad.size(), hipMemcpyHostToDevice));
    for (int c = 0; c < 2; ++c) {
        if ((c == 0 && n_omega == 0) || (c == 1 && !electro())) continue;
        for (int k = 0; k < 2; ++k)
            for (int par = 0; par < 2; ++par) HIP_TRY(lrk_hat[c][k][par].alloc(Ms));
    }
    if (n_omega > 0) HIP_TRY(lrk_u.alloc(Ms));
    if (!lrk_params.p) {
        HIP_TRY(lrk_params.alloc(FTS_LRK_NPARAMS));
        HIP_TRY(hipMemset(lrk_params.p, 0, sizeof(float) * FTS_LRK_NPARAMS));
    }
    HIP_TRY(lrk_state.alloc(1));
    HIP_TRY(hipMemset(lrk_state.p, 0, sizeof(LrkDevState)));
    lrk_ready = true;
    lrk_have_prev = false;
    return FTS_OK;
}

// One judged step: called by the semi-implicit update once u(w_t) is known (lrk_u after
// k_mode0_drift, dpsi after psi_drift).  Judges the previous step, writes its record,
// decides this step's kernel and (the learned schemes) rebuilds the P tables the update applies.
int fts_engine::lrk_step()
{
    int rc;
    const int cur = lrk_par, prv = lrk_par ^ 1;
    const bool f0 = n_omega > 0, f1 = electro();
    FftJob fj[4];
    int nf = 0;
    if (f0) {
        hipLaunchKernelGGL(k_mode0_drift, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, mc, omega.p, rho_type.p, M, lrk_u.p);
        HIP_TRY(hipGetLastError());
        fj[nf++] = FftJob{omega.p, lrk_hat[0][0][cur].p, KOp{}};
        fj[nf++] = FftJob{lrk_u.p, lrk_hat[0][1][cur].p, KOp{}};
    }
    // an FP32 engine keeps the judge's spectra in complex FP32 (its transforms are FP32):
    // half the bytes for k_lrk_bin_chunks' gather
    const bool s32 = fields_fp32();
    if (f1) {
        if (psi_hat_valid(psi.p)) {
            if (s32) {
                hipLaunchKernelGGL((k_cast<double, float>), dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, psi_hat.p,
                                   reinterpret_cast<float2*>(lrk_hat[1][0][cur].p), 1.0, static_cast<size_t>(M));
                HIP_TRY(hipGetLastError());
            } else {
                HIP_TRY(hipMemcpyAsync(lrk_hat[1][0][cur].p, psi_hat.p, sizeof(double2) * M, hipMemcpyDeviceToDevice, stream));
            }
        } else {
            fj[nf++] = FftJob{psi.p, lrk_hat[1][0][cur].p, KOp{}};
        }
        fj[nf++] = FftJob{dpsi.p, lrk_hat[1][1][cur].p, KOp{}};
    }
    if ((rc = s32 ? fft32_n(3, fj, nf) : fft_forward_n(fj, nf))) return rc;
    HIP_TRY(lrk_moment_sums(f0 ? omega.p : nullptr, f1 ? psi.p : nullptr, M, partials.p, stream));
    if ((rc = reduce_finalize(4, SC_LRK, 1.0 / static_cast<double>(M)))) return rc;
    HIP_TRY(lrk_infer(lrk_params.p, scal.p + SC_LRK, lrk_state.p, stream));
    if (lrk_have_prev)
        HIP_TRY(lrk_bin_chunks(lrk_nchunks, lrk_chunk_start.p, lrk_perm.p, lrk_ad_perm.p, lrk_nperm, f0 ? lrk_hat[0][0][cur].p : nullptr,
                               f0 ? lrk_hat[0][0][prv].p : nullptr, f0 ? lrk_hat[0][1][cur].p : nullptr, f0 ? lrk_hat[0][1][prv].p : nullptr,
                               f1 ? lrk_hat[1][0][cur].p : nullptr, f1 ? lrk_hat[1][0][prv].p : nullptr, f1 ? lrk_hat[1][1][cur].p : nullptr,
                               f1 ? lrk_hat[1][1][prv].p : nullptr, s32 ? 1 : 0, lrk_chunk_sums.p, stream));
    HIP_TRY(lrk_judge(lrk_state.p, lrk_have_prev ? 1 : 0, lrk_bin_chunk.p, lrk_chunk_sums.p, lrk_threshold,
                      learned_scheme() ? 1 : 0, lrk_ring.p, lrk_capacity, stream));
    if (learned_scheme())
        HIP_TRY(lrk_ptable(lrk_state.p, lrk_bin.p, f0 ? si_A_omega.p : nullptr, f1 ? si_A_psi.p : nullptr,
                           f0 ? dt * lambda_omega[0] : 0.0, dt * lambda_psi, M, f0 ? si_table_omega.p : nullptr, f1 ? si_table_psi.p : nullptr,
                           stream));
    lrk_par = prv;
    lrk_have_prev = true;
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// Primitives
// ---------------------------------------------------------------------------
// dst = IFFT[K(k) FFT[src]]; callers fold 1/M into K (fft_apply_real_kernel, common.h:1079-1085)
int fts_engine::fft_apply(const double2* src, double2* dst, const KOp& op)
{
    if (fields_fp32()) {
        const FftJob j{src, dst, op};
        return fft32_n(0, &j, 1);
    }
    if (plane_ok) {
        PlaneJob<double> pj{src, dst, nullptr, nullptr};
        HIP_TRY(launch_plane<double>(g64, PLANE_FWD, &pj, 1, stream));
        PencilJob<double> xj{dst, op};
        HIP_TRY(launch_pencil<double>(g64, PENCIL_FWD_MUL_INV, &xj, 1, stream));
        PlaneJob<double> ij{dst, dst, nullptr, nullptr};
        HIP_TRY(launch_plane<double>(g64, PLANE_INV, &ij, 1, stream));
        return FTS_OK;
    }
    if (dst != src) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double2) * M, hipMemcpyDeviceToDevice, stream));
    LineJob<double> j{dst};
    HIP_TRY(launch_line<double>(g64, 2, LINE_FWD, tw64_z.p, &j, 1, stream));
    HIP_TRY(launch_line<double>(g64, 1, LINE_FWD, tw64_yz.p, &j, 1, stream));
    LineJob<double> xj{dst, nullptr, nullptr, op};
    HIP_TRY(launch_line<double>(g64, 0, LINE_FWD_MUL_INV, tw64_x.p, &xj, 1, stream));
    HIP_TRY(launch_line<double>(g64, 1, LINE_INV, tw64_yz.p, &j, 1, stream));
    HIP_TRY(launch_line<double>(g64, 2, LINE_INV, tw64_z.p, &j, 1, stream));
    return FTS_OK;
}

// fft_apply from a spectrum fft_forward made: the x pass multiplies and transforms back
// out of place, then the yz pass; the same arithmetic as fft_apply on the field
int fts_engine::fft_apply_hat(const double2* hat, double2* dst, const KOp& op)
{
    if (fields_fp32()) {
        const FftJob j{hat, dst, op};
        return fft32_n(1, &j, 1);
    }
    PencilJob<double> xj{dst, op, hat};
    HIP_TRY(launch_pencil<double>(g64, PENCIL_MUL_INV, &xj, 1, stream));
    PlaneJob<double> ij{dst, dst, nullptr, nullptr};
    HIP_TRY(launch_plane<double>(g64, PLANE_INV, &ij, 1, stream));
    return FTS_OK;
}

// forward transforms src -> dst, and src2 -> dst2 in the same launches when given
template <typename T> int fts_engine::fft_forward_t(const V2<T>* src, V2<T>* dst, const V2<T>* src2, V2<T>* dst2)
{
    const FftGeom& g = sizeof(T) == 8 ? g64 : g32;
    const int n = src2 ? 2 : 1;
    if (plane_ok) {
        PlaneJob<T> pj[2] = {{src, dst, nullptr, nullptr}, {src2, dst2, nullptr, nullptr}};
        HIP_TRY(launch_plane<T>(g, PLANE_FWD, pj, n, stream));
        PencilJob<T> xj[2] = {{dst, KOp{}}, {dst2, KOp{}}};
        HIP_TRY(launch_pencil<T>(g, PENCIL_FWD, xj, n, stream));
        return FTS_OK;
    }
    const bool d = sizeof(T) == 8;
    // the z pass reads src and writes dst; y and x transform dst in place
    LineJob<T> zj[2] = {{dst, nullptr, nullptr, KOp{}, nullptr, nullptr, nullptr, src}, {dst2, nullptr, nullptr, KOp{}, nullptr, nullptr, nullptr, src2}};
    LineJob<T> j[2] = {{dst}, {dst2}};
    HIP_TRY(launch_line<T>(g, 2, LINE_FWD, d ? static_cast<const void*>(tw64_z.p) : static_cast<const void*>(tw32_z.p), zj, n, stream));
    HIP_TRY(launch_line<T>(g, 1, LINE_FWD, d ? static_cast<const void*>(tw64_yz.p) : static_cast<const void*>(tw32_yz.p), j, n, stream));
    HIP_TRY(launch_line<T>(g, 0, LINE_FWD, d ? static_cast<const void*>(tw64_x.p) : static_cast<const void*>(tw32_x.p), j, n, stream));
    return FTS_OK;
}

int fts_engine::fft_forward(const double2* src, double2* dst)
{
    if (fields_fp32()) {
        const FftJob j{src, dst, KOp{}};
        return fft32_n(2, &j, 1);
    }
    return fft_forward_t<double>(src, dst);
}

// FP64 fields through FP32 plane / pencil passes, two jobs per launch.  kind 3: forward
// into a complex FP32 spectrum (dst holds float2 values; no scratch)
int fts_engine::fft32_n(int kind, const FftJob* jobs, int n)
{
    if (kind == 3) {
        for (int k = 0; k < n; k += 2) {
            const int c = std::min(2, n - k);
            PlaneJob<float> pj[2] = {};
            PencilJob<float> xj[2] = {};
            for (int u = 0; u < c; ++u) {
                float2* d = reinterpret_cast<float2*>(jobs[k + u].dst);
                pj[u].dst = d;
                pj[u].src64 = jobs[k + u].src;
                xj[u].data = d;
            }
            HIP_TRY(launch_plane<float>(g32, PLANE_FWD_D2F, pj, c, stream));
            HIP_TRY(launch_pencil<float>(g32, PENCIL_FWD, xj, c, stream));
        }
        return FTS_OK;
    }
    float2* scr[2] = {static_cast<float2*>(tmpF), static_cast<float2*>(tmpB)};
    for (int k = 0; k < n; k += 2) {
        const int c = std::min(2, n - k);
        PlaneJob<float> pj[2] = {};
        PencilJob<float> xj[2] = {};
        for (int u = 0; u < c; ++u) {
            const FftJob& j = jobs[k + u];
            xj[u].data = scr[u];
            xj[u].op = j.op;
            if (kind == 0 || kind == 2) {
                pj[u].dst = scr[u];
                pj[u].src64 = j.src;
            } else {
                xj[u].src64 = j.src;
            }
            if (kind == 2) xj[u].dst64 = j.dst;
        }
        if (kind == 0 || kind == 2) HIP_TRY(launch_plane<float>(g32, PLANE_FWD_D2F, pj, c, stream));
        HIP_TRY(launch_pencil<float>(g32, kind == 0 ? PENCIL_FWD_MUL_INV : kind == 1 ? PENCIL_MUL_INV_D2F : PENCIL_FWD_F2D, xj, c, stream));
        if (kind == 2) continue;
        for (int u = 0; u < c; ++u) {
            pj[u] = PlaneJob<float>{};
            pj[u].src = scr[u];
            pj[u].dst64 = jobs[k + u].dst;
        }
        HIP_TRY(launch_plane<float>(g32, PLANE_INV_F2D, pj, c, stream));
    }
    return FTS_OK;
}

int fts_engine::fft_apply_n(const FftJob* jobs, int n)
{
    if (fields_fp32()) return fft32_n(0, jobs, n);
    for (int k = 0; k < n; k += 2) {
        const int c = std::min(2, n - k);
        if (!plane_ok) {
            for (int u = 0; u < c; ++u)
                if (int rc = fft_apply(jobs[k + u].src, jobs[k + u].dst, jobs[k + u].op)) return rc;
            continue;
        }
        PlaneJob<double> pj[2], ij[2];
        PencilJob<double> xj[2];
        for (int u = 0; u < c; ++u) {
            const FftJob& j = jobs[k + u];
            pj[u] = PlaneJob<double>{j.src, j.dst, nullptr, nullptr};
            xj[u] = PencilJob<double>{j.dst, j.op};
            ij[u] = PlaneJob<double>{j.dst, j.dst, nullptr, nullptr};
        }
        HIP_TRY(launch_plane<double>(g64, PLANE_FWD, pj, c, stream));
        HIP_TRY(launch_pencil<double>(g64, PENCIL_FWD_MUL_INV, xj, c, stream));
        HIP_TRY(launch_plane<double>(g64, PLANE_INV, ij, c, stream));
    }
    return FTS_OK;
}

int fts_engine::fft_apply_hat_n(const FftJob* jobs, int n)
{
    if (fields_fp32()) return fft32_n(1, jobs, n);
    for (int k = 0; k < n; k += 2) {
        const int c = std::min(2, n - k);
        PencilJob<double> xj[2];
        PlaneJob<double> ij[2];
        for (int u = 0; u < c; ++u) {
            const FftJob& j = jobs[k + u];
            xj[u] = PencilJob<double>{j.dst, j.op, j.src};
            ij[u] = PlaneJob<double>{j.dst, j.dst, nullptr, nullptr};
        }
        HIP_TRY(launch_pencil<double>(g64, PENCIL_MUL_INV, xj, c, stream));
        HIP_TRY(launch_plane<double>(g64, PLANE_INV, ij, c, stream));
    }
    return FTS_OK;
}

int fts_engine::fft_forward_n(const FftJob* jobs, int n)
{
    if (fields_fp32()) return fft32_n(2, jobs, n);
    for (int k = 0; k < n; k += 2) {
        const int c = std::min(2, n - k);
        if (!plane_ok) {
            for (int u = 0; u < c; ++u)
                if (int rc = fft_forward(jobs[k + u].src, jobs[k + u].dst)) return rc;
            continue;
        }
        PlaneJob<double> pj[2];
        PencilJob<double> xj[2];
        for (int u = 0; u < c; ++u) {
            pj[u] = PlaneJob<double>{jobs[k + u].src, jobs[k + u].dst, nullptr, nullptr};
            xj[u] = PencilJob<double>{jobs[k + u].dst, KOp{}};
        }
        HIP_TRY(launch_plane<double>(g64, PLANE_FWD, pj, c, stream));
        HIP_TRY(launch_pencil<double>(g64, PENCIL_FWD, xj, c, stream));
    }
    return FTS_OK;
}

int fts_engine::reduce_finalize(int Kc, int slot, double scale)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RED_THREADS), 0, stream, partials.p, Kc, scal.p + slot, scale);
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

int fts_engine::sum_field(const double2* x, int slot, double scale)
{
    hipLaunchKernelGGL(k_sum<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, x, M, partials.p);
    HIP_TRY(hipGetLastError());
    return reduce_finalize(1, slot, scale);
}

// a -= mean(a)  (remove_complex_mean, common.h:1008-1012)
int fts_engine::remove_mean(double2* x)
{
    int rc = sum_field(x, SC_TMP, 1.0 / static_cast<double>(M));
    if (rc) return rc;
    hipLaunchKernelGGL(k_add_scalar, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, x, scal.p + SC_TMP, -1.0, M);
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

int fts_engine::read_scal(int slot, int count, double2* out)
{
    HIP_TRY(hipMemcpyAsync(out, scal.p + slot, sizeof(double2) * count, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return FTS_OK;
}

// host copies of the scales of the last solve: LF(N-1) per species (per-bead
// renormalisation) and the FP32 class shifts with their sum over the chain
int fts_engine::read_log_gain()
{
    double2 v[FTS_MAX_SPECIES];
    int rc = read_scal(SC_LF, static_cast<int>(species.size()), v);
    if (rc) return rc;
    for (size_t p = 0; p < species.size(); ++p) {
        HostSpecies& sp = species[p];
        q_log_gain[p] = v[p].x;
        HIP_TRY(hipMemcpyAsync(sp.cls_shift.data(), sp.d_cls_shift.p, sizeof(double) * sp.cls_shift.size(), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        double total = 0.0;
        for (int s = 0; s < sp.N; ++s) total += sp.cls_shift[static_cast<size_t>(sp.cls_of_bead[static_cast<size_t>(s)])];
        q_shift_total[p] = total;
    }
    return FTS_OK;
}

hipEvent_t fts_engine::get_event()
{
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

int fts_engine::resolve_timing()
{
    if (pending.empty()) return FTS_OK;
    HIP_TRY(hipEventSynchronize(pending.back().second));
    for (auto& pr : pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
        tstat.sweep_ms += ms;
        event_pool.push_back(pr.first);
        event_pool.push_back(pr.second);
    }
    pending.clear();
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// Both propagator sweeps of species p for nsteps bead steps (polymer.cu:402-416):
//   forward  qF_{s+1} = e^{-W_{s+1}} S_{phi_s}[qF_s],       s = 0 .. nsteps-1
//   backward qB_{t-1} = e^{-W_{t-1}} S_{phi_{t-1}}[qB_t],   t = N-1 .. N-nsteps
// Every launch carries the two sweeps as two jobs.  Each pass that produces a
// propagator rescales it by a power of two from the previous bead's maximum
// (ping-pong norm buffers) and records the exponent in logc.
// ---------------------------------------------------------------------------
template <typename T> int fts_engine::run_sweeps(int p, int schedule, int nsteps, bool timed)
{
    using Cv = V2<T>;
    HostSpecies& sp = species[static_cast<size_t>(p)];
    const FftGeom& g = sizeof(T) == 8 ? g64 : g32;
    const size_t Ms = static_cast<size_t>(M);
    Cv* qf = static_cast<Cv*>(qF);
    Cv* qb = static_cast<Cv*>(qB);
    Cv* e = static_cast<Cv*>(E);
    Cv* tf = static_cast<Cv*>(tmpF);
    Cv* tb = static_cast<Cv*>(tmpB);
    const int N = sp.N;
    auto E_of = [&](int s) { return e + static_cast<size_t>(sp.cls_of_bead[static_cast<size_t>(s)]) * Ms; };
    // a mirrored species runs the forward sweep only (one job per launch)
    const int nj = sp.mirror ? 1 : 2;
    double* nF[2] = {norms.p, norms.p + norm_stride};
    double* nB[2] = {norms.p + 2 * norm_stride, norms.p + 3 * norm_stride};
    double* lcF = logc.p;
    double* lcB = logc.p + Nmax;
    const double invM = 1.0 / static_cast<double>(M);
    const size_t ks = static_cast<size_t>(m1 + m2 + m3);
    const void* twx = sizeof(T) == 8 ? static_cast<const void*>(tw64_x.p) : static_cast<const void*>(tw32_x.p);
    const void* twy = sizeof(T) == 8 ? static_cast<const void*>(tw64_yz.p) : static_cast<const void*>(tw32_yz.p);
    const void* twz = sizeof(T) == 8 ? static_cast<const void*>(tw64_z.p) : static_cast<const void*>(tw32_z.p);
    hipEvent_t ev0 = nullptr;
    if (sched_is_conv(schedule)) {
        // bead step s: pencil (x factor and 1/M) q_s -> tmp, plane (y, z factors,
        // e^{-W}, renormalisation) tmp -> q_{s+1}; the first plane pass stores q_0
        // (its transformed copy in tmp is not used)
        PlaneJob<T> init[2] = {{nullptr, tf, qf, E_of(0), nullptr, nF[0], lcF},
                               {nullptr, tb, qb + static_cast<size_t>(N - 1) * Ms, E_of(N - 1), nullptr, nB[0], lcB + (N - 1)}};
        HIP_TRY(launch_plane<T>(g, PLANE_INIT, init, nj, stream));
        if (timed) {
            ev0 = get_event();
            HIP_TRY(hipEventRecord(ev0, stream));
        }
        for (int s = 0; s < nsteps; ++s) {
            const int bf = sp.bond_cls[static_cast<size_t>(s)], bb = sp.bond_cls[static_cast<size_t>(N - 2 - s)];
            const KOp of = kop_sep(sep_bond.p + static_cast<size_t>(bf) * ks, invM), ob = kop_sep(sep_bond.p + static_cast<size_t>(bb) * ks, invM);
            PencilJob<T> pj[2] = {{tf, of, qf + static_cast<size_t>(s) * Ms}, {tb, ob, qb + static_cast<size_t>(N - 1 - s) * Ms}};
            HIP_TRY(launch_pencil<T>(g, conv_pencil_mode(schedule), pj, nj, stream));
            const int sf = s + 1, sb = N - 2 - s;
            PlaneJob<T> st[2] = {{tf, nullptr, qf + static_cast<size_t>(sf) * Ms, E_of(sf), nF[s % 2], nF[(s + 1) % 2], lcF + sf, of},
                                 {tb, nullptr, qb + static_cast<size_t>(sb) * Ms, E_of(sb), nB[s % 2], nB[(s + 1) % 2], lcB + sb, ob}};
            HIP_TRY(launch_plane<T>(g, PLANE_CONV_E, st, nj, stream));
        }
    } else if (schedule == FTS_SCHED_PLANE) {
        PlaneJob<T> init[2] = {{nullptr, tf, qf, E_of(0), nullptr, nF[0], lcF},
                               {nullptr, tb, qb + static_cast<size_t>(N - 1) * Ms, E_of(N - 1), nullptr, nB[0], lcB + (N - 1)}};
        HIP_TRY(launch_plane<T>(g, PLANE_INIT, init, nj, stream));
        if (timed) {
            ev0 = get_event();
            HIP_TRY(hipEventRecord(ev0, stream));
        }
        for (int s = 0; s < nsteps; ++s) {
            const int bf = sp.bond_cls[static_cast<size_t>(s)], bb = sp.bond_cls[static_cast<size_t>(N - 2 - s)];
            PencilJob<T> pj[2] = {{tf, kop_sep(sep_bond.p + static_cast<size_t>(bf) * ks, invM)},
                                  {tb, kop_sep(sep_bond.p + static_cast<size_t>(bb) * ks, invM)}};
            HIP_TRY(launch_pencil<T>(g, PENCIL_FWD_MUL_INV, pj, nj, stream));
            const int sf = s + 1, sb = N - 2 - s;
            PlaneJob<T> st[2] = {{tf, tf, qf + static_cast<size_t>(sf) * Ms, E_of(sf), nF[s % 2], nF[(s + 1) % 2], lcF + sf},
                                 {tb, tb, qb + static_cast<size_t>(sb) * Ms, E_of(sb), nB[s % 2], nB[(s + 1) % 2], lcB + sb}};
            HIP_TRY(launch_plane<T>(g, s < nsteps - 1 ? PLANE_STEP : PLANE_LAST, st, nj, stream));
        }
    } else if (schedule == FTS_SCHED_SEPARABLE) {
        // bead step s, one axis of the separable bond kernel per pass: x line pass (x
        // factor and 1/M) q_s -> tmp, z line pass (z factor) in place, y line pass (y
        // factor, e^{-W}, renormalisation) tmp -> q_{s+1}; the first y pass stores
        // q_0 = e^{-W_0} alone
        LineJob<T> init[2] = {{nullptr, qf, E_of(0), KOp{}, nullptr, nF[0], lcF},
                              {nullptr, qb + static_cast<size_t>(N - 1) * Ms, E_of(N - 1), KOp{}, nullptr, nB[0], lcB + (N - 1)}};
        HIP_TRY(launch_line<T>(g, 1, LINE_E, twy, init, nj, stream));
        if (timed) {
            ev0 = get_event();
            HIP_TRY(hipEventRecord(ev0, stream));
        }
        for (int s = 0; s < nsteps; ++s) {
            const int bf = sp.bond_cls[static_cast<size_t>(s)], bb = sp.bond_cls[static_cast<size_t>(N - 2 - s)];
            const KOp of = kop_sep(sep_bond.p + static_cast<size_t>(bf) * ks, invM), ob = kop_sep(sep_bond.p + static_cast<size_t>(bb) * ks, invM);
            const KOp of1 = kop_sep(of.sep, 1.0), ob1 = kop_sep(ob.sep, 1.0);
            LineJob<T> xj[2] = {{tf, nullptr, nullptr, of, nullptr, nullptr, nullptr, qf + static_cast<size_t>(s) * Ms},
                                {tb, nullptr, nullptr, ob, nullptr, nullptr, nullptr, qb + static_cast<size_t>(N - 1 - s) * Ms}};
            HIP_TRY(launch_line<T>(g, 0, LINE_CONV, twx, xj, nj, stream));
            LineJob<T> zj[2] = {{tf, nullptr, nullptr, of1}, {tb, nullptr, nullptr, ob1}};
            HIP_TRY(launch_line<T>(g, 2, LINE_CONV, twz, zj, nj, stream));
            const int sf = s + 1, sb = N - 2 - s;
            LineJob<T> yj[2] = {{tf, qf + static_cast<size_t>(sf) * Ms, E_of(sf), of1, nF[s % 2], nF[(s + 1) % 2], lcF + sf},
                                {tb, qb + static_cast<size_t>(sb) * Ms, E_of(sb), ob1, nB[s % 2], nB[(s + 1) % 2], lcB + sb}};
            HIP_TRY(launch_line<T>(g, 1, LINE_CONV_E, twy, yj, nj, stream));
        }
    } else {
        LineJob<T> init[2] = {{tf, qf, E_of(0), KOp{}, nullptr, nF[0], lcF},
                              {tb, qb + static_cast<size_t>(N - 1) * Ms, E_of(N - 1), KOp{}, nullptr, nB[0], lcB + (N - 1)}};
        HIP_TRY(launch_line<T>(g, 2, LINE_E_FWD, twz, init, nj, stream));
        LineJob<T> yf[2] = {{tf}, {tb}};
        HIP_TRY(launch_line<T>(g, 1, LINE_FWD, twy, yf, nj, stream));
        if (timed) {
            ev0 = get_event();
            HIP_TRY(hipEventRecord(ev0, stream));
        }
        for (int s = 0; s < nsteps; ++s) {
            const int bf = sp.bond_cls[static_cast<size_t>(s)], bb = sp.bond_cls[static_cast<size_t>(N - 2 - s)];
            LineJob<T> xj[2] = {{tf, nullptr, nullptr, kop_sep(sep_bond.p + static_cast<size_t>(bf) * ks, invM)},
                                {tb, nullptr, nullptr, kop_sep(sep_bond.p + static_cast<size_t>(bb) * ks, invM)}};
            HIP_TRY(launch_line<T>(g, 0, LINE_FWD_MUL_INV, twx, xj, nj, stream));
            HIP_TRY(launch_line<T>(g, 1, LINE_INV, twy, yf, nj, stream));
            const int sf = s + 1, sb = N - 2 - s;
            const bool last = s == nsteps - 1;
            LineJob<T> zj[2] = {{tf, qf + static_cast<size_t>(sf) * Ms, E_of(sf), KOp{}, nF[s % 2], nF[(s + 1) % 2], lcF + sf},
                                {tb, qb + static_cast<size_t>(sb) * Ms, E_of(sb), KOp{}, nB[s % 2], nB[(s + 1) % 2], lcB + sb}};
            HIP_TRY(launch_line<T>(g, 2, last ? LINE_INV_E : LINE_INV_E_FWD, twz, zj, nj, stream));
            if (!last) HIP_TRY(launch_line<T>(g, 1, LINE_FWD, twy, yf, nj, stream));
        }
    }
    if (timed) {
        hipEvent_t ev1 = get_event();
        HIP_TRY(hipEventRecord(ev1, stream));
        pending.emplace_back(ev0, ev1);
        tstat.bead_steps += static_cast<long long>(nj) * nsteps;
        tstat.step_launches += nsteps;
        tstat.sweeps += 1;
        tstat.algorithmic_bytes += static_cast<double>(nj) * nsteps * 2.0 * static_cast<double>(sizeof(Cv)) * static_cast<double>(M);
    }
    return FTS_OK;
}

// One bead step on a single field with the engine's schedule (fts_bead_step): q is
// overwritten with intermediates, qout = e^{-W} S_phi[q].
template <typename T> int fts_engine::bead_hook(V2<T>* q, const V2<T>* emw, V2<T>* qout, const KOp& op, hipStream_t stream)
{
    const FftGeom& g = sizeof(T) == 8 ? g64 : g32;
    const bool d = sizeof(T) == 8;
    if (sched_is_conv(sched)) {
        PencilJob<T> j1{q, op};
        HIP_TRY(launch_pencil<T>(g, conv_pencil_mode(sched), &j1, 1, stream));
        PlaneJob<T> j2{q, nullptr, qout, emw, nullptr, nullptr, nullptr, op};
        HIP_TRY(launch_plane<T>(g, PLANE_CONV_E, &j2, 1, stream));
        return FTS_OK;
    }
    if (sched == FTS_SCHED_PLANE) {
        PlaneJob<T> j1{q, q, nullptr, nullptr};
        HIP_TRY(launch_plane<T>(g, PLANE_FWD, &j1, 1, stream));
        PencilJob<T> j2{q, op};
        HIP_TRY(launch_pencil<T>(g, PENCIL_FWD_MUL_INV, &j2, 1, stream));
        PlaneJob<T> j3{q, nullptr, qout, emw};
        HIP_TRY(launch_plane<T>(g, PLANE_LAST, &j3, 1, stream));
        return FTS_OK;
    }
    const void* twx = d ? static_cast<const void*>(tw64_x.p) : static_cast<const void*>(tw32_x.p);
    const void* twy = d ? static_cast<const void*>(tw64_yz.p) : static_cast<const void*>(tw32_yz.p);
    const void* twz = d ? static_cast<const void*>(tw64_z.p) : static_cast<const void*>(tw32_z.p);
    if (sched == FTS_SCHED_SEPARABLE) {
        LineJob<T> xj{q, nullptr, nullptr, op};
        HIP_TRY(launch_line<T>(g, 0, LINE_CONV, twx, &xj, 1, stream));
        LineJob<T> zj{q, nullptr, nullptr, kop_sep(op.sep, 1.0)};
        HIP_TRY(launch_line<T>(g, 2, LINE_CONV, twz, &zj, 1, stream));
        LineJob<T> yj{q, qout, emw, kop_sep(op.sep, 1.0)};
        HIP_TRY(launch_line<T>(g, 1, LINE_CONV_E, twy, &yj, 1, stream));
        return FTS_OK;
    }
    LineJob<T> j{q};
    HIP_TRY(launch_line<T>(g, 2, LINE_FWD, twz, &j, 1, stream));
    HIP_TRY(launch_line<T>(g, 1, LINE_FWD, twy, &j, 1, stream));
    LineJob<T> xj{q, nullptr, nullptr, op};
    HIP_TRY(launch_line<T>(g, 0, LINE_FWD_MUL_INV, twx, &xj, 1, stream));
    HIP_TRY(launch_line<T>(g, 1, LINE_INV, twy, &j, 1, stream));
    LineJob<T> zj{q, qout, emw};
    HIP_TRY(launch_line<T>(g, 2, LINE_INV_E, twz, &zj, 1, stream));
    return FTS_OK;
}

template int fts_engine::bead_hook<float>(V2<float>*, const V2<float>*, V2<float>*, const KOp&, hipStream_t);
template int fts_engine::bead_hook<double>(V2<double>*, const V2<double>*, V2<double>*, const KOp&, hipStream_t);

// Bead-step schedule chosen by timing conv and separable on this grid and precision
// (the FFT autotuner): 2 x 8 bead steps of species 0 at zero fields, best of three;
// tune_ms = {conv, separable}.  The plane schedule is taken only when asked for: it is
// within a few percent of conv either way, so timing it too would make the choice,
// and the last bits of the results, vary from run to run.  The line schedule moves
// more bytes than the separable one for the same transforms and is never chosen.
int fts_engine::autotune_schedule()
{
    if (!plane_ok) {
        sched = FTS_SCHED_SEPARABLE;
        return FTS_OK;
    }
    const HostSpecies& sp = species[0];
    if (sp.N < 2) {
        sched = FTS_SCHED_CONV;
        return FTS_OK;
    }
    const int steps = std::min(8, sp.N - 1);
    const int cand[2] = {FTS_SCHED_CONV, FTS_SCHED_SEPARABLE};
    double best[2] = {1e30, 1e30};
    hipEvent_t a = get_event(), b = get_event();
    for (int rep = 0; rep < 3; ++rep)
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(hipEventRecord(a, stream));
            int rc = precision == FTS_PREC_FP32 ? run_sweeps<float>(0, cand[k], steps, false) : run_sweeps<double>(0, cand[k], steps, false);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(b, stream));
            HIP_TRY(hipEventSynchronize(b));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, a, b));
            best[k] = std::min(best[k], static_cast<double>(ms));
        }
    event_pool.push_back(a);
    event_pool.push_back(b);
    int pick = 0;
    for (int k = 0; k < 2; ++k) {
        tune_ms[k] = best[k];
        if (best[k] < best[pick]) pick = k;
    }
    sched = cand[pick];
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// Propagator solve for species p (polymer.cu:418-426 + accumulate, 434-447)
// ---------------------------------------------------------------------------
// Before the sweeps: FP32 class shifts and the Boltzmann factors e^{-W_c}.
template <typename T> int fts_engine::solve_begin(int p)
{
    using Cv = V2<T>;
    HostSpecies& sp = species[static_cast<size_t>(p)];
    last_species = p;
    const size_t Ms = static_cast<size_t>(M);
    Cv* e = static_cast<Cv*>(E);
    const int nc = static_cast<int>(sp.cls_type.size());
    // FP32: shift each class's potential so that max_i exp(-Re W_c + shift_c) = exp(BOLTZ_CAP)
    // (shift_c = min Re W_c + BOLTZ_CAP, kept in FP64 on the device)
    if (sizeof(T) == 4) {
        for (int c = 0; c < nc; ++c) {
            hipLaunchKernelGGL(k_min_class_w, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, f(W_type_s, sp.cls_type[static_cast<size_t>(c)]),
                               psi_s.p, electro() ? sp.cls_z[static_cast<size_t>(c)] : 0.0, M, partials.p);
            hipLaunchKernelGGL(k_finalize_min, dim3(1), dim3(RED_THREADS), 0, stream, partials.p, BOLTZ_CAP, sp.d_cls_shift.p + c);
            HIP_TRY(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(k_boltzmann<T>, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, W_type_s.p, psi_s.p, sp.d_cls_type.p, sp.d_cls_z.p,
                       sp.d_cls_shift.p, nc, M, e);
    HIP_TRY(hipGetLastError());
    if (sp.N == 1) {
        const Cv* e0 = e + static_cast<size_t>(sp.cls_of_bead[0]) * Ms;
        HIP_TRY(hipMemcpyAsync(qF, e0, sizeof(Cv) * Ms, hipMemcpyDeviceToDevice, stream));
        if (!sp.mirror) HIP_TRY(hipMemcpyAsync(qB, e0, sizeof(Cv) * Ms, hipMemcpyDeviceToDevice, stream));
    }
    return FTS_OK;
}

// After the sweeps: renormalisation factors, Q = <qF_{N-1}> (polymer.cu:425) and,
// with accumulate, the density contraction (polymer.cu:434-447).
template <typename T> int fts_engine::solve_end(int p, bool accumulate)
{
    using Cv = V2<T>;
    const HostSpecies& sp = species[static_cast<size_t>(p)];
    const size_t Ms = static_cast<size_t>(M);
    Cv* qf = static_cast<Cv*>(qF);
    Cv* qb = static_cast<Cv*>(qB);
    const int nc = static_cast<int>(sp.cls_type.size());
    const int N = sp.N;
    if (N > 1) {
        hipLaunchKernelGGL(k_scale_prefix, dim3(1), dim3(256), 2 * sizeof(double) * static_cast<size_t>(N), stream, logc.p, logc.p + Nmax, N,
                           sp.mirror ? 1 : 0, rfac.p, rfac2.p,
                           scal.p + SC_LF + p);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sum<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, qf + static_cast<size_t>(N - 1) * Ms, M, partials.p);
    HIP_TRY(hipGetLastError());
    int rc = reduce_finalize(1, SC_Q + p, 1.0 / static_cast<double>(M));
    if (rc) return rc;
    if (!accumulate) return FTS_OK;
    hipLaunchKernelGGL(k_contract_density<T>, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, qf, qb,
                       sp.mirror ? sp.d_half_start.p : sp.d_cls_start.p, sp.mirror ? sp.d_half_beads.p : sp.d_cls_beads.p, sp.d_cls_type.p,
                       sp.d_cls_z.p, sp.d_cls_shift.p, nc, W_type_s.p, psi_s.p, scal.p + SC_Q + p, static_cast<double>(sp.n) / V,
                       electro() ? 1 : 0, N > 1 ? rfac.p : nullptr, M, sp.mirror ? N : 0, rho_hat.p, rho_c_hat.p);
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

template int fts_engine::solve_begin<float>(int);
template int fts_engine::solve_begin<double>(int);
template int fts_engine::solve_end<float>(int, bool);
template int fts_engine::solve_end<double>(int, bool);

template <typename T> int fts_engine::solve_species(int p, bool accumulate)
{
    int rc = solve_begin<T>(p);
    if (rc) return rc;
    const int N = species[static_cast<size_t>(p)].N;
    if (N > 1 && (rc = run_sweeps<T>(p, sched, N - 1, timing))) return rc;
    return solve_end<T>(p, accumulate);
}

// complexLangevin::compute_densities (complexLangevin.cu:255-260)
int fts_engine::compute_densities(const double2* om, const double2* ps)
{
    int rc = densities_begin(om, ps);
    if (rc) return rc;
    for (int p = 0; p < static_cast<int>(species.size()); ++p) {
        rc = precision == FTS_PREC_FP32 ? solve_species<float>(p) : solve_species<double>(p);
        if (rc) return rc;
    }
    return densities_end();
}

// smoothed sources of the state (om, ps) and cleared density accumulators
int fts_engine::densities_begin(const double2* om, const double2* ps)
{
    const size_t Ms = static_cast<size_t>(M);
    const double invM = 1.0 / static_cast<double>(M);
    int rc;
    dens_psi = ps;
    // build_smoothed_sources (polymer.cu:341-363)
    for (int l = 0; l < n_omega; ++l)
        if ((rc = fft_apply(om + l * Ms, f(omega_s, l), kop_sep(sep_gamma.p, invM)))) return rc;
    hipLaunchKernelGGL(k_combine_modes, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, omega_s.p, mode_coef.p, n_omega, nT, M, W_type_s.p);
    HIP_TRY(hipGetLastError());
    if (electro()) {
        if (psi_hat.p) {
            if ((rc = fft_forward(ps, psi_hat.p)) || (rc = fft_apply_hat(psi_hat.p, psi_s.p, kop_sep(sep_gamma.p, invM)))) return rc;
            psi_hat_of = ps;
        } else if ((rc = fft_apply(ps, psi_s.p, kop_sep(sep_gamma.p, invM)))) {
            return rc;
        }
    } else {
        if ((rc = zero_async(psi_s.p, Ms))) return rc;
    }
    if ((rc = zero_async(rho_hat.p, rho_hat.n)) || (rc = zero_async(rho_c_hat.p, Ms))) return rc;
    return FTS_OK;
}

// smoothing of the densities (polymer.cu:467-476) and the ion densities (ion.cu:54-72)
int fts_engine::densities_end()
{
    const size_t Ms = static_cast<size_t>(M);
    const double invM = 1.0 / static_cast<double>(M);
    int rc;
    const double2* ps = dens_psi;
    // the smoothed type densities and the polymer charge density: one convolution each
    FftJob jb[FTS_MAX_TYPES + 1];
    int nj = 0;
    for (int al = 0; al < nT; ++al) jb[nj++] = FftJob{f(rho_hat, al), f(rho_type, al), kop_sep(sep_gamma.p, invM)};
    if (electro()) jb[nj++] = FftJob{rho_c_hat.p, rho_c_poly.p, kop_sep(sep_gamma.p, invM)};
    if ((rc = fft_apply_n(jb, nj))) return rc;
    hipLaunchKernelGGL(k_sum_fields, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, rho_type.p, nT, M, rho_p.p);
    HIP_TRY(hipGetLastError());
    if (electro()) {
        HIP_TRY(hipMemcpyAsync(rho_c.p, rho_c_poly.p, sizeof(double2) * Ms, hipMemcpyDeviceToDevice, stream));
    } else {
        if ((rc = zero_async(rho_c_poly.p, Ms)) || (rc = zero_async(rho_c.p, Ms))) return rc;
    }
    if ((rc = zero_async(rho_c_ion.p, Ms))) return rc;
    // ion::compute_densities (ion.cu:54-72)
    const size_t nI = ions.size();
    auto sep_of = [&](size_t j) { return sep_ion.p + j * static_cast<size_t>(m1 + m2 + m3); };
    if (!electro()) {
        for (size_t j = 0; j < nI; ++j) {
            hipLaunchKernelGGL(k_fill, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, f(ion_rho, j), double2{static_cast<double>(ions[j].n) / V, 0.0}, M);
            HIP_TRY(hipGetLastError());
        }
        return FTS_OK;
    }
    // the smoothed potential of every ion species (all from psi)
    FftJob ij[FTS_MAX_IONS];
    const bool hat = psi_hat_valid(ps);
    for (size_t j = 0; j < nI; ++j) ij[j] = FftJob{hat ? psi_hat.p : ps, f(ion_psi, j), kop_sep(sep_of(j), invM)};
    if ((rc = hat ? fft_apply_hat_n(ij, static_cast<int>(nI)) : fft_apply_n(ij, static_cast<int>(nI)))) return rc;
    for (size_t j = 0; j < nI; ++j) {
        hipLaunchKernelGGL(k_ion_boltzmann, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, f(ion_psi, j), ions[j].z, M, f(ion_b, j), partials.p);
        HIP_TRY(hipGetLastError());
        if ((rc = reduce_finalize(1, SC_QION + static_cast<int>(j), invM))) return rc;
        hipLaunchKernelGGL(k_ion_density, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, f(ion_b, j), scal.p + SC_QION + j,
                           static_cast<double>(ions[j].n) / V, M, f(ion_rho, j));
        HIP_TRY(hipGetLastError());
    }
    // ion charges z S_{Gamma_j} rho_j (fft_apply_real_kernel_scaled, common.h:1087-1091), two
    // species at a time through t1 and t2, added in species order
    double2* tmp[2] = {t1.p, t2.p};
    for (size_t j0 = 0; j0 < nI; j0 += 2) {
        const int c = static_cast<int>(std::min<size_t>(2, nI - j0));
        for (int u = 0; u < c; ++u) ij[u] = FftJob{f(ion_rho, j0 + u), tmp[u], kop_sep(sep_of(j0 + u), ions[j0 + u].z * invM)};
        if ((rc = fft_apply_n(ij, c))) return rc;
        for (int u = 0; u < c; ++u) {
            hipLaunchKernelGGL(k_add2, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, rho_c_ion.p, rho_c.p, tmp[u], M);
            HIP_TRY(hipGetLastError());
        }
    }
    return FTS_OK;
}

// u_psi = S_{k^2} psi / (4 pi lB) + i rho, minus its mean (complexLangevin.cu:299-311)
int fts_engine::psi_drift(const double2* ps, const double2* rho_src, double2* out)
{
    if (!electro()) return zero_async(out, static_cast<size_t>(M));
    const KOp lap{nullptr, nullptr, 0.0, 1.0 / static_cast<double>(M), 0.0, 0, 0};
    int rc = psi_hat_valid(ps) ? fft_apply_hat(psi_hat.p, t2.p, lap) : fft_apply(ps, t2.p, lap);
    if (rc) return rc;
    hipLaunchKernelGGL(k_psi_drift, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, t2.p, rho_src, 1.0 / (4.0 * Pi * lB), M, out);
    HIP_TRY(hipGetLastError());
    return remove_mean(out);
}

int fts_engine::key_queue()
{
    hkey[0] = seed;
    hkey[1] = static_cast<unsigned long long>(step_counter);
    HIP_TRY(hipMemcpyAsync(nkey.p, hkey, 2 * sizeof(unsigned long long), hipMemcpyHostToDevice, stream));
    return FTS_OK;
}

int fts_engine::zero_async(double2* p, size_t n)
{
    if (n == 0) return FTS_OK;
    hipLaunchKernelGGL(k_fill, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream, p, double2{0.0, 0.0}, static_cast<int>(n));
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

int fts_engine::finite_queue(int* host_out)
{
    if (inject_step >= 0 && step_counter == inject_step) {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(EW_THREADS), 0, stream, psi.p, double2{std::nan(""), 0.0}, 1);
        psi_hat_of = nullptr;
        inject_step = -1;
    }
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, stream, flag.p, 1);
    const size_t no = static_cast<size_t>(n_omega) * M;
    if (no) hipLaunchKernelGGL(k_finite, dim3(ew_blocks(no)), dim3(EW_THREADS), 0, stream, omega.p, no, flag.p);
    hipLaunchKernelGGL(k_finite, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, psi.p, static_cast<size_t>(M), flag.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host_out, flag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    return FTS_OK;
}

int fts_engine::q_queue(double2* host_out)
{
    HIP_TRY(hipMemcpyAsync(host_out, scal.p + SC_Q, sizeof(double2) * Q_SLOTS, hipMemcpyDeviceToHost, stream));
    return FTS_OK;
}

int fts_engine::check_finite(bool& ok)
{
    int h = 0;
    int rc = finite_queue(&h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    ok = h != 0;
    return FTS_OK;
}

// dst[c] = sum_k w_c(k) S(k) A(k) B(-k), c = xx yy zz xy xz yz 1 (device slots)
int fts_engine::contract_to(const void* A, const void* B, bool fp32, const double* sep, double2* dst)
{
    if (fp32)
        hipLaunchKernelGGL(k_contract_kspace<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, static_cast<const float2*>(A),
                           static_cast<const float2*>(B), kvec.p, sep, m1, m2, m3, partials.p);
    else
        hipLaunchKernelGGL(k_contract_kspace<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, static_cast<const double2*>(A),
                           static_cast<const double2*>(B), kvec.p, sep, m1, m2, m3, partials.p);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RED_THREADS), 0, stream, partials.p, 7, dst, 1.0);
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

size_t fts_engine::stress_chain_base(int p) const
{
    size_t base = 0;
    for (int q = 0; q < p; ++q) base += static_cast<size_t>(std::max(1, species[static_cast<size_t>(q)].N - 1));
    return base;
}

// Chain-connectivity sums of species p into stress_acc rows base + s:
// sum_k w_c phi_s(k) FFT[qB_{s+1}](k) FFT[qF_s](-k), the k-space form of
// <qF_s S_{-k_a k_b}[e^{W_s} qB_s]> with FFT[e^{W_s} qB_s] = phi_s FFT[qB_{s+1}].
template <typename T> int fts_engine::stress_chain(int p)
{
    using Cv = V2<T>;
    const HostSpecies& sp = species[static_cast<size_t>(p)];
    const size_t Ms = static_cast<size_t>(M);
    Cv* qf = static_cast<Cv*>(qF);
    Cv* qb = static_cast<Cv*>(qB);
    double2* rows = stress_acc.p + stress_chain_base(p) * 7;
    for (int s = 0; s + 1 < sp.N; ++s) {
        // both spectra of bead s in one pair of launches
        int rc = fft_forward_t<T>(qf + static_cast<size_t>(s) * Ms, static_cast<Cv*>(kA),
                                  sp.mirror ? qf + static_cast<size_t>(sp.N - 2 - s) * Ms : qb + static_cast<size_t>(s + 1) * Ms, static_cast<Cv*>(kB));
        if (rc) return rc;
        const double* sep = sep_bond.p + static_cast<size_t>(sp.bond_cls[static_cast<size_t>(s)]) * (m1 + m2 + m3);
        if ((rc = contract_to(kB, kA, sizeof(T) == 4, sep, rows + static_cast<size_t>(s) * 7))) return rc;
    }
    return FTS_OK;
}

namespace {
double2 cinv_host(double2 q)
{
    // Smith's scaled division (complex_inverse, common.h:127-139)
    const double ax = std::fabs(q.x), ay = std::fabs(q.y);
    if (ax >= ay) {
        const double r = q.y / q.x, d = q.x * (1.0 + r * r);
        return double2{1.0 / d, -r / d};
    }
    const double r = q.x / q.y, d = q.y * (1.0 + r * r);
    return double2{r / d, -1.0 / d};
}
} // namespace

// observable::measure (observable.cu:53-67) at (om, ps), after compute_densities at
// the same state.  All device reductions are queued first; their results, the
// scalar slots and the renormalisation data come back in one copy to pinned memory.
int fts_engine::measure(const double2* om, const double2* ps, fts_observables* rec)
{
    int rc = measure_queue(om, ps, rec);
    if (rc || defer_measure) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    return measure_finish(rec);
}

// the host fields of a record, set when its device work is queued
void fts_engine::measure_header(fts_observables* rec) const
{
    std::memset(rec, 0, sizeof(*rec));
    rec->step = step_counter;
    rec->n_species = static_cast<int>(species.size());
    rec->n_ions = static_cast<int>(ions.size());
    rec->stress_mask = stress_mask;
}

int fts_engine::measure_queue(const double2* om, const double2* ps, fts_observables* rec)
{
    measure_header(rec);
    const double invM = 1.0 / static_cast<double>(M);
    const size_t nsp = species.size();
    const bool want_stress = (stress_mask & 0x7F) != 0;
    int rc;
    // ---- queue ----------------------------------------------------------------
    // H/V field terms: (c_l/2) <omega_l^2> and (1/(8 pi lB)) <psi S_{k^2} psi>  (observable.cu:160-182)
    for (int l = 0; l < n_omega; ++l) {
        const double2* w = om + static_cast<size_t>(l) * M;
        hipLaunchKernelGGL(k_sum_product, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, w, w, M, partials.p);
        HIP_TRY(hipGetLastError());
        if ((rc = reduce_finalize(1, SC_H + l, 0.5 * omega_coefficient(l) * invM))) return rc;
    }
    if (electro()) {
        const double2* hat = t3.p;
        if (psi_hat_valid(ps)) hat = psi_hat.p;
        else if ((rc = fft_forward(ps, t3.p))) return rc;
        if ((rc = contract_to(hat, hat, false, nullptr, scal.p + SC_PSI_S))) return rc;
    }
    // pinned staging: [scal][stress_acc][rfac2: nsp x Nmax doubles][shifts: nsp x MAX_CLASSES doubles]
    double2* h_scal = static_cast<double2*>(pin);
    double2* h_acc = h_scal + SC_COUNT;
    double* h_rfac2 = reinterpret_cast<double*>(h_acc + stress_acc.n);
    double* h_shift = h_rfac2 + nsp * static_cast<size_t>(Nmax);
    const size_t fixed_rows = stress_chain_base(static_cast<int>(nsp)); // rows after the chains
    if (want_stress) {
        for (size_t p = 0; p < nsp; ++p) {
            if (nsp > 1) {
                // the propagator store holds one species: re-solve species p at the same state
                rc = precision == FTS_PREC_FP32 ? solve_species<float>(static_cast<int>(p), false) : solve_species<double>(static_cast<int>(p), false);
                if (rc) return rc;
            }
            rc = precision == FTS_PREC_FP32 ? stress_chain<float>(static_cast<int>(p)) : stress_chain<double>(static_cast<int>(p));
            if (rc) return rc;
            if (species[p].N > 1)
                HIP_TRY(hipMemcpyAsync(h_rfac2 + p * Nmax, rfac2.p, sizeof(double) * (species[p].N - 1), hipMemcpyDeviceToHost, stream));
        }
        // polymer field coupling (one row per type) and, with electrostatics, the psi row and one row per ion
        double2* fixed = stress_acc.p + fixed_rows * 7;
        auto pair = [&](const double2* a, const double2* b, size_t row) {
            const FftJob jb[2] = {{a, t1.p, KOp{}}, {b, t2.p, KOp{}}};
            int r = fft_forward_n(jb, 2);
            return r ? r : contract_to(t1.p, t2.p, false, nullptr, fixed + row * 7);
        };
        for (int al = 0; al < nT; ++al)
            if ((rc = pair(f(W_type_s, al), f(rho_hat, al), static_cast<size_t>(al)))) return rc;
        if (electro()) {
            if ((rc = pair(psi_s.p, rho_c_hat.p, static_cast<size_t>(nT)))) return rc;
            // the ion rows contract psi~_j with the normalised ion density rho_j = (n_j/V) b_j / Q_j,
            // not with b_j: the same sum by linearity, and b_j = e^{-i z_j psi~_j} leaves the FP32
            // range of the FP32 form's transforms where |Im psi~_j| is large (~e^184 measured in a
            // rare C2 excursion) while rho_j stays of the order of the density
            for (size_t j = 0; j < ions.size(); ++j)
                if ((rc = pair(f(ion_psi, j), f(ion_rho, j), static_cast<size_t>(nT) + 1 + j))) return rc;
        }
        HIP_TRY(hipMemcpyAsync(h_acc, stress_acc.p, sizeof(double2) * stress_acc.n, hipMemcpyDeviceToHost, stream));
    }
    HIP_TRY(hipMemcpyAsync(h_scal, scal.p, sizeof(double2) * SC_COUNT, hipMemcpyDeviceToHost, stream));
    for (size_t p = 0; p < nsp; ++p)
        HIP_TRY(hipMemcpyAsync(h_shift + p * MAX_CLASSES, species[p].d_cls_shift.p, sizeof(double) * species[p].cls_shift.size(),
                               hipMemcpyDeviceToHost, stream));
    return FTS_OK;
}

// host arithmetic of measure over the pinned copies measure_queue made (after a sync)
int fts_engine::measure_finish(fts_observables* rec)
{
    const double invM = 1.0 / static_cast<double>(M);
    const double invM2 = invM * invM;
    const size_t nsp = species.size();
    const bool want_stress = (stress_mask & 0x7F) != 0;
    const double2* h_scal = static_cast<const double2*>(pin);
    const double2* h_acc = h_scal + SC_COUNT;
    const double* h_rfac2 = reinterpret_cast<const double*>(h_acc + stress_acc.n);
    const double* h_shift = h_rfac2 + nsp * static_cast<size_t>(Nmax);
    const size_t fixed_rows = stress_chain_base(static_cast<int>(nsp));
    const double2* Q = h_scal + SC_Q;
    const double2* Qi = h_scal + SC_QION;
    for (size_t p = 0; p < nsp; ++p) {
        HostSpecies& sp = species[p];
        q_log_gain[p] = h_scal[SC_LF + p].x;
        double total = 0.0;
        for (size_t c = 0; c < sp.cls_shift.size(); ++c) sp.cls_shift[c] = h_shift[p * MAX_CLASSES + c];
        for (int s = 0; s < sp.N; ++s) total += sp.cls_shift[static_cast<size_t>(sp.cls_of_bead[static_cast<size_t>(s)])];
        q_shift_total[p] = total;
    }
    // |z| by hypot: finite for every finite z (the squares overflow from |z| ~ 1e154)
    auto neglog = [](double2 z) { return double2{-std::log(std::hypot(z.x, z.y)), -std::atan2(z.y, z.x)}; };
    double2 h{0.0, 0.0};
    for (size_t p = 0; p < nsp; ++p) {
        double2 mu = neglog(Q[p]);
        mu.x += q_shift_total[p] + q_log_gain[p]; // Q = Q~ e^{-(sum of shifts + LF(N-1))}
        const double scale_q = std::exp(-q_shift_total[p] - q_log_gain[p]);
        rec->Q_poly[p][0] = Q[p].x * scale_q;
        rec->Q_poly[p][1] = Q[p].y * scale_q;
        rec->mu_ex_poly[p][0] = mu.x;
        rec->mu_ex_poly[p][1] = mu.y;
        rec->mu_id_poly[p] = std::log(static_cast<double>(species[p].n) / V);
        rec->mu_self_poly[p] = -lB * species[p].charge_square_sum() / (2.0 * std::sqrt(Pi) * a);
        const double w = static_cast<double>(species[p].n) / V;
        h.x += w * mu.x;
        h.y += w * mu.y;
    }
    for (size_t j = 0; j < ions.size(); ++j) {
        double2 mu{0.0, 0.0}, Qj{1.0, 0.0};
        if (electro()) {
            Qj = Qi[j];
            mu = neglog(Qj);
        }
        rec->Q_ion[j][0] = Qj.x;
        rec->Q_ion[j][1] = Qj.y;
        rec->mu_ex_ion[j][0] = mu.x;
        rec->mu_ex_ion[j][1] = mu.y;
        rec->mu_id_ion[j] = std::log(static_cast<double>(ions[j].n) / V);
        rec->mu_self_ion[j] = -lB * ions[j].z * ions[j].z / (2.0 * std::sqrt(Pi) * ions[j].a);
        const double w = static_cast<double>(ions[j].n) / V;
        h.x += w * mu.x;
        h.y += w * mu.y;
    }
    for (int l = 0; l < n_omega; ++l) {
        h.x += h_scal[SC_H + l].x;
        h.y += h_scal[SC_H + l].y;
    }
    const double2* S_psi = h_scal + SC_PSI_S;
    if (electro()) {
        const double sc = 1.0 / (8.0 * Pi * lB) * invM2;
        h.x += sc * (S_psi[0].x + S_psi[1].x + S_psi[2].x);
        h.y += sc * (S_psi[0].y + S_psi[1].y + S_psi[2].y);
    }
    rec->H[0] = h.x;
    rec->H[1] = h.y;
    if (!want_stress) return FTS_OK;

    // stress tensor (stress.cu:161-321) in k-space
    double2 sig[6];
    long long ion_count = 0, chains = 0;
    for (const HostIon& ion : ions) ion_count += ion.n;
    for (const HostSpecies& sp : species) chains += sp.n;
    const double ideal = (static_cast<double>(chains) + static_cast<double>(ion_count)) / V;
    for (int c = 0; c < 6; ++c) sig[c] = double2{(ideal + P_sr) * (c < 3 ? 1.0 : 0.0) + mode_term[c], 0.0};
    if (electro()) {
        const double sc = 1.0 / (4.0 * Pi * lB) * invM2;
        const double2 k2s{S_psi[0].x + S_psi[1].x + S_psi[2].x, S_psi[0].y + S_psi[1].y + S_psi[2].y};
        for (int c = 0; c < 6; ++c) {
            const double d = c < 3 ? 1.0 / 3.0 : 0.0;
            sig[c].x += sc * (S_psi[c].x - d * k2s.x);
            sig[c].y += sc * (S_psi[c].y - d * k2s.y);
        }
    }
    // chain connectivity: + (n/(V Q M^2)) sum_s (b_s^2/3) rfac2_s sum_k k_a k_b phi_s qB^_{s+1}(k) qF^_s(-k)
    for (size_t p = 0; p < nsp; ++p) {
        const HostSpecies& sp = species[p];
        const double2* rows = h_acc + stress_chain_base(static_cast<int>(p)) * 7;
        double2 tot[6] = {};
        for (int s = 0; s + 1 < sp.N; ++s) {
            const double w = sp.bond[static_cast<size_t>(s)] * sp.bond[static_cast<size_t>(s)] / 3.0 * h_rfac2[p * Nmax + s];
            for (int c = 0; c < 6; ++c) {
                tot[c].x += w * rows[static_cast<size_t>(s) * 7 + c].x;
                tot[c].y += w * rows[static_cast<size_t>(s) * 7 + c].y;
            }
        }
        // prefactor (n / V) / Q~ (the shifts cancel between qF_s qB_{s+1} and Q~)
        const double2 inv = cinv_host(Q[p]);
        const double sc = static_cast<double>(sp.n) / V * invM2;
        const double2 pref{inv.x * sc, inv.y * sc};
        for (int c = 0; c < 6; ++c) {
            sig[c].x += pref.x * tot[c].x - pref.y * tot[c].y;
            sig[c].y += pref.x * tot[c].y + pref.y * tot[c].x;
        }
    }
    // polymer field coupling: sum_alpha [ -a^2/M^2 S_ab(W^_a, P^_a) + (delta/2) S_1/M^2 ] (+ psi terms)
    const double a2 = a * a;
    const double2* fixed = h_acc + fixed_rows * 7;
    for (int al = 0; al < nT; ++al) {
        const double2* S = fixed + static_cast<size_t>(al) * 7;
        for (int c = 0; c < 6; ++c) {
            const double d = c < 3 ? 0.5 : 0.0;
            sig[c].x += invM2 * (-a2 * S[c].x + d * S[6].x);
            sig[c].y += invM2 * (-a2 * S[c].y + d * S[6].y);
        }
    }
    if (electro()) {
        const double2* S = fixed + static_cast<size_t>(nT) * 7;
        for (int c = 0; c < 6; ++c) {
            const double d = c < 3 ? 1.0 / 6.0 : 0.0;
            const double2 v{invM2 * (-a2 * S[c].x + d * S[6].x), invM2 * (-a2 * S[c].y + d * S[6].y)};
            sig[c].x += -v.y; // i * v
            sig[c].y += v.x;
        }
        // ions: i (n_j z_j / (V Q_j)) [ -a_j^2/M^2 S_ab(psi~^_j, b^_j) + (delta/6) S_1/M^2 ], with
        // (n_j / (V Q_j)) b_j = rho_j contracted directly: i z_j [ ... S(psi~^_j, rho^_j) ... ]
        for (size_t j = 0; j < ions.size(); ++j) {
            const double2* Sj = fixed + (static_cast<size_t>(nT) + 1 + j) * 7;
            const double2 pref{0.0, ions[j].z}; // i z_j
            const double aj2 = ions[j].a * ions[j].a;
            for (int c = 0; c < 6; ++c) {
                const double d = c < 3 ? 1.0 / 6.0 : 0.0;
                const double2 v{invM2 * (-aj2 * Sj[c].x + d * Sj[6].x), invM2 * (-aj2 * Sj[c].y + d * Sj[6].y)};
                sig[c].x += pref.x * v.x - pref.y * v.y;
                sig[c].y += pref.x * v.y + pref.y * v.x;
            }
        }
    }
    for (int c = 0; c < 6; ++c)
        if (stress_mask & (1 << c)) {
            rec->stress[c][0] = sig[c].x;
            rec->stress[c][1] = sig[c].y;
        }
    if (stress_mask & FTS_STRESS_PISO) {
        // p_iso = tr(sigma)/3 (observable.cu:297-317 is the same sum with k_a k_b -> k^2/3)
        rec->p_iso[0] = (sig[0].x + sig[1].x + sig[2].x) / 3.0;
        rec->p_iso[1] = (sig[0].y + sig[1].y + sig[2].y) / 3.0;
    }
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// CL steps
// ---------------------------------------------------------------------------
namespace {
bool finite2(double2 z) { return std::isfinite(z.x) && std::isfinite(z.y); }
} // namespace

// A step whose polymer or ion partition functions are not finite fails (an ion Q that
// overflows leaves finite fields but zero ion densities and H = -inf for the step)
int fts_engine::check_Q(const char* what)
{
    double2 Q[Q_SLOTS];
    int rc = read_scal(SC_Q, Q_SLOTS, Q);
    if (rc) return rc;
    if (!partition_finite(Q)) return fail(FTS_NONFINITE, std::string("non-finite ") + what + " at step " + std::to_string(step_counter + 1));
    return FTS_OK;
}

bool fts_engine::partition_finite(const double2* Q) const
{
    for (size_t p = 0; p < species.size(); ++p)
        if (!finite2(Q[SC_Q + p])) return false;
    if (electro())
        for (size_t j = 0; j < ions.size(); ++j)
            if (!finite2(Q[SC_QION + j])) return false;
    return true;
}

// run_step (complexLangevin.cu:669-698): predictor-corrector with the same noise in both halves
int fts_engine::step_pc(bool record, fts_observables* rec)
{
    int rc = compute_densities(omega.p, psi.p);
    if (rc || (rc = check_Q("Q")) || (rc = pc_predict())) return rc;
    if ((rc = compute_densities(omega_bar.p, psi_bar.p)) || (rc = check_Q("Qbar"))) return rc;
    if ((rc = pc_correct(record, rec))) return rc;
    return finish_step();
}

// predictor (complexLangevin.cu:328-345), after the densities at (w, psi)
int fts_engine::pc_predict()
{
    if (psi_hat_of == psi_bar.p) psi_hat_of = nullptr; // psi_bar is rewritten below
    const size_t Ms = static_cast<size_t>(M);
    int rc;
    if ((rc = key_queue())) return rc; // the step's noise key (the corrector draws the same eta)
    const double sigma_psi = electro() ? (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_psi * dt / dV) : 0.0;
    if (electro()) {
        hipLaunchKernelGGL(k_sum_noise, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, nkey.p, 0x100u, sigma_psi, M, partials.p);
        HIP_TRY(hipGetLastError());
        if ((rc = reduce_finalize(1, SC_ETA_MEAN, 1.0 / static_cast<double>(M)))) return rc;
    }
    if (n_omega > 0) {
        hipLaunchKernelGGL(k_omega_predict, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, mc, omega.p, rho_type.p, nkey.p, M, drift.p, omega_bar.p);
        HIP_TRY(hipGetLastError());
    }
    // predictor psi drift uses rho_c of the current fields (complexLangevin.cu:331, 674)
    if ((rc = psi_drift(psi.p, rho_c.p, dpsi.p))) return rc;
    if (electro()) {
        hipLaunchKernelGGL(k_psi_update, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, psi.p, dpsi.p, -dt * lambda_psi,
                           static_cast<const double2*>(nullptr), 0.0, sigma_psi, scal.p + SC_ETA_MEAN, nkey.p, M, psi_bar.p);
        HIP_TRY(hipGetLastError());
        if ((rc = remove_mean(psi_bar.p))) return rc;
    } else if ((rc = zero_async(psi_bar.p, Ms))) {
        return rc;
    }
    return FTS_OK;
}

// corrector (complexLangevin.cu:347-368), after the densities at (w_bar, psi_bar);
// the record is taken at the predictor state (complexLangevin.cu:680-684)
int fts_engine::pc_correct(bool record, fts_observables* rec)
{
    const size_t Ms = static_cast<size_t>(M);
    int rc;
    const double sigma_psi = electro() ? (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_psi * dt / dV) : 0.0;
    if ((rc = psi_drift(psi_bar.p, rho_c.p, dpsi_bar.p))) return rc;
    if (drift_stats && (rc = drift_queue(true))) return rc;
    if (record && (rc = measure(omega_bar.p, psi_bar.p, rec))) return rc;
    if (n_omega > 0) {
        hipLaunchKernelGGL(k_omega_correct, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, mc, drift.p, omega_bar.p, rho_type.p, nkey.p, M, omega.p);
        HIP_TRY(hipGetLastError());
    }
    psi_hat_of = nullptr; // psi changes below
    if (electro()) {
        hipLaunchKernelGGL(k_psi_update, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, psi.p, dpsi.p, -0.5 * dt * lambda_psi,
                           static_cast<const double2*>(dpsi_bar.p), -0.5 * dt * lambda_psi, sigma_psi, scal.p + SC_ETA_MEAN, nkey.p, M, psi.p);
        HIP_TRY(hipGetLastError());
        if ((rc = remove_mean(psi.p))) return rc;
    } else if ((rc = zero_async(psi.p, Ms))) {
        return rc;
    }
    return FTS_OK;
}

// write_drift_max / corrector_drift_max (complexLangevin.cu:318-324, 474-497): per mode
// |lam_l / 2| max |u_l + u_bar_l| and |lam_psi / 2| max |u_psi + u_psi_bar| after the
// corrector's drifts (pc); EM and semi-implicit: |lam| max |u| at the state the step starts from
int fts_engine::drift_queue(bool pc)
{
    for (int l = 0; l < n_omega; ++l) {
        const double lam = lambda_omega[static_cast<size_t>(l)];
        hipLaunchKernelGGL(k_mode_drift_max2, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, mc, l, pc ? drift.p : nullptr,
                           pc ? omega_bar.p : omega.p, rho_type.p, M, partials.p);
        hipLaunchKernelGGL(k_finalize_max, dim3(1), dim3(RED_THREADS), 0, stream, partials.p, pc ? 0.5 * lam : lam, scal.p + SC_DRIFT + l);
    }
    hipLaunchKernelGGL(k_sum_max2, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, dpsi.p, pc ? static_cast<const double2*>(dpsi_bar.p) : nullptr,
                       M, partials.p);
    hipLaunchKernelGGL(k_finalize_max, dim3(1), dim3(RED_THREADS), 0, stream, partials.p, pc ? 0.5 * lambda_psi : lambda_psi,
                       scal.p + SC_DRIFT + n_omega);
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

int fts_engine::finish_step()
{
    ++step_counter;
    bool ok = true;
    int rc = check_finite(ok);
    if (rc) return rc;
    if (!ok) return fail(FTS_NONFINITE, "updated fields have non-finite values at step " + std::to_string(step_counter));
    return FTS_OK;
}

// Plain Euler-Maruyama, and the semi-implicit variant w <- w + S_P[-dt lam u + eta],
// P(k) = 1/(1 + dt lam A(k)).  Records are taken at the state the step starts from.
int fts_engine::step_em(bool record, fts_observables* rec, bool si)
{
    int rc = compute_densities(omega.p, psi.p);
    if (rc || (rc = check_Q("Q")) || (rc = em_update(record, rec, si))) return rc;
    return finish_step();
}

int fts_engine::em_update(bool record, fts_observables* rec, bool si)
{
    const size_t Ms = static_cast<size_t>(M);
    const double invM = 1.0 / static_cast<double>(M);
    int rc;
    if (record && (rc = measure(omega.p, psi.p, rec))) return rc;
    if ((rc = key_queue())) return rc; // the step's noise key
    const double sigma_psi = electro() ? (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_psi * dt / dV) : 0.0;
    // EM adds real-space noise (psi's mean-free); the semi-implicit step adds it in k-space
    if (electro() && !si) {
        hipLaunchKernelGGL(k_sum_noise, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, nkey.p, 0x100u, sigma_psi, M, partials.p);
        HIP_TRY(hipGetLastError());
        if ((rc = reduce_finalize(1, SC_ETA_MEAN, invM))) return rc;
    }
    if ((rc = psi_drift(psi.p, rho_c.p, dpsi.p))) return rc;
    if (drift_stats && (rc = drift_queue(false))) return rc;
    if (si && lrk_active() && (rc = lrk_step())) return rc;
    if (n_omega > 0) {
        hipLaunchKernelGGL(k_omega_increment, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, mc, rho_type.p, nkey.p, M, si ? 0 : 1, omega.p, drift.p,
                           si ? 0 : 1);
        HIP_TRY(hipGetLastError());
    }
    if (electro()) {
        hipLaunchKernelGGL(k_psi_increment, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, dpsi.p, -dt * lambda_psi, sigma_psi,
                           scal.p + SC_ETA_MEAN, nkey.p, M, dpsi.p, si ? 0 : 1);
        HIP_TRY(hipGetLastError());
    }
    if (si) {
        // w += S_P[-dt lam u] + S_Pn[eta]: the multiply pass of each increment's transform
        // adds sigma sqrt(M) Pn(k) g(k) / M (kop_noise), g the spectrum of the step's white
        // noise; psi's drops k = 0 (mean-free noise).  Independent fields share launches.
        const double rM = 1.0 / std::sqrt(static_cast<double>(M));
        FftJob jb[FTS_MAX_TYPES + 1];
        int nj = 0;
        for (int l = 0; l < n_omega; ++l) {
            KOp op{nullptr, si_table_omega.p + l * Ms, invM, 0.0, 0.0, 0, 0};
            op.nkey = nkey.p;
            op.nfield = SI_NOISE_FIELD + 2u * static_cast<unsigned>(l);
            op.namp = mc.sigma[l] * rM;
            jb[nj++] = FftJob{f(drift, l), f(drift, l), op};
        }
        if (electro()) {
            KOp op{nullptr, si_table_psi.p, invM, 0.0, 0.0, 0, 0};
            op.nkey = nkey.p;
            op.nfield = SI_NOISE_FIELD_PSI;
            op.namp = sigma_psi * rM;
            op.nzero = 1;
            jb[nj++] = FftJob{dpsi.p, dpsi.p, op};
        }
        if ((rc = fft_apply_n(jb, nj))) return rc;
        if (n_omega > 0) {
            const size_t n = static_cast<size_t>(n_omega) * Ms;
            hipLaunchKernelGGL(k_axpy1, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream, omega.p, drift.p, n);
            HIP_TRY(hipGetLastError());
        }
    }
    if (electro()) {
        psi_hat_of = nullptr; // psi changes below
        hipLaunchKernelGGL(k_axpy1, dim3(ew_blocks(Ms)), dim3(EW_THREADS), 0, stream, psi.p, dpsi.p, Ms);
        HIP_TRY(hipGetLastError());
        if ((rc = remove_mean(psi.p))) return rc;
    } else if ((rc = zero_async(psi.p, Ms))) {
        return rc;
    }
    return FTS_OK;
}

// Semi-implicit predictor-corrector (FTS_SCHEME_SIPC / _SIPC_LEARNED; no reference
// counterpart), with P = 1 / (1 + a), a = dt lam A(k) from this step's tables, and
// P2 = 2 P / (1 + P) = 1 / (1 + a/2):
//   predictor  d = S_P[-dt lam u(w)] + S_Pn[eta],  w_bar = w + d        (the semi-implicit step)
//   corrector  w <- w + S_P2[-(dt lam / 2)(u(w) + u(w_bar))] + S_P2[eta] + S_{1-P2}[d]
// with one k-space draw of eta for both.  For a linear drift u = A w this is
// Crank-Nicolson, w <- (1 - a/2)/(1 + a/2) w + P2 eta, whose stationary variance is 1 / A
// at any dt.  The record is taken at the predictor state (as the predictor-corrector's,
// complexLangevin.cu:680-684).
int fts_engine::step_sipc(bool record, fts_observables* rec)
{
    int rc = compute_densities(omega.p, psi.p);
    if (rc || (rc = check_Q("Q")) || (rc = sipc_predict())) return rc;
    if ((rc = compute_densities(omega_bar.p, psi_bar.p)) || (rc = check_Q("Qbar"))) return rc;
    if ((rc = sipc_correct(record, rec))) return rc;
    return finish_step();
}

// predictor, after the densities at (w, psi): drift and dpsi keep -dt lam u(w), sipc_d and
// sipc_dpsi the increments d
int fts_engine::sipc_predict()
{
    if (psi_hat_of == psi_bar.p) psi_hat_of = nullptr; // psi_bar is rewritten below
    const size_t Ms = static_cast<size_t>(M);
    const double invM = 1.0 / static_cast<double>(M), rM = 1.0 / std::sqrt(static_cast<double>(M));
    int rc;
    if ((rc = key_queue())) return rc; // the step's noise key (the corrector draws the same eta)
    const double sigma_psi = electro() ? (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_psi * dt / dV) : 0.0;
    if ((rc = psi_drift(psi.p, rho_c.p, dpsi.p))) return rc;
    if (drift_stats && (rc = drift_queue(false))) return rc;
    if (lrk_active() && (rc = lrk_step())) return rc;
    if (n_omega > 0) {
        hipLaunchKernelGGL(k_omega_increment, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, mc, rho_type.p, nkey.p, M, 0, omega.p, drift.p, 0);
        HIP_TRY(hipGetLastError());
    }
    if (electro()) {
        hipLaunchKernelGGL(k_psi_increment, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, dpsi.p, -dt * lambda_psi, sigma_psi,
                           scal.p + SC_ETA_MEAN, nkey.p, M, dpsi.p, 0);
        HIP_TRY(hipGetLastError());
    }
    FftJob jb[FTS_MAX_TYPES + 1];
    int nj = 0;
    for (int l = 0; l < n_omega; ++l) {
        KOp op{nullptr, si_table_omega.p + l * Ms, invM, 0.0, 0.0, 0, 0};
        op.nkey = nkey.p;
        op.nfield = SI_NOISE_FIELD + 2u * static_cast<unsigned>(l);
        op.namp = mc.sigma[l] * rM;
        jb[nj++] = FftJob{f(drift, l), f(sipc_d, l), op};
    }
    if (electro()) {
        KOp op{nullptr, si_table_psi.p, invM, 0.0, 0.0, 0, 0};
        op.nkey = nkey.p;
        op.nfield = SI_NOISE_FIELD_PSI;
        op.namp = sigma_psi * rM;
        op.nzero = 1;
        jb[nj++] = FftJob{dpsi.p, sipc_dpsi.p, op};
    }
    if ((rc = fft_apply_n(jb, nj))) return rc;
    if (n_omega > 0) {
        const size_t n = static_cast<size_t>(n_omega) * Ms;
        hipLaunchKernelGGL(k_add_out, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream, omega.p, sipc_d.p, n, omega_bar.p);
        HIP_TRY(hipGetLastError());
    }
    if (electro()) {
        hipLaunchKernelGGL(k_add_out, dim3(ew_blocks(Ms)), dim3(EW_THREADS), 0, stream, psi.p, sipc_dpsi.p, Ms, psi_bar.p);
        HIP_TRY(hipGetLastError());
        if ((rc = remove_mean(psi_bar.p))) return rc;
    } else if ((rc = zero_async(psi_bar.p, Ms))) {
        return rc;
    }
    return FTS_OK;
}

// corrector, after the densities at (w_bar, psi_bar); the record is taken at the predictor state
int fts_engine::sipc_correct(bool record, fts_observables* rec)
{
    const size_t Ms = static_cast<size_t>(M);
    const double invM = 1.0 / static_cast<double>(M), rM = 1.0 / std::sqrt(static_cast<double>(M));
    int rc;
    const double sigma_psi = electro() ? (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_psi * dt / dV) : 0.0;
    if ((rc = psi_drift(psi_bar.p, rho_c.p, dpsi_bar.p))) return rc;
    if (record && (rc = measure(omega_bar.p, psi_bar.p, rec))) return rc;
    // r = (-dt lam u(w) - dt lam u(w_bar)) / 2 in place of -dt lam u(w)
    if (n_omega > 0) {
        hipLaunchKernelGGL(k_omega_midpoint, dim3(ew_blocks(M)), dim3(EW_THREADS), 0, stream, mc, omega_bar.p, rho_type.p, M, drift.p);
        HIP_TRY(hipGetLastError());
    }
    if (electro()) {
        hipLaunchKernelGGL(k_midpoint, dim3(ew_blocks(Ms)), dim3(EW_THREADS), 0, stream, dpsi_bar.p, -dt * lambda_psi, Ms, dpsi.p);
        HIP_TRY(hipGetLastError());
    }
    // S_P2[r] + S_P2[eta] and S_{1-P2}[d], every field's in the same launches
    FftJob jb[2 * (FTS_MAX_TYPES + 1)];
    int nj = 0;
    for (int l = 0; l < n_omega; ++l) {
        KOp op{nullptr, si_table_omega.p + l * Ms, invM, 0.0, 0.0, 0, 0};
        op.table_form = KOP_CN;
        op.nkey = nkey.p;
        op.nfield = SI_NOISE_FIELD + 2u * static_cast<unsigned>(l);
        op.namp = mc.sigma[l] * rM;
        op.ncn = 1;
        jb[nj++] = FftJob{f(drift, l), f(drift, l), op};
        KOp od{nullptr, si_table_omega.p + l * Ms, invM, 0.0, 0.0, 0, 0};
        od.table_form = KOP_CNR;
        jb[nj++] = FftJob{f(sipc_d, l), f(sipc_d, l), od};
    }
    if (electro()) {
        KOp op{nullptr, si_table_psi.p, invM, 0.0, 0.0, 0, 0};
        op.table_form = KOP_CN;
        op.nkey = nkey.p;
        op.nfield = SI_NOISE_FIELD_PSI;
        op.namp = sigma_psi * rM;
        op.nzero = 1;
        op.ncn = 1;
        jb[nj++] = FftJob{dpsi.p, dpsi.p, op};
        KOp od{nullptr, si_table_psi.p, invM, 0.0, 0.0, 0, 0};
        od.table_form = KOP_CNR;
        jb[nj++] = FftJob{sipc_dpsi.p, sipc_dpsi.p, od};
    }
    if ((rc = fft_apply_n(jb, nj))) return rc;
    if (n_omega > 0) {
        const size_t n = static_cast<size_t>(n_omega) * Ms;
        hipLaunchKernelGGL(k_axpy2, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream, omega.p, drift.p, sipc_d.p, n);
        HIP_TRY(hipGetLastError());
    }
    psi_hat_of = nullptr; // psi changes below
    if (electro()) {
        hipLaunchKernelGGL(k_axpy2, dim3(ew_blocks(Ms)), dim3(EW_THREADS), 0, stream, psi.p, dpsi.p, sipc_dpsi.p, Ms);
        HIP_TRY(hipGetLastError());
        if ((rc = remove_mean(psi.p))) return rc;
    } else if ((rc = zero_async(psi.p, Ms))) {
        return rc;
    }
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// Checkpoint and non-finite restart
// ---------------------------------------------------------------------------
int fts_engine::maybe_checkpoint()
{
    if (ckpt_every <= 0) return FTS_OK;
    if (ckpt_step >= 0 && (step_counter % ckpt_every != 0 || step_counter == ckpt_step)) return FTS_OK;
    const size_t Ms = static_cast<size_t>(M);
    const size_t no = static_cast<size_t>(n_omega) * Ms;
    if (!ckpt_psi.p) {
        HIP_TRY(ckpt_omega.alloc(std::max<size_t>(1, no)));
        HIP_TRY(ckpt_psi.alloc(Ms));
    }
    if (no) HIP_TRY(hipMemcpyAsync(ckpt_omega.p, omega.p, sizeof(double2) * no, hipMemcpyDeviceToDevice, stream));
    HIP_TRY(hipMemcpyAsync(ckpt_psi.p, psi.p, sizeof(double2) * Ms, hipMemcpyDeviceToDevice, stream));
    ckpt_step = step_counter;
    return FTS_OK;
}

int fts_engine::restart_from_checkpoint(long long failed_step)
{
    const std::string why = g_error;
    if (ckpt_every <= 0 || ckpt_step < 0) return fail(FTS_NONFINITE, why);
    if (restarts >= max_restarts)
        return fail(FTS_NONFINITE, why + "; maximum non-finite restarts (" + std::to_string(max_restarts) + ") exceeded");
    const size_t Ms = static_cast<size_t>(M);
    const size_t no = static_cast<size_t>(n_omega) * Ms;
    if (no) HIP_TRY(hipMemcpyAsync(omega.p, ckpt_omega.p, sizeof(double2) * no, hipMemcpyDeviceToDevice, stream));
    HIP_TRY(hipMemcpyAsync(psi.p, ckpt_psi.p, sizeof(double2) * Ms, hipMemcpyDeviceToDevice, stream));
    ++restarts;
    ++noise_epoch;
    seed = seed0 + static_cast<unsigned long long>(noise_epoch) * 0x9E3779B97F4A7C15ull;
    last_failed_step = failed_step;
    step_counter = ckpt_step;
    psi_hat_of = nullptr;
    lrk_have_prev = false;
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* fts_last_error(void) { return g_error.c_str(); }

int fts_create(const fts_system* s, fts_engine** out) { return engine_create(s, nullptr, true, out); }

} // extern "C"

// ---- the no-alloc scratch of fts_bead_step_noalloc -------------------------------
// fts_bead_step_noalloc (below fts_bead_step) is the bead step of fts_bead_step on
// device arrays, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// Its intermediates (the bond kernel's axis factors and the work fields of the
// transform passes) live in one file-scope device array, `scratch`, of
// NOALLOC_BYTES bytes.  Host code never uses the array's name as an address: the
// name is a host-side stand-in.  noalloc_setup looks up the array's address on the
// current GPU once per device with hipGetSymbolAddress (each GPU holds its own copy
// until the process ends) and every engine keeps the address of its GPU; the kernels
// receive pointers derived from it.  Every call on a GPU shares that GPU's array and
// finds the previous call's contents there: each call writes every scratch byte it
// reads before reading it, so nothing needs clearing.  Calls that overlap on one GPU
// (two streams, two host threads) race on the array.
namespace {

// grids of at most 2^20 points (96^3 = 884 736 fits): the FP32 form stages q, e^{-W}
// and q' in complex FP32 (3 x 8 B per point), the FP64 form the work field
// (16 B per point); the bond factors take m1 + m2 + m3 <= 768 doubles
constexpr size_t NOALLOC_MAX_POINTS = size_t(1) << 20;
constexpr size_t NOALLOC_TABLE_BYTES = 8192;
constexpr size_t NOALLOC_BYTES = NOALLOC_TABLE_BYTES + 3 * sizeof(float2) * NOALLOC_MAX_POINTS;
constexpr int NOALLOC_MAX_DEVICES = 64;

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

} // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

int noalloc_setup(void** scratch_out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(FTS_DEVICE_ERROR, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(FTS_DEVICE_ERROR, "hipGetSymbolAddress returned no address for the no-alloc scratch");
        g_scratch_ptr[dev] = ptr;
    }
    *scratch_out = g_scratch_ptr[dev];
    return FTS_OK;
}


int engine_create(const fts_system* s, hipStream_t shared, bool tune, fts_engine** out)
{
    if (!s || !out) return fail(FTS_INVALID, "null argument");
    *out = nullptr;
    if (s->m1 <= 0 || s->m2 <= 0 || s->m3 <= 0) return fail(FTS_INVALID, "mesh");
    const long long msize = static_cast<long long>(s->m1) * s->m2 * s->m3;
    if (msize > 2147483647LL) return fail(FTS_INVALID, "mesh size");
    if (s->precision != FTS_PREC_FP64 && s->precision != FTS_PREC_FP32) return fail(FTS_INVALID, "precision");
    if (!line_supported(s->m1, s->m2, s->m3))
        return fail(FTS_INVALID, "mesh not supported by the engine FFT: lengths must be in {8,12,16,24,32,48,64,96,128,192,256}");
    const bool plane_fits = fft_supported(s->m1, s->m2, s->m3, s->precision == FTS_PREC_FP32);
    if ((s->fft_schedule == FTS_SCHED_PLANE || s->fft_schedule == FTS_SCHED_CONV) && !plane_fits)
        return fail(FTS_INVALID, "the plane schedules need m2 == m3 and the m2 x m3 plane in LDS (FP64 up to 96, FP32 up to 128)");
    if (s->fft_schedule == FTS_SCHED_CONV_MFMA && !conv_mfma_supported(s->m1, s->m2, s->m3, s->precision == FTS_PREC_FP32))
        return fail(FTS_INVALID, "the matrix-core conv schedule needs FP32, m1 in {32, 48, 64, 96} and the plane schedules' mesh");
    if (s->fft_schedule < 0 || s->fft_schedule > 5) return fail(FTS_INVALID, "fft_schedule");
    if (s->chain_symmetry != FTS_SYMMETRY_AUTO && s->chain_symmetry != FTS_SYMMETRY_OFF) return fail(FTS_INVALID, "chain_symmetry");
    if (s->a <= 0.0 || s->dL <= 0.0 || s->lB < 0.0) return fail(FTS_INVALID, "a, dL, or lB");
    if (s->dt < 0.0 || s->lambda_psi < 0.0) return fail(FTS_INVALID, "dt or lambdaPsi");
    if (s->n_species <= 0 || s->n_species > FTS_MAX_SPECIES || !s->species) return fail(FTS_INVALID, "polymer species_count");
    if (s->scheme < 0 || s->scheme > 5) return fail(FTS_INVALID, "scheme");
    fts_engine* e = new fts_engine();
    if (shared) {
        e->stream = shared;
        e->own_stream = false;
    }
    auto bail = [&](int rc) {
        delete e;
        return rc;
    };
    e->m1 = s->m1;
    e->m2 = s->m2;
    e->m3 = s->m3;
    e->M = static_cast<int>(msize);
    e->L1 = s->dL * s->m1;
    e->L2 = s->dL * s->m2;
    e->L3 = s->dL * s->m3;
    e->V = e->L1 * e->L2 * e->L3;
    e->dV = s->dL * s->dL * s->dL;
    e->a = s->a;
    e->lB = s->lB;
    e->dt = s->dt;
    e->lambda_psi = s->lambda_psi;
    e->noise = s->noise != 0;
    e->scheme = s->scheme;
    e->precision = s->precision;
    e->seed = s->seed;
    e->seed0 = s->seed;
    e->stress_mask = s->stress_mask;
    e->plane_ok = plane_fits;
    e->sched = s->fft_schedule == FTS_SCHED_AUTO ? (plane_fits ? FTS_SCHED_CONV : FTS_SCHED_SEPARABLE) : s->fft_schedule;
    int rc = e->build_modes(s);
    if (rc) return bail(rc);
    e->total_type_segments.assign(static_cast<size_t>(e->nT), 0);
    for (int p = 0; p < s->n_species; ++p) {
        const fts_species_desc& d = s->species[p];
        if (d.N <= 0 || d.num_chains <= 0 || !d.types || !d.charges) return bail(fail(FTS_INVALID, "polymer species"));
        HostSpecies sp;
        sp.n = d.num_chains;
        sp.N = d.N;
        sp.type.assign(d.types, d.types + d.N);
        sp.charge.assign(d.charges, d.charges + d.N);
        sp.bond.assign(static_cast<size_t>(std::max(0, d.N - 1)), 1.0);
        for (int b = 0; b + 1 < d.N; ++b) {
            if (d.bonds) sp.bond[static_cast<size_t>(b)] = d.bonds[b];
            if (!(sp.bond[static_cast<size_t>(b)] > 0.0)) return bail(fail(FTS_INVALID, "sequence bond length"));
        }
        for (int t : sp.type)
            if (t < 0 || t >= e->nT) return bail(fail(FTS_INVALID, "sequence monomer type"));
        // bond classes (distinct lengths)
        for (double b : sp.bond) {
            int cls = -1;
            for (size_t k = 0; k < e->bond_cls_len.size(); ++k)
                if (e->bond_cls_len[k] == b) cls = static_cast<int>(k);
            if (cls < 0) {
                cls = static_cast<int>(e->bond_cls_len.size());
                e->bond_cls_len.push_back(b);
            }
            sp.bond_cls.push_back(cls);
        }
        // bead classes: distinct (type, charge seen by the field)
        for (int s2 = 0; s2 < d.N; ++s2) {
            const int t = sp.type[static_cast<size_t>(s2)];
            const double z = s->lB > 0.0 ? sp.charge[static_cast<size_t>(s2)] : 0.0;
            int cls = -1;
            for (size_t k = 0; k < sp.cls_type.size(); ++k)
                if (sp.cls_type[k] == t && sp.cls_z[k] == z) cls = static_cast<int>(k);
            if (cls < 0) {
                cls = static_cast<int>(sp.cls_type.size());
                sp.cls_type.push_back(t);
                sp.cls_z.push_back(z);
            }
            sp.cls_of_bead.push_back(cls);
        }
        const int nc = static_cast<int>(sp.cls_type.size());
        if (nc > MAX_CLASSES) return bail(fail(FTS_INVALID, "too many distinct (type, charge) bead classes"));
        sp.cls_start.assign(static_cast<size_t>(nc + 1), 0);
        for (int c = 0; c < nc; ++c) {
            for (int s2 = 0; s2 < d.N; ++s2)
                if (sp.cls_of_bead[static_cast<size_t>(s2)] == c) sp.cls_beads.push_back(s2);
            sp.cls_start[static_cast<size_t>(c + 1)] = static_cast<int>(sp.cls_beads.size());
        }
        // palindrome: bead classes and bond classes read the same from both ends
        bool pal = s->chain_symmetry == FTS_SYMMETRY_AUTO;
        for (int s2 = 0; pal && s2 < d.N; ++s2)
            pal = sp.cls_of_bead[static_cast<size_t>(s2)] == sp.cls_of_bead[static_cast<size_t>(d.N - 1 - s2)];
        for (int s2 = 0; pal && s2 + 1 < d.N; ++s2)
            pal = sp.bond_cls[static_cast<size_t>(s2)] == sp.bond_cls[static_cast<size_t>(d.N - 2 - s2)];
        sp.mirror = pal;
        if (pal) {
            sp.half_start.assign(static_cast<size_t>(nc + 1), 0);
            for (int c = 0; c < nc; ++c) {
                for (int s2 = 0; 2 * s2 <= d.N - 1; ++s2)
                    if (sp.cls_of_bead[static_cast<size_t>(s2)] == c) sp.half_beads.push_back(s2);
                sp.half_start[static_cast<size_t>(c + 1)] = static_cast<int>(sp.half_beads.size());
            }
        }
        e->max_cls = std::max(e->max_cls, nc);
        e->Nmax = std::max(e->Nmax, d.N);
        e->total_segments += sp.n * d.N;
        for (int t : sp.type) e->total_type_segments[static_cast<size_t>(t)] += sp.n;
        e->species.push_back(std::move(sp));
    }
    e->rho0 = static_cast<double>(e->total_segments) / e->V;
    {
        // short_range_self_pressure (param.cu:135-149)
        long double wd = 0.0L;
        for (int al = 0; al < e->nT; ++al) wd += static_cast<long double>(e->total_type_segments[static_cast<size_t>(al)]) * e->K[static_cast<size_t>(al) * e->nT + al];
        const long double md = wd / static_cast<long double>(e->total_segments);
        e->P_sr = static_cast<double>(md / (16.0L * static_cast<long double>(Pi) * std::sqrt(static_cast<long double>(Pi)) * e->a * e->a * e->a));
    }
    e->lambda_omega.assign(static_cast<size_t>(e->n_omega), 0.0);
    if (s->n_lambda_omega == 1 && s->lambda_omega) e->lambda_omega.assign(static_cast<size_t>(e->n_omega), s->lambda_omega[0]);
    else if (s->n_lambda_omega == e->n_omega && s->lambda_omega) e->lambda_omega.assign(s->lambda_omega, s->lambda_omega + e->n_omega);
    else if (s->n_lambda_omega != 0) return bail(fail(FTS_INVALID, "lambda_omega length must be 1 or the number of active interaction modes"));
    for (double x : e->lambda_omega)
        if (!std::isfinite(x) || x < 0.0) return bail(fail(FTS_INVALID, "lambda_omega"));
    if ((rc = e->build_ions(s))) return bail(rc);
    {
        const hipError_t he = hipGetDevice(&e->device);
        if (he != hipSuccess) return bail(fail(FTS_DEVICE_ERROR, std::string("hipGetDevice: ") + hipGetErrorString(he)));
    }
    if ((rc = noalloc_setup(&e->noalloc_scratch))) return bail(rc);
    if ((rc = e->setup_device())) return bail(rc);
    if (e->si_scheme() && (rc = e->build_si_tables())) return bail(rc);
    if (e->learned_scheme() && (rc = e->lrk_setup())) return bail(rc);
    if (tune && s->fft_schedule == FTS_SCHED_AUTO && (rc = e->autotune_schedule())) return bail(rc);
    *out = e;
    return FTS_OK;
}

extern "C" {

void fts_destroy(fts_engine* h) { delete h; }

int fts_get_info(fts_engine* h, fts_info* o)
{
    if (!h || !o) return fail(FTS_INVALID, "null argument");
    std::memset(o, 0, sizeof(*o));
    o->grid_size = h->M;
    o->step = h->step_counter;
    o->n_omega = h->n_omega;
    o->monomer_types = h->nT;
    o->n_species = static_cast<int>(h->species.size());
    o->n_ions = static_cast<int>(h->ions.size());
    o->n_segment_classes = h->max_cls;
    o->precision = h->precision;
    o->scheme = h->scheme;
    o->fft_schedule = h->sched;
    for (int k = 0; k < 2; ++k) o->tune_ms[k] = h->tune_ms[k];
    o->restarts = h->restarts;
    o->noise_epoch = h->noise_epoch;
    o->checkpoint_step = h->ckpt_step;
    o->failed_step = h->last_failed_step;
    o->mirrored_species = 0;
    for (size_t p = 0; p < h->species.size(); ++p)
        if (h->species[p].mirror) o->mirrored_species |= 1 << p;
    o->volume = h->V;
    o->rho0 = h->rho0;
    o->short_range_self_pressure = h->P_sr;
    for (int l = 0; l < h->n_omega; ++l) {
        o->mode_eigenvalue[l] = h->eig[static_cast<size_t>(h->active[static_cast<size_t>(l)])];
        o->omega_coefficient[l] = h->omega_coefficient(l);
        o->gamma_is_imag[l] = h->gamma_imag(l) ? 1.0 : 0.0;
        for (int al = 0; al < h->nT; ++al) o->mode_O[l][al] = h->Ocoef(l, al);
    }
    for (size_t j = 0; j < h->ions.size(); ++j) {
        o->ion_valence[j] = h->ions[j].z;
        o->ion_count[j] = h->ions[j].n;
        o->ion_smearing[j] = h->ions[j].a;
    }
    return FTS_OK;
}

int fts_set_fields(fts_engine* h, const double* om, const double* ps)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    ++h->config_epoch;
    const size_t Ms = static_cast<size_t>(h->M);
    const size_t no = static_cast<size_t>(h->n_omega) * Ms;
    if (om) HIP_TRY(hipMemcpyAsync(h->omega.p, om, sizeof(double2) * no, hipMemcpyHostToDevice, h->stream));
    else HIP_TRY(hipMemsetAsync(h->omega.p, 0, sizeof(double2) * no, h->stream));
    if (ps) HIP_TRY(hipMemcpyAsync(h->psi.p, ps, sizeof(double2) * Ms, hipMemcpyHostToDevice, h->stream));
    else HIP_TRY(hipMemsetAsync(h->psi.p, 0, sizeof(double2) * Ms, h->stream));
    // remove_psi_zero_mode on load (complexLangevin.cu:183, 313-317)
    int rc = h->electro() ? h->remove_mean(h->psi.p) : FTS_OK;
    if (rc) return rc;
    if (!h->electro()) HIP_TRY(hipMemsetAsync(h->psi.p, 0, sizeof(double2) * Ms, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->ckpt_step = -1; // new fields: the next step takes a new checkpoint
    h->lrk_have_prev = false;
    h->psi_hat_of = nullptr;
    return FTS_OK;
}

// Binary state file (include/fts.h): "FTSSTAT1", int32 m1 m2 m3 n_omega, int64 step,
// uint64 seed0, int32 noise_epoch restarts, then omega [n_omega][M] and psi [M] as
// interleaved doubles, host byte order.  Written to path.new and renamed into place.
int fts_save_state(fts_engine* h, const char* path)
{
    if (!h || !path) return fail(FTS_INVALID, "null argument");
    const size_t Ms = static_cast<size_t>(h->M), no = static_cast<size_t>(h->n_omega) * Ms;
    std::vector<double2> w(no), p(Ms);
    if (no) HIP_TRY(hipMemcpyAsync(w.data(), h->omega.p, sizeof(double2) * no, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(p.data(), h->psi.p, sizeof(double2) * Ms, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const std::string tmp = std::string(path) + ".new";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return fail(FTS_INVALID, std::string("cannot open ") + tmp);
    const int32_t hdr[4] = {h->m1, h->m2, h->m3, h->n_omega};
    const int64_t step = h->step_counter;
    const uint64_t seed0 = h->seed0;
    const int32_t epoch[2] = {h->noise_epoch, h->restarts};
    bool ok = std::fwrite("FTSSTAT1", 1, 8, fp) == 8 && std::fwrite(hdr, sizeof(hdr), 1, fp) == 1 && std::fwrite(&step, sizeof(step), 1, fp) == 1 &&
              std::fwrite(&seed0, sizeof(seed0), 1, fp) == 1 && std::fwrite(epoch, sizeof(epoch), 1, fp) == 1 &&
              (no == 0 || std::fwrite(w.data(), sizeof(double2), no, fp) == no) && std::fwrite(p.data(), sizeof(double2), Ms, fp) == Ms;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) return fail(FTS_INVALID, std::string("cannot write ") + path);
    return FTS_OK;
}

int fts_load_state(fts_engine* h, const char* path)
{
    if (!h || !path) return fail(FTS_INVALID, "null argument");
    ++h->config_epoch;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(FTS_INVALID, std::string("cannot open ") + path);
    char magic[8];
    int32_t hdr[4];
    int64_t step = 0;
    uint64_t seed0 = 0;
    int32_t epoch[2];
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, "FTSSTAT1", 8) == 0 && std::fread(hdr, sizeof(hdr), 1, fp) == 1 &&
              std::fread(&step, sizeof(step), 1, fp) == 1 && std::fread(&seed0, sizeof(seed0), 1, fp) == 1 && std::fread(epoch, sizeof(epoch), 1, fp) == 1;
    if (!ok) {
        std::fclose(fp);
        return fail(FTS_INVALID, std::string(path) + ": not a state file");
    }
    if (hdr[0] != h->m1 || hdr[1] != h->m2 || hdr[2] != h->m3 || hdr[3] != h->n_omega || step < 0 || epoch[0] < 0 || epoch[1] < 0) {
        std::fclose(fp);
        return fail(FTS_INVALID, std::string(path) + ": state of another mesh or mode count");
    }
    const size_t Ms = static_cast<size_t>(h->M), no = static_cast<size_t>(h->n_omega) * Ms;
    std::vector<double2> w(no), p(Ms);
    ok = (no == 0 || std::fread(w.data(), sizeof(double2), no, fp) == no) && std::fread(p.data(), sizeof(double2), Ms, fp) == Ms;
    std::fclose(fp);
    if (!ok) return fail(FTS_INVALID, std::string(path) + ": truncated state file");
    if (no) HIP_TRY(hipMemcpyAsync(h->omega.p, w.data(), sizeof(double2) * no, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->psi.p, p.data(), sizeof(double2) * Ms, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->step_counter = step;
    h->seed0 = seed0;
    h->noise_epoch = epoch[0];
    h->restarts = epoch[1];
    h->seed = seed0 + static_cast<unsigned long long>(epoch[0]) * 0x9E3779B97F4A7C15ull;
    h->ckpt_step = -1;
    h->lrk_have_prev = false;
    h->psi_hat_of = nullptr;
    return FTS_OK;
}

int fts_get_field(fts_engine* h, int id, double* out, size_t n_doubles)
{
    if (!h || !out) return fail(FTS_INVALID, "null argument");
    const size_t Ms = static_cast<size_t>(h->M);
    const double2* src = nullptr;
    size_t count = 0;
    switch (id) {
    case FTS_FIELD_OMEGA: src = h->omega.p; count = static_cast<size_t>(h->n_omega) * Ms; break;
    case FTS_FIELD_PSI: src = h->psi.p; count = Ms; break;
    case FTS_FIELD_RHO_TYPE: src = h->rho_type.p; count = static_cast<size_t>(h->nT) * Ms; break;
    case FTS_FIELD_RHO_P: src = h->rho_p.p; count = Ms; break;
    case FTS_FIELD_RHO_C: src = h->rho_c.p; count = Ms; break;
    case FTS_FIELD_RHO_C_POLY: src = h->rho_c_poly.p; count = Ms; break;
    case FTS_FIELD_RHO_C_ION: src = h->rho_c_ion.p; count = Ms; break;
    case FTS_FIELD_ION_RHO: src = h->ion_rho.p; count = h->ions.size() * Ms; break;
    case FTS_FIELD_QF:
    case FTS_FIELD_QB: {
        if (int rc = h->read_log_gain()) return rc;
        const HostSpecies& sp = h->species[static_cast<size_t>(h->last_species)];
        count = static_cast<size_t>(sp.N) * Ms;
        if (n_doubles != 2 * count) return fail(FTS_INVALID, "field size mismatch");
        const void* q = id == FTS_FIELD_QF || sp.mirror ? h->qF : h->qB;
        // undo the FP32 potential shifts: qF_s = q~F_s e^{-sum_{t<=s} shift}, qB_s = q~B_s e^{-sum_{t>=s} shift}
        std::vector<double> lc(2 * static_cast<size_t>(h->Nmax), 0.0);
        HIP_TRY(hipMemcpy(lc.data(), h->logc.p, sizeof(double) * lc.size(), hipMemcpyDeviceToHost));
        for (int s = 0; s < sp.N; ++s) {
            double shift = 0.0, k2sum = 0.0;
            for (int t = 0; t < sp.N; ++t)
                if ((id == FTS_FIELD_QF && t <= s) || (id == FTS_FIELD_QB && t >= s)) {
                    shift += sp.cls_shift[static_cast<size_t>(sp.cls_of_bead[static_cast<size_t>(t)])];
                    k2sum += lc[static_cast<size_t>(id == FTS_FIELD_QF ? t : sp.mirror ? sp.N - 1 - t : h->Nmax + t)];
                }
            const double scale = std::exp(-shift) * std::exp2(-k2sum);
            const int src_s = id == FTS_FIELD_QB && sp.mirror ? sp.N - 1 - s : s;
            double2* dst = h->t1.p;
            if (h->precision == FTS_PREC_FP32)
                hipLaunchKernelGGL((k_cast<float, double>), dim3(ew_blocks(Ms)), dim3(EW_THREADS), 0, h->stream,
                                   static_cast<const float2*>(q) + static_cast<size_t>(src_s) * Ms, dst, scale, Ms);
            else
                hipLaunchKernelGGL((k_cast<double, double>), dim3(ew_blocks(Ms)), dim3(EW_THREADS), 0, h->stream,
                                   static_cast<const double2*>(q) + static_cast<size_t>(src_s) * Ms, dst, scale, Ms);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(out + 2 * static_cast<size_t>(s) * Ms, dst, sizeof(double2) * Ms, hipMemcpyDeviceToHost, h->stream));
        }
        HIP_TRY(hipStreamSynchronize(h->stream));
        return FTS_OK;
    }
    default: return fail(FTS_INVALID, "field id");
    }
    if (n_doubles != 2 * count) return fail(FTS_INVALID, "field size mismatch");
    HIP_TRY(hipMemcpyAsync(out, src, sizeof(double) * n_doubles, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FTS_OK;
}

int fts_compute_densities(fts_engine* h, double* Q_out)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    int rc = h->compute_densities(h->omega.p, h->psi.p);
    if (rc) return rc;
    double2 Q[FTS_MAX_SPECIES];
    if ((rc = h->read_scal(SC_Q, static_cast<int>(h->species.size()), Q))) return rc;
    if ((rc = h->read_log_gain())) return rc;
    bool ok = true;
    for (size_t p = 0; p < h->species.size(); ++p) {
        const double sc = std::exp(-h->q_shift_total[p] - h->q_log_gain[p]);
        if (Q_out) {
            Q_out[2 * p] = Q[p].x * sc;
            Q_out[2 * p + 1] = Q[p].y * sc;
        }
        ok = ok && finite2(Q[p]);
    }
    return ok ? FTS_OK : fail(FTS_NONFINITE, "non-finite Q");
}

int fts_measure(fts_engine* h, fts_observables* out)
{
    if (!h || !out) return fail(FTS_INVALID, "null argument");
    int rc = h->compute_densities(h->omega.p, h->psi.p);
    if (rc) return rc;
    if ((rc = h->measure(h->omega.p, h->psi.p, out))) return rc;
    for (int p = 0; p < out->n_species; ++p)
        if (!std::isfinite(out->mu_ex_poly[p][0]) || !std::isfinite(out->mu_ex_poly[p][1])) return fail(FTS_NONFINITE, "non-finite Q");
    return FTS_OK;
}

int fts_step(fts_engine* h, int n, int record_every, fts_observables* records, int capacity, int* n_records)
{
    if (!h || n < 0) return fail(FTS_INVALID, "null engine or negative step count");
    int nrec = 0;
    if (n_records) *n_records = 0;
    int rc = FTS_OK;
    const long long target = h->step_counter + n;
    while (h->step_counter < target) {
        if ((rc = h->maybe_checkpoint())) break;
        const long long step = h->step_counter + 1; // 1-based step label (complexLangevin.cu:206)
        const bool record = record_every > 0 && step % record_every == 0 && records && nrec < capacity;
        fts_observables* rec = record ? &records[nrec] : nullptr;
        rc = h->scheme == FTS_SCHEME_PC_REFERENCE ? h->step_pc(record, rec)
             : scheme_two_stage(h->scheme)        ? h->step_sipc(record, rec)
                                                  : h->step_em(record, rec, h->si_scheme());
        if (record) {
            rec->step = step;
            ++nrec;
        }
        if (rc == FTS_OK && h->drift_stats) {
            double2 v[FTS_MAX_TYPES + 1];
            if ((rc = h->read_scal(SC_DRIFT, h->n_omega + 1, v))) break;
            h->drift_rows.push_back(static_cast<double>(step));
            for (int k = 0; k <= h->n_omega; ++k) h->drift_rows.push_back(v[k].x);
        }
        if (rc == FTS_NONFINITE && h->ckpt_every > 0) {
            // restart from the checkpoint with fresh noise; drop the records taken after it
            if ((rc = h->restart_from_checkpoint(step))) break;
            while (nrec > 0 && records[nrec - 1].step > h->ckpt_step) --nrec;
            const size_t w = static_cast<size_t>(h->n_omega) + 2;
            while (!h->drift_rows.empty() && h->drift_rows[h->drift_rows.size() - w] > static_cast<double>(h->ckpt_step))
                h->drift_rows.resize(h->drift_rows.size() - w);
        }
        if (n_records) *n_records = nrec;
        if (rc != FTS_OK) break;
        if (h->timing && h->pending.size() > 256 && (rc = h->resolve_timing())) break;
    }
    return rc;
}

int fts_set_drift_stats(fts_engine* h, int enabled)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    ++h->config_epoch;
    h->drift_stats = enabled != 0;
    h->drift_rows.clear();
    return FTS_OK;
}

int fts_get_drift_rows(fts_engine* h, double* out, int capacity, int* n_rows)
{
    if (!h || !n_rows || capacity < 0 || (capacity > 0 && !out)) return fail(FTS_INVALID, "bad argument");
    const size_t w = static_cast<size_t>(h->n_omega) + 2;
    const size_t n = std::min(h->drift_rows.size() / w, static_cast<size_t>(capacity));
    if (n) std::memcpy(out, h->drift_rows.data(), sizeof(double) * n * w);
    h->drift_rows.erase(h->drift_rows.begin(), h->drift_rows.begin() + static_cast<std::ptrdiff_t>(n * w));
    *n_rows = static_cast<int>(n);
    return FTS_OK;
}

int fts_inject_nonfinite(fts_engine* h, long long step)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    ++h->config_epoch;
    h->inject_step = step;
    return FTS_OK;
}

int fts_set_restart(fts_engine* h, int checkpoint_every, int max_restarts)
{
    if (!h || checkpoint_every < 0 || max_restarts < 0) return fail(FTS_INVALID, "bad restart settings");
    h->ckpt_every = checkpoint_every;
    h->max_restarts = max_restarts;
    h->restarts = 0;
    h->ckpt_step = -1;
    return FTS_OK;
}

int fts_set_step(fts_engine* h, long long step)
{
    if (!h || step < 0) return fail(FTS_INVALID, "bad step");
    ++h->config_epoch;
    h->step_counter = step;
    h->ckpt_step = -1;
    return FTS_OK;
}

int fts_fft3d(fts_engine* h, const double* in, double* out, int sign)
{
    if (!h || !in || !out || (sign != 1 && sign != -1)) return fail(FTS_INVALID, "bad argument");
    const size_t Ms = static_cast<size_t>(h->M);
    HIP_TRY(hipMemcpyAsync(h->t1.p, in, sizeof(double2) * Ms, hipMemcpyHostToDevice, h->stream));
    if (h->plane_ok) {
        PlaneJob<double> pj{h->t1.p, h->t2.p, nullptr, nullptr};
        HIP_TRY(launch_plane<double>(h->g64, sign < 0 ? PLANE_FWD : PLANE_INV, &pj, 1, h->stream));
        PencilJob<double> xj{h->t2.p, KOp{}};
        HIP_TRY(launch_pencil<double>(h->g64, sign < 0 ? PENCIL_FWD : PENCIL_INV, &xj, 1, h->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(h->t2.p, h->t1.p, sizeof(double2) * Ms, hipMemcpyDeviceToDevice, h->stream));
        LineJob<double> j{h->t2.p};
        const int mode = sign < 0 ? LINE_FWD : LINE_INV;
        HIP_TRY(launch_line<double>(h->g64, 2, mode, h->tw64_z.p, &j, 1, h->stream));
        HIP_TRY(launch_line<double>(h->g64, 1, mode, h->tw64_yz.p, &j, 1, h->stream));
        HIP_TRY(launch_line<double>(h->g64, 0, mode, h->tw64_x.p, &j, 1, h->stream));
    }
    HIP_TRY(hipMemcpyAsync(out, h->t2.p, sizeof(double2) * Ms, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FTS_OK;
}

int fts_bead_step(fts_engine* h, const double* q, const double* emw, double bond, double* q_out)
{
    if (!h || !q || !emw || !q_out || !(bond > 0.0)) return fail(FTS_INVALID, "bad argument");
    const size_t Ms = static_cast<size_t>(h->M);
    const int nk = h->m1 + h->m2 + h->m3;
    DevBuf<double> dsep;
    HIP_TRY(dsep.alloc(static_cast<size_t>(nk)));
    hipLaunchKernelGGL(k_bond_table, dim3((nk + 255) / 256), dim3(256), 0, h->stream, h->kvec.p, bond, dsep.p, nk);
    HIP_TRY(hipMemcpyAsync(h->t1.p, q, sizeof(double2) * Ms, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->t2.p, emw, sizeof(double2) * Ms, hipMemcpyHostToDevice, h->stream));
    const KOp op = h->kop_sep(dsep.p, 1.0 / static_cast<double>(h->M));
    const int nb = ew_blocks(Ms);
    int rc = FTS_OK;
    if (h->precision == FTS_PREC_FP32) {
        float2* a = static_cast<float2*>(h->tmpF);
        float2* e = static_cast<float2*>(h->E);
        float2* o = static_cast<float2*>(h->qF);
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(EW_THREADS), 0, h->stream, h->t1.p, a, 1.0, Ms);
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(EW_THREADS), 0, h->stream, h->t2.p, e, 1.0, Ms);
        rc = h->bead_hook<float>(a, e, o, op, h->stream);
        hipLaunchKernelGGL((k_cast<float, double>), dim3(nb), dim3(EW_THREADS), 0, h->stream, o, h->t3.p, 1.0, Ms);
    } else {
        rc = h->bead_hook<double>(h->t1.p, h->t2.p, h->t3.p, op, h->stream);
    }
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(q_out, h->t3.p, sizeof(double2) * Ms, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    dsep.release();
    return FTS_OK;
}

int fts_bead_step_noalloc(fts_engine* h, const double* q, const double* exp_minus_w, double bond, double* q_out,
                                     void* stream)
{
    if (!h || !q || !exp_minus_w || !q_out || !(bond > 0.0)) return fail(FTS_INVALID, "bead_step_noalloc: bad argument");
    const size_t Ms = static_cast<size_t>(h->M);
    if (Ms > NOALLOC_MAX_POINTS)
        return fail(FTS_INVALID, "bead_step_noalloc: a grid of " + std::to_string(Ms) + " points exceeds the no-alloc scratch (at most " +
                                     std::to_string(NOALLOC_MAX_POINTS) + " points)");
    const int nk = h->m1 + h->m2 + h->m3;
    if (static_cast<size_t>(nk) * sizeof(double) > NOALLOC_TABLE_BYTES) return fail(FTS_INVALID, "bead_step_noalloc: axis lengths");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device) return fail(FTS_INVALID, "bead_step_noalloc: the engine's GPU must be the current device");
    if (!h->noalloc_scratch) return fail(FTS_DEVICE_ERROR, "bead_step_noalloc: scratch address not set up");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
    unsigned char* base = static_cast<unsigned char*>(h->noalloc_scratch);
    double* table = reinterpret_cast<double*>(base);
    unsigned char* fields = base + NOALLOC_TABLE_BYTES;
    hipLaunchKernelGGL(k_bond_table, dim3((nk + 255) / 256), dim3(256), 0, s, h->kvec.p, bond, table, nk);
    const KOp op = h->kop_sep(table, 1.0 / static_cast<double>(h->M));
    const int nb = ew_blocks(Ms);
    const double2* qd = reinterpret_cast<const double2*>(q);
    const double2* ed = reinterpret_cast<const double2*>(exp_minus_w);
    double2* od = reinterpret_cast<double2*>(q_out);
    int rc = FTS_OK;
    if (h->precision == FTS_PREC_FP32) {
        float2* a = reinterpret_cast<float2*>(fields);
        float2* e = a + Ms;
        float2* o = e + Ms;
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(EW_THREADS), 0, s, qd, a, 1.0, Ms);
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(EW_THREADS), 0, s, ed, e, 1.0, Ms);
        rc = h->bead_hook<float>(a, e, o, op, s);
        if (rc) return rc;
        hipLaunchKernelGGL((k_cast<float, double>), dim3(nb), dim3(EW_THREADS), 0, s, o, od, 1.0, Ms);
    } else {
        // the passes overwrite their input with intermediates: work on a copy of q
        double2* a = reinterpret_cast<double2*>(fields);
        hipLaunchKernelGGL((k_cast<double, double>), dim3(nb), dim3(EW_THREADS), 0, s, qd, a, 1.0, Ms);
        rc = h->bead_hook<double>(a, ed, od, op, s);
        if (rc) return rc;
    }
    HIP_TRY(hipGetLastError());
    return FTS_OK;
}

int fts_set_timing(fts_engine* h, int enabled)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    int rc = h->resolve_timing();
    if (rc) return rc;
    h->timing = enabled != 0;
    h->tstat = fts_timing{};
    return FTS_OK;
}

int fts_get_timing(fts_engine* h, fts_timing* out)
{
    if (!h || !out) return fail(FTS_INVALID, "null argument");
    int rc = h->resolve_timing();
    if (rc) return rc;
    *out = h->tstat;
    return FTS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// C ABI: the learned linear-response kernel of an engine
// ---------------------------------------------------------------------------
extern "C" {

int fts_lrk_set_network(fts_engine* h, const float* params, double threshold)
{
    if (!h || !params) return fail(FTS_INVALID, "lrk_set_network: null argument");
    ++h->config_epoch;
    if (std::isnan(threshold)) return fail(FTS_INVALID, "lrk_set_network: threshold is NaN");
    for (int j = 0; j < FTS_LRK_NPARAMS; ++j)
        if (!std::isfinite(params[j])) return fail(FTS_INVALID, "lrk_set_network: non-finite parameter");
    if (!h->lrk_params.p) HIP_TRY(h->lrk_params.alloc(FTS_LRK_NPARAMS));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->lrk_params.p, params, sizeof(float) * FTS_LRK_NPARAMS, hipMemcpyHostToDevice));
    h->lrk_threshold = threshold;
    return FTS_OK;
}

int fts_lrk_set_recording(fts_engine* h, int capacity)
{
    if (!h || capacity < 0) return fail(FTS_INVALID, "lrk_set_recording: bad argument");
    ++h->config_epoch;
    if (capacity > 0 && !h->si_scheme())
        return fail(FTS_INVALID, "training records need a semi-implicit scheme");
    int rc;
    if (capacity > 0 && (rc = h->lrk_setup())) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->lrk_ring.release();
    if (capacity > 0) HIP_TRY(h->lrk_ring.alloc(static_cast<size_t>(capacity) * FTS_LRK_RECORD));
    h->lrk_capacity = capacity;
    if (h->lrk_state.p) {
        LrkDevState st;
        HIP_TRY(hipMemcpy(&st, h->lrk_state.p, sizeof(st), hipMemcpyDeviceToHost));
        st.records = 0;
        h->lrk_read_idx = st.write_idx;
        HIP_TRY(hipMemcpy(h->lrk_state.p, &st, sizeof(st), hipMemcpyHostToDevice));
    }
    return FTS_OK;
}

int fts_lrk_get_records(fts_engine* h, double* out, int max_records, int* n_out)
{
    if (!h || !n_out || max_records < 0 || (max_records > 0 && !out)) return fail(FTS_INVALID, "lrk_get_records: bad argument");
    *n_out = 0;
    if (!h->lrk_state.p || h->lrk_capacity <= 0) return FTS_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    LrkDevState st;
    HIP_TRY(hipMemcpy(&st, h->lrk_state.p, sizeof(st), hipMemcpyDeviceToHost));
    const long long first = std::max(h->lrk_read_idx, st.write_idx - h->lrk_capacity);
    const long long n = std::min<long long>(max_records, st.write_idx - first);
    for (long long k = 0; k < n; ++k) {
        const size_t slot = static_cast<size_t>((first + k) % h->lrk_capacity);
        HIP_TRY(hipMemcpy(out + static_cast<size_t>(k) * FTS_LRK_RECORD, h->lrk_ring.p + slot * FTS_LRK_RECORD, sizeof(double) * FTS_LRK_RECORD,
                          hipMemcpyDeviceToHost));
    }
    h->lrk_read_idx = st.write_idx;
    *n_out = static_cast<int>(n);
    return FTS_OK;
}

int fts_lrk_get_stats(fts_engine* h, fts_lrk_stats* out)
{
    if (!h || !out) return fail(FTS_INVALID, "lrk_get_stats: null argument");
    std::memset(out, 0, sizeof(*out));
    if (!h->lrk_state.p) return FTS_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    LrkDevState st;
    HIP_TRY(hipMemcpy(&st, h->lrk_state.p, sizeof(st), hipMemcpyDeviceToHost));
    out->judged = st.judged;
    out->learned_steps = st.learned;
    out->fallback_steps = st.fallback;
    out->records = st.records;
    out->last_excess = st.last_excess;
    for (int c = 0; c < 2; ++c) {
        out->sum_rel_residual_learned[c] = st.sum_rl[c];
        out->sum_rel_residual_debye[c] = st.sum_rd[c];
    }
    return FTS_OK;
}

} // extern "C"
