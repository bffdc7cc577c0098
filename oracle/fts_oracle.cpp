// fts_oracle.cpp -- FP64 CPU oracle of the complex-Langevin hot path.
//
// TEST INFRASTRUCTURE.  This is the checker, not the product: only tests/, the
// smoke check and bench.py's cpu_baseline leg load it.  It restates the
// reference's algorithm, file by file, in plain C++ over host vectors:
//   - grid, kernels, smoothing        common.h:290-328, 1079-1102; spectral.cu:3-35
//   - interaction modes               interactionMode.cu:9-298; param.cu:115-149
//   - ions (build, densities)         ion.cu:3-72, 117-134; common.h:547-558, 1104-1130
//   - propagators and densities       polymer.cu:341-478; common.h:560-580
//   - observables, H, mu, p_iso       observable.cu:53-317
//   - stress tensor                   stress.cu:106-321
//   - drifts, noise, PC step          complexLangevin.cu:255-368, 669-698
// and adds the schemes the reference lacks (plain Euler-Maruyama, the semi-implicit
// update and the semi-implicit predictor-corrector, each with the Debye or the learned
// kernel), with the same formulas the engine uses.  Noise is Philox4x32-10 (philox_host.h) instead of cuRAND, the FFT is
// oracle_fft.h instead of cuFFT, and reductions are long-double block sums.
//
// It exports the engine's C ABI (include/fts.h) under the prefix fto_ so that a
// test can drive the oracle and the HIP engine with identical calls.
//
// Parity pins (tests/): brute-force DFT and numpy for the FFT; the Philox KAT
// fixture; closed forms at w = psi = 0 and at uniform fields; the propagator
// invariant <qF_s qB_s e^{W_s}> = Q; exact density means; p_iso = tr(sigma)/3;
// M = 1 / M = 2 mode closed forms; the Heun step's second-order convergence;
// the discrete-chain linear response for the semi-implicit kernel; the semi-implicit
// steps' damping factors and noise spectra; the predictor-corrector's Crank-Nicolson
// limit, its record at the predictor state and its second order; the electrostatics at
// psi != 0 and the stress as a strain derivative (tests/test_oracle_pins.py).
#include "../include/fts.h"
#include "lrk_mlp.h"
#include "oracle_fft.h"
#include "philox_host.h"

#include <algorithm>
#include <cstdint>
#include <deque>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

using oracle::Cx;

namespace {

constexpr double Pi = 3.141592653589793238462643383279502884;

thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

// complex helpers: the same formulas as common.h:78-139, 191-201
inline Cx cx(double x, double y) { return Cx{x, y}; }
inline Cx cadd(Cx a, Cx b) { return Cx{a.x + b.x, a.y + b.y}; }
inline Cx csub(Cx a, Cx b) { return Cx{a.x - b.x, a.y - b.y}; }
inline Cx cmul(Cx a, Cx b) { return Cx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
inline Cx cscale(Cx z, double s) { return Cx{s * z.x, s * z.y}; }
inline Cx cinv(Cx z)
{
    const double ax = std::fabs(z.x), ay = std::fabs(z.y);
    if (ax >= ay) {
        const double r = z.y / z.x;
        const double d = z.x * (1.0 + r * r);
        return Cx{1.0 / d, -r / d};
    }
    const double r = z.x / z.y;
    const double d = z.y * (1.0 + r * r);
    return Cx{r / d, -1.0 / d};
}
inline Cx cexp(Cx z)
{
    const double r = std::exp(z.x);
    return Cx{r * std::cos(z.y), r * std::sin(z.y)};
}
// -log z; |z| by hypot, which stays finite for every finite z (the squares overflow from
// |z| ~ 1e154, which a hot spot can give an ion partition function)
inline Cx cneglog(Cx z)
{
    return Cx{-std::log(std::hypot(z.x, z.y)), -std::atan2(z.y, z.x)};
}
inline bool finite(Cx z) { return std::isfinite(z.x) && std::isfinite(z.y); }

inline int spectral_mode(int i, int m) { return i > m / 2 ? i - m : i; } // common.h:219-222
inline double wave_number(int i, int m, double L) { return 2.0 * Pi * spectral_mode(i, m) / L; }

using Field = std::vector<Cx>;

struct Species {
    long long n = 0;
    int N = 0;
    std::vector<int> type;
    std::vector<double> charge;
    std::vector<double> bond;
    size_t bond_offset = 0;
    double charge_square_sum() const
    {
        double t = 0.0;
        for (double z : charge) t += z * z;
        return t;
    }
    double net_charge() const
    {
        double t = 0.0;
        for (double z : charge) t += z;
        return t;
    }
};

struct Ion {
    double z = 0.0;
    long long n = 0;
    double a = 0.0;
};

} // namespace

struct fts_engine {
    // ---- parameters -------------------------------------------------------
    int m1 = 0, m2 = 0, m3 = 0, M = 0;
    double L1 = 0, L2 = 0, L3 = 0, V = 0, dV = 0, a = 0, lB = 0, dt = 0, lambda_psi = 0;
    bool noise = true;
    int scheme = 0;
    unsigned long long seed = 0;
    long long step_counter = 0;
    int stress_mask = 0;

    // checkpoint and non-finite restart (complexLangevin.cu:518-667); the restart
    // moves the noise to the key seed0 + epoch * 0x9E3779B97F4A7C15
    int ckpt_every = 0, max_restarts = 0, restarts = 0, noise_epoch = 0;
    long long last_failed_step = -1;
    long long ckpt_step = -1, inject_step = -1;
    unsigned long long seed0 = 0;
    std::vector<Cx> ckpt_omega, ckpt_psi;

    // modes
    int nT = 0;                      // monomer types
    double xi = 0;
    std::vector<double> K;           // [nT*nT]
    std::vector<double> eig;         // sorted descending [nT]
    std::vector<double> O;           // [nT][nT] rows = modes
    std::vector<int> active;
    std::vector<double> lambda_omega;
    int n_omega = 0;
    double rho0 = 0;
    long long total_segments = 0;
    std::vector<long long> total_type_segments;

    std::vector<Species> species;
    std::vector<Ion> ions;
    int Nmax = 0;
    size_t total_bonds = 0;

    // kernels
    std::vector<double> k2, Gamma, PHI, ion_kernels, kab, minus_kab;
    double mode_term[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> A_omega, A_psi; // semi-implicit linear-response kernels

    // learned linear-response kernel (FTS_SCHEME_SI_LEARNED, include/fts.h FTS_LRK_*)
    std::vector<double> lrk_P = std::vector<double>(FTS_LRK_NPARAMS, 0.0);
    double lrk_threshold = 1.0e-3;
    std::vector<int> lrk_binof;   // [M] |k| bin of every wave vector (k = 0 in bin 0)
    bool lrk_have_prev = false;
    Field lrk_W[2], lrk_U[2];     // k-space w and u(w) of the previous step, field c
    double lrk_m_prev[FTS_LRK_MOMENTS] = {};
    double lrk_r_prev[2][FTS_LRK_BINS] = {};
    fts_lrk_stats lrk_st{};
    int lrk_capacity = 0;
    std::deque<std::vector<double>> lrk_ring;
    bool learned_scheme() const { return scheme == FTS_SCHEME_SI_LEARNED || scheme == FTS_SCHEME_SIPC_LEARNED; }
    bool si_scheme() const { return scheme == FTS_SCHEME_SEMI_IMPLICIT || scheme == FTS_SCHEME_SIPC || learned_scheme(); }
    bool lrk_active() const { return learned_scheme() || lrk_capacity > 0; }
    void lrk_build_bins();
    void lrk_moments(double* m);
    void lrk_step(std::vector<double> (&mult)[2]);

    oracle::Fft3d* fft = nullptr;

    // drift statistics (complexLangevin.cu:318-324, 466-497): values of the step being
    // taken (modes, then psi) and the rows [step, values] of the steps taken
    bool drift_stats = false;
    std::vector<double> drift_cur, drift_rows;
    // corrector_drift_max: |scale| sqrt(max_i |a_i + b_i|^2), a non-finite |.|^2 counting as +inf
    double drift_max(const Cx* a, const Cx* b, double scale) const
    {
        double m = 0.0;
        for (int i = 0; i < M; ++i) {
            const Cx z = b ? Cx{a[i].x + b[i].x, a[i].y + b[i].y} : a[i];
            double v = z.x * z.x + z.y * z.y;
            if (!std::isfinite(v)) v = INFINITY;
            m = std::max(m, v);
        }
        return std::fabs(scale) * std::sqrt(std::max(0.0, m));
    }
    void drift_statistics(bool pc)
    {
        if (!drift_stats) return;
        drift_cur.assign(static_cast<size_t>(n_omega) + 1, 0.0);
        const double f = pc ? 0.5 : 1.0;
        for (int l = 0; l < n_omega; ++l)
            drift_cur[static_cast<size_t>(l)] = drift_max(fld(drift, l), pc ? fld(drift_bar, l) : nullptr, f * lambda_omega[static_cast<size_t>(l)]);
        drift_cur[static_cast<size_t>(n_omega)] = drift_max(dpsi.data(), pc ? dpsi_bar.data() : nullptr, f * lambda_psi);
    }

    // ---- state (deviceState.h:5-53) ----------------------------------------
    Field omega, omega_bar, omega_s, drift, drift_bar;
    std::vector<double> eta_omega, etaPsi;
    Field psi, psi_bar, psi_s, dpsi, dpsi_bar;
    Field W_type_s, rho_type, rho_c, rho_c_poly, rho_c_ion, rho_c_old, rho_p;
    Field W, exp_mW, exp_W, qF, qB;
    Field ion_psi, ion_boltzmann, ion_charge, ion_rho;
    Field obs_a, obs_b, sa, sb, sc, dfield, dpsi_s;
    std::vector<Cx> Q_ion;
    // the step's check: polymer Qs (Qp) and the ion Qs of the same densities finite (an
    // ion Q that overflows leaves finite fields, zero ion densities and H = -inf)
    bool partition_finite(const std::vector<Cx>& Qp) const
    {
        for (const Cx& q : Qp)
            if (!finite(q)) return false;
        for (const Cx& q : Q_ion)
            if (!finite(q)) return false;
        return true;
    }
    int last_species = 0;

    ~fts_engine() { delete fft; }

    // ---- primitives --------------------------------------------------------
    Cx* fld(Field& f, size_t idx) { return f.data() + idx * static_cast<size_t>(M); }

    Cx sum(const Cx* d) const
    {
        // fixed-block long-double sum: deterministic for any thread count
        const int block = 4096;
        const int nb = (M + block - 1) / block;
        std::vector<long double> px(static_cast<size_t>(nb)), py(static_cast<size_t>(nb));
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nb; ++b) {
            long double sx = 0.0L, sy = 0.0L;
            const int end = std::min(M, (b + 1) * block);
            for (int i = b * block; i < end; ++i) {
                sx += d[i].x;
                sy += d[i].y;
            }
            px[static_cast<size_t>(b)] = sx;
            py[static_cast<size_t>(b)] = sy;
        }
        long double sx = 0.0L, sy = 0.0L;
        for (int b = 0; b < nb; ++b) {
            sx += px[static_cast<size_t>(b)];
            sy += py[static_cast<size_t>(b)];
        }
        return Cx{static_cast<double>(sx), static_cast<double>(sy)};
    }
    double sum_real(const double* d) const
    {
        long double s = 0.0L;
        for (int i = 0; i < M; ++i) s += d[i];
        return static_cast<double>(s);
    }
    Cx mean(const Cx* d) const { return cscale(sum(d), 1.0 / static_cast<double>(M)); }

    // fft_apply_real_kernel (common.h:1079-1085): dst = IFFT[kernel * FFT[src] / M]
    void apply_kernel(const Cx* src, Cx* dst, const double* kernel)
    {
        if (dst != src) std::memcpy(dst, src, sizeof(Cx) * static_cast<size_t>(M));
        fft->run(dst, -1);
        const double inv = 1.0 / static_cast<double>(M);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {
            dst[i] = cscale(dst[i], inv);
            dst[i] = cscale(dst[i], kernel[i]);
        }
        fft->run(dst, +1);
    }
    void mul(Cx* d, const Cx* s)
    {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) d[i] = cmul(d[i], s[i]);
    }
    void axpy(Cx* d, const Cx* s, double scale)
    {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) d[i] = cadd(d[i], cscale(s[i], scale));
    }
    void caxpy(Cx* d, const Cx* s, Cx scale)
    {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) d[i] = cadd(d[i], cmul(s[i], scale));
    }
    void clear(Cx* d) { std::fill(d, d + M, Cx{0.0, 0.0}); }
    void copy(Cx* d, const Cx* s) { std::memcpy(d, s, sizeof(Cx) * static_cast<size_t>(M)); }
    Cx sum_product(const Cx* x, const Cx* y)
    {
        std::vector<Cx> t(static_cast<size_t>(M));
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) t[static_cast<size_t>(i)] = cmul(x[i], y[i]);
        return sum(t.data());
    }
    Cx mean_product(const Cx* x, const Cx* y) { return cscale(sum_product(x, y), 1.0 / static_cast<double>(M)); }
    Cx sum_sq(const Cx* w) { return sum_product(w, w); }

    bool electrostatics() const { return lB > 0.0; }
    Cx gamma(int l) const { return eig[static_cast<size_t>(active[static_cast<size_t>(l)])] > 0.0 ? cx(0, 1) : cx(1, 0); }
    double Ocoef(int l, int alpha) const { return O[static_cast<size_t>(active[static_cast<size_t>(l)]) * nT + alpha]; }
    double omega_coefficient(int l) const
    {
        return rho0 * (1.0 / std::fabs(eig[static_cast<size_t>(active[static_cast<size_t>(l)])]));
    }

    // ---- setup -------------------------------------------------------------
    int build_modes(const fts_system* s);
    int build_ions(const fts_system* s);
    void build_kernels();
    void build_si_kernels();

    // ---- hot path (reference restatement) ------------------------------------
    void build_smoothed_sources(const Cx* om, const Cx* ps);
    Cx solve_species_prepared(int idx);
    void accumulate_species_densities(int idx, Cx Q);
    std::vector<Cx> solve_and_compute_densities(const Cx* om, const Cx* ps);
    void ion_compute_densities(const Cx* ps);
    std::vector<Cx> compute_densities(const Cx* om, const Cx* ps)
    {
        std::vector<Cx> Q = solve_and_compute_densities(om, ps);
        ion_compute_densities(ps);
        return Q;
    }
    Cx ion_partition_function(Cx* smoothed, Cx* boltz, const Cx* ps, const double* kernel, double z);
    void compute_drifts(const Cx* om, Cx* u);
    void compute_psi_drift(const Cx* ps, const Cx* rho_src, Cx* out);
    void remove_psi_zero_mode(Cx* ps);
    void generate_noises(long long t);
    void measure(const std::vector<Cx>& Qp, const Cx* om, const Cx* ps, fts_observables* rec);
    void stress_measure(const std::vector<Cx>& Qp, const Cx* om, const Cx* ps, Cx* out6);
    Cx isotropic_pressure(const std::vector<Cx>& Qp, const Cx* om, const Cx* ps);
    bool fields_finite() const;
    int step_pc(bool record, fts_observables* rec);
    // test hook: the step that ends at inject_step leaves a NaN in psi (once)
    void inject()
    {
        if (inject_step >= 0 && step_counter == inject_step) {
            psi[0].x = std::numeric_limits<double>::quiet_NaN();
            inject_step = -1;
        }
    }
    int step_em(bool record, fts_observables* rec, bool semi_implicit);
    int step_sipc(bool record, fts_observables* rec);
    void add_kspace_noise(Cx* w, double sigma, unsigned field, bool drop_zero, const std::vector<double>& P, bool cn);
};

// ---- interaction modes: interactionMode.cu:9-298 --------------------------
namespace {

double frob(const std::vector<double>& a)
{
    long double s = 0.0L;
    for (double x : a) s += static_cast<long double>(x) * x;
    return std::sqrt(static_cast<double>(s));
}

double offdiag(const std::vector<double>& a, int n)
{
    long double s = 0.0L;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j) s += static_cast<long double>(a[static_cast<size_t>(i) * n + j]) * a[static_cast<size_t>(i) * n + j];
    return std::sqrt(static_cast<double>(s));
}

// cyclic Jacobi with the rotation of interactionMode.cu:122-212
bool jacobi(const std::vector<double>& in, int n, std::vector<double>& ev, std::vector<double>& V)
{
    std::vector<double> A = in;
    V.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) V[static_cast<size_t>(i) * n + i] = 1.0;
    if (n == 1) {
        ev.assign(1, A[0]);
        return true;
    }
    const double scale = std::max(1.0, frob(in));
    const double tol = 1.0e-14 * scale;
    const int max_sweeps = std::max(50, 100 * n * n);
    bool converged = false;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        if (offdiag(A, n) <= tol) {
            converged = true;
            break;
        }
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const size_t pp = static_cast<size_t>(p) * n + p, qq = static_cast<size_t>(q) * n + q;
                const size_t pq = static_cast<size_t>(p) * n + q, qp = static_cast<size_t>(q) * n + p;
                const double apq = A[pq];
                if (std::fabs(apq) <= 2.220446049250313e-16 * scale) continue;
                const double app = A[pp], aqq = A[qq];
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau))
                                            : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = t * c;
                for (int k = 0; k < n; ++k) {
                    if (k == p || k == q) continue;
                    const size_t kp = static_cast<size_t>(k) * n + p, kq = static_cast<size_t>(k) * n + q;
                    const double akp = A[kp], akq = A[kq];
                    const double nkp = c * akp - s * akq, nkq = s * akp + c * akq;
                    A[kp] = nkp;
                    A[static_cast<size_t>(p) * n + k] = nkp;
                    A[kq] = nkq;
                    A[static_cast<size_t>(q) * n + k] = nkq;
                }
                A[pp] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
                A[qq] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
                A[pq] = 0.0;
                A[qp] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const size_t kp = static_cast<size_t>(k) * n + p, kq = static_cast<size_t>(k) * n + q;
                    const double vkp = V[kp], vkq = V[kq];
                    V[kp] = c * vkp - s * vkq;
                    V[kq] = s * vkp + c * vkq;
                }
                rotated = true;
            }
        }
        if (!rotated) {
            converged = true;
            break;
        }
    }
    if (!converged && offdiag(A, n) > 1.0e-10 * scale) return false;
    ev.assign(static_cast<size_t>(n), 0.0);
    for (int i = 0; i < n; ++i) ev[static_cast<size_t>(i)] = A[static_cast<size_t>(i) * n + i];
    return true;
}

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
    for (double x : chi) max_chi = std::max(max_chi, std::fabs(x));
    const double sym_tol = 1.0e-12 * std::max(1.0, max_chi);
    for (int al = 0; al < nT; ++al) {
        if (std::fabs(chi[static_cast<size_t>(al) * nT + al]) > sym_tol) return fail(FTS_INVALID, "chi diagonal entries must be zero");
        chi[static_cast<size_t>(al) * nT + al] = 0.0;
        for (int be = al + 1; be < nT; ++be) {
            const size_t ab = static_cast<size_t>(al) * nT + be, ba = static_cast<size_t>(be) * nT + al;
            if (std::fabs(chi[ab] - chi[ba]) > sym_tol) return fail(FTS_INVALID, "chi matrix must be symmetric");
            const double avg = 0.5 * (chi[ab] + chi[ba]);
            chi[ab] = avg;
            chi[ba] = avg;
        }
    }
    K.assign(static_cast<size_t>(nT) * nT, 0.0);
    for (int al = 0; al < nT; ++al)
        for (int be = 0; be < nT; ++be) K[static_cast<size_t>(al) * nT + be] = al == be ? xi : xi + chi[static_cast<size_t>(al) * nT + be];
    std::vector<double> raw_ev, raw_V;
    if (!jacobi(K, nT, raw_ev, raw_V)) return fail(FTS_INVALID, "Jacobi diagonalization did not converge");
    std::vector<int> order(static_cast<size_t>(nT));
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        const double dx = raw_ev[static_cast<size_t>(x)], dy = raw_ev[static_cast<size_t>(y)];
        if (dx != dy) return dx > dy;
        return x < y;
    });
    eig.assign(static_cast<size_t>(nT), 0.0);
    O.assign(static_cast<size_t>(nT) * nT, 0.0);
    for (int l = 0; l < nT; ++l) {
        const int raw = order[static_cast<size_t>(l)];
        eig[static_cast<size_t>(l)] = raw_ev[static_cast<size_t>(raw)];
        for (int al = 0; al < nT; ++al) O[static_cast<size_t>(l) * nT + al] = raw_V[static_cast<size_t>(al) * nT + raw];
        int pivot = 0;
        double pabs = std::fabs(O[static_cast<size_t>(l) * nT]);
        for (int al = 1; al < nT; ++al) {
            const double mag = std::fabs(O[static_cast<size_t>(l) * nT + al]);
            if (mag > pabs) {
                pabs = mag;
                pivot = al;
            }
        }
        if (O[static_cast<size_t>(l) * nT + pivot] < 0.0)
            for (int al = 0; al < nT; ++al) O[static_cast<size_t>(l) * nT + al] *= -1.0;
    }
    double max_abs = 0.0;
    for (double d : eig) max_abs = std::max(max_abs, std::fabs(d));
    const double tol_abs = tol_rel * std::max(1.0, max_abs);
    active.clear();
    for (int l = 0; l < nT; ++l)
        if (std::fabs(eig[static_cast<size_t>(l)]) > tol_abs) active.push_back(l);
    n_omega = static_cast<int>(active.size());
    // reconstruction / orthogonality checks (interactionMode.cu:265-298)
    std::vector<double> res(static_cast<size_t>(nT) * nT), orth(static_cast<size_t>(nT) * nT);
    for (int al = 0; al < nT; ++al)
        for (int be = 0; be < nT; ++be) {
            double v = 0.0;
            for (int l = 0; l < nT; ++l) v += O[static_cast<size_t>(l) * nT + al] * eig[static_cast<size_t>(l)] * O[static_cast<size_t>(l) * nT + be];
            res[static_cast<size_t>(al) * nT + be] = K[static_cast<size_t>(al) * nT + be] - v;
        }
    for (int l = 0; l < nT; ++l)
        for (int mm = 0; mm < nT; ++mm) {
            double dot = 0.0;
            for (int al = 0; al < nT; ++al) dot += O[static_cast<size_t>(l) * nT + al] * O[static_cast<size_t>(mm) * nT + al];
            orth[static_cast<size_t>(l) * nT + mm] = dot - (l == mm ? 1.0 : 0.0);
        }
    if (frob(res) / std::max(1.0, frob(K)) > 1.0e-10) return fail(FTS_INVALID, "interaction eigenmode reconstruction check failed");
    if (frob(orth) > 1.0e-10) return fail(FTS_INVALID, "interaction eigenvector orthogonality check failed");
    return FTS_OK;
}

// ---- ions: ion.cu:3-46, 117-134 --------------------------------------------
int fts_engine::build_ions(const fts_system* s)
{
    ions.clear();
    double polymer_charge = 0.0;
    for (const Species& sp : species) polymer_charge += static_cast<double>(sp.n) * sp.net_charge();
    const double def_a = s->default_smearing_length > 0.0 ? s->default_smearing_length : a;
    const bool combine = s->salt.combine_identical_species != 0;
    auto append = [&](double z, long long n, double aj) {
        if (n == 0) return;
        if (combine) {
            for (Ion& ion : ions) {
                if (std::fabs(ion.z - z) <= 1.0e-12 && std::fabs(ion.a - aj) <= 1.0e-12) {
                    ion.n += n;
                    return;
                }
            }
        }
        ions.push_back(Ion{z, n, aj});
    };
    if (s->ions_enabled && s->counterions.enabled) {
        const double z = s->counterions.valence;
        const double aj = s->counterions.smearing_length > 0.0 ? s->counterions.smearing_length : def_a;
        if (electrostatics() && z == 0.0) return fail(FTS_INVALID, "counterion valence");
        long long n = s->counterions.number;
        if (s->counterions.auto_number) {
            if (electrostatics()) {
                const double needed = -polymer_charge / z;
                if (needed < -1.0e-10) return fail(FTS_INVALID, "counterion sign");
                const double r = std::round(needed);
                if (std::fabs(needed - r) > 1.0e-8 * std::max(1.0, std::fabs(needed))) return fail(FTS_INVALID, "counterion number");
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
        if (electrostatics()) {
            if (sl.cation_valence == 0.0 || sl.anion_valence == 0.0) return fail(FTS_INVALID, "salt");
            if (std::fabs(sl.cation_stoich * sl.cation_valence + sl.anion_stoich * sl.anion_valence) > 1.0e-10) return fail(FTS_INVALID, "salt is not neutral");
        }
        append(sl.cation_valence, sl.num_salt * static_cast<long long>(sl.cation_stoich), aj);
        append(sl.anion_valence, sl.num_salt * static_cast<long long>(sl.anion_stoich), aj);
    }
    double total = polymer_charge, scale = std::fabs(polymer_charge);
    for (const Ion& ion : ions) {
        if (ion.n < 0 || (electrostatics() && ion.z == 0.0) || ion.a <= 0.0) return fail(FTS_INVALID, "ion species");
        total += ion.z * static_cast<double>(ion.n);
        scale += std::fabs(ion.z * static_cast<double>(ion.n));
    }
    if (electrostatics() && std::fabs(total) > 1.0e-8 * std::max(1.0, scale)) return fail(FTS_INVALID, "electroneutrality failed");
    if (ions.size() > FTS_MAX_IONS) return fail(FTS_INVALID, "too many ion species");
    return FTS_OK;
}

// ---- kernels: spectral.cu:3-35, ion.cu:48-51, stress.cu:106-149 ------------
void fts_engine::build_kernels()
{
    k2.assign(static_cast<size_t>(M), 0.0);
    Gamma.assign(static_cast<size_t>(M), 0.0);
    PHI.assign(std::max<size_t>(1, total_bonds) * static_cast<size_t>(M), 1.0);
    kab.assign(6 * static_cast<size_t>(M), 0.0);
    minus_kab.assign(6 * static_cast<size_t>(M), 0.0);
    ion_kernels.assign(ions.size() * static_cast<size_t>(M), 0.0);
    double sum_over_k2[6] = {0, 0, 0, 0, 0, 0};
    long long modes = 0;
    for (int x = 0; x < m1; ++x) {
        const double kx = wave_number(x, m1, L1);
        for (int y = 0; y < m2; ++y) {
            const double ky = wave_number(y, m2, L2);
            for (int z = 0; z < m3; ++z) {
                const double kz = wave_number(z, m3, L3);
                const size_t i = (static_cast<size_t>(x) * m2 + y) * m3 + z;
                const double ksq = kx * kx + ky * ky + kz * kz;
                k2[i] = ksq;
                Gamma[i] = std::exp(-0.5 * a * a * ksq);
                for (const Species& sp : species)
                    for (int s = 0; s < sp.N - 1; ++s) {
                        const double b = sp.bond[static_cast<size_t>(s)];
                        PHI[(sp.bond_offset + s) * static_cast<size_t>(M) + i] = std::exp(-b * b * ksq / 6.0);
                    }
                for (size_t j = 0; j < ions.size(); ++j) ion_kernels[j * M + i] = std::exp(-0.5 * ions[j].a * ions[j].a * ksq);
                const double v[6] = {kx * kx, ky * ky, kz * kz, kx * ky, kx * kz, ky * kz};
                for (int c = 0; c < 6; ++c) {
                    kab[static_cast<size_t>(c) * M + i] = v[c];
                    minus_kab[static_cast<size_t>(c) * M + i] = -v[c];
                }
                if (ksq > 0.0) {
                    ++modes;
                    for (int c = 0; c < 6; ++c) sum_over_k2[c] += v[c] / ksq;
                }
            }
        }
    }
    for (int c = 0; c < 6; ++c) {
        const double delta = c < 3 ? 1.0 : 0.0;
        mode_term[c] = electrostatics() ? ((1.0 / 3.0) * static_cast<double>(modes) * delta - sum_over_k2[c]) / V : 0.0;
    }
}

// ---- learned linear-response kernel (no reference counterpart) --------------
// The formulas are the ones include/fts.h states for FTS_SCHEME_SI_LEARNED.
void fts_engine::lrk_build_bins()
{
    lrk_binof.assign(static_cast<size_t>(M), 0);
    for (int x = 0; x < m1; ++x)
        for (int y = 0; y < m2; ++y)
            for (int z = 0; z < m3; ++z)
                lrk_binof[(static_cast<size_t>(x) * m2 + y) * m3 + z] =
                    oracle::lrk_bin(spectral_mode(x, m1), spectral_mode(y, m2), spectral_mode(z, m3), m1, m2, m3);
}

// std Re w0, std Im w0, mean Re w0, mean Im w0, std Re psi, std Im psi (population moments)
void fts_engine::lrk_moments(double* m)
{
    auto mom = [&](const Cx* f, double* out4) {
        long double s[4] = {0.0L, 0.0L, 0.0L, 0.0L};
        for (int i = 0; i < M; ++i) {
            s[0] += f[i].x;
            s[1] += static_cast<long double>(f[i].x) * f[i].x;
            s[2] += f[i].y;
            s[3] += static_cast<long double>(f[i].y) * f[i].y;
        }
        const double inv = 1.0 / static_cast<double>(M);
        const double mx = static_cast<double>(s[0]) * inv, my = static_cast<double>(s[2]) * inv;
        out4[0] = std::sqrt(std::max(0.0, static_cast<double>(s[1]) * inv - mx * mx));
        out4[1] = std::sqrt(std::max(0.0, static_cast<double>(s[3]) * inv - my * my));
        out4[2] = mx;
        out4[3] = my;
    };
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    if (n_omega > 0) mom(omega.data(), a);
    if (electrostatics()) mom(psi.data(), b);
    m[0] = a[0];
    m[1] = a[1];
    m[2] = a[2];
    m[3] = a[3];
    m[4] = b[0];
    m[5] = b[1];
}

// Called by the semi-implicit step after the drifts u(w_t) are known (drift, dpsi):
// judges the previous step with the exact force, writes its training record, and
// returns in mult[c] the per-k factor exp(r) that multiplies A_D for this step
// (empty: A_D unchanged).
void fts_engine::lrk_step(std::vector<double> (&mult)[2])
{
    constexpr int NB = FTS_LRK_BINS;
    Field Wc[2], Uc[2];
    for (int c = 0; c < 2; ++c) {
        Wc[c].assign(static_cast<size_t>(M), Cx{0.0, 0.0});
        Uc[c].assign(static_cast<size_t>(M), Cx{0.0, 0.0});
    }
    if (n_omega > 0) {
        copy(Wc[0].data(), omega.data());
        copy(Uc[0].data(), drift.data());
        fft->run(Wc[0].data(), -1);
        fft->run(Uc[0].data(), -1);
    }
    if (electrostatics()) {
        copy(Wc[1].data(), psi.data());
        copy(Uc[1].data(), dpsi.data());
        fft->run(Wc[1].data(), -1);
        fft->run(Uc[1].data(), -1);
    }
    double m[FTS_LRK_MOMENTS];
    lrk_moments(m);
    double r[2][NB];
    for (int b = 0; b < NB; ++b) {
        double f[FTS_LRK_FEATURES];
        oracle::lrk_features(b, m, f);
        oracle::LrkActs act;
        oracle::lrk_forward(lrk_P.data(), f, act);
        r[0][b] = act.r[0];
        r[1][b] = act.r[1];
    }
    bool use = false;
    if (lrk_have_prev) {
        double E[2] = {0.0, 0.0}, S1[2][NB] = {}, S2[2][NB] = {};
        for (int i = 1; i < M; ++i) {
            const int b = lrk_binof[static_cast<size_t>(i)];
            for (int c = 0; c < 2; ++c) {
                const Cx dW = csub(Wc[c][static_cast<size_t>(i)], lrk_W[c][static_cast<size_t>(i)]);
                const Cx dU = csub(Uc[c][static_cast<size_t>(i)], lrk_U[c][static_cast<size_t>(i)]);
                const double AD = c == 0 ? (n_omega > 0 ? A_omega[static_cast<size_t>(i)] : 0.0) : A_psi[static_cast<size_t>(i)];
                E[c] += dU.x * dU.x + dU.y * dU.y;
                S1[c][b] += AD * (dU.x * dW.x + dU.y * dW.y);
                S2[c][b] += AD * AD * (dW.x * dW.x + dW.y * dW.y);
            }
        }
        double ex = -std::numeric_limits<double>::infinity();
        bool any = false;
        for (int c = 0; c < 2; ++c) {
            if (!(E[c] > 0.0)) continue;
            double RL = E[c], RD = E[c];
            for (int b = 0; b < NB; ++b) {
                const double e = std::exp(lrk_r_prev[c][b]);
                RL += -2.0 * e * S1[c][b] + e * e * S2[c][b];
                RD += -2.0 * S1[c][b] + S2[c][b];
            }
            lrk_st.sum_rel_residual_learned[c] += RL / E[c];
            lrk_st.sum_rel_residual_debye[c] += RD / E[c];
            ex = std::max(ex, (RL - RD) / E[c]);
            any = true;
        }
        if (!any) ex = 0.0;
        ++lrk_st.judged;
        lrk_st.last_excess = ex;
        use = ex <= lrk_threshold;
        if (lrk_capacity > 0) {
            std::vector<double> rec(FTS_LRK_RECORD, 0.0);
            for (int j = 0; j < FTS_LRK_MOMENTS; ++j) rec[static_cast<size_t>(FTS_LRK_REC_MOMENTS + j)] = lrk_m_prev[j];
            for (int c = 0; c < 2; ++c) {
                rec[static_cast<size_t>(FTS_LRK_REC_E + c)] = E[c];
                for (int b = 0; b < NB; ++b) {
                    rec[static_cast<size_t>(FTS_LRK_REC_S1 + c * NB + b)] = S1[c][b];
                    rec[static_cast<size_t>(FTS_LRK_REC_S2 + c * NB + b)] = S2[c][b];
                }
            }
            lrk_ring.push_back(std::move(rec));
            if (static_cast<int>(lrk_ring.size()) > lrk_capacity) lrk_ring.pop_front();
            ++lrk_st.records;
        }
    }
    if (learned_scheme()) {
        if (use) ++lrk_st.learned_steps;
        else ++lrk_st.fallback_steps;
        if (use)
            for (int c = 0; c < 2; ++c) {
                mult[c].assign(static_cast<size_t>(M), 1.0);
                for (int i = 0; i < M; ++i) mult[c][static_cast<size_t>(i)] = std::exp(r[c][lrk_binof[static_cast<size_t>(i)]]);
            }
    }
    for (int c = 0; c < 2; ++c) {
        lrk_W[c] = std::move(Wc[c]);
        lrk_U[c] = std::move(Uc[c]);
        for (int b = 0; b < NB; ++b) lrk_r_prev[c][b] = r[c][b];
    }
    for (int j = 0; j < FTS_LRK_MOMENTS; ++j) lrk_m_prev[j] = m[j];
    lrk_have_prev = true;
}

// Linear-response kernels of the semi-implicit scheme (no reference counterpart).
//   A_l(k)   = c_l + s_l Gamma^2 sum_p (n_p/V) sum_{s,s'} O_{l a(s)} O_{l a(s')} phi_{ss'}(k)
//   A_psi(k) = k^2/(4 pi lB) + Gamma^2 sum_p (n_p/V) sum_{s,s'} z_s z_s' phi_{ss'}(k)
//              + sum_j z_j^2 (n_j/V) Gamma_j^2
// with phi_{ss'} the product of the bond factors between beads s and s', and
// s_l = +1 for gamma_l = i, -1 for gamma_l = 1.  The double sum uses the O(N)
// recurrence h_{t+1} = (h_t + w_t) phi_t.
void fts_engine::build_si_kernels()
{
    A_omega.assign(static_cast<size_t>(n_omega) * M, 0.0);
    A_psi.assign(static_cast<size_t>(M), 0.0);
    auto pair_sum = [&](const Species& sp, const std::vector<double>& w, double ksq) {
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
    for (int i = 0; i < M; ++i) {
        const double ksq = k2[static_cast<size_t>(i)];
        const double G2 = Gamma[static_cast<size_t>(i)] * Gamma[static_cast<size_t>(i)];
        for (int l = 0; l < n_omega; ++l) {
            double resp = 0.0;
            for (const Species& sp : species) {
                std::vector<double> w(static_cast<size_t>(sp.N));
                for (int s = 0; s < sp.N; ++s) w[static_cast<size_t>(s)] = Ocoef(l, sp.type[static_cast<size_t>(s)]);
                resp += static_cast<double>(sp.n) / V * pair_sum(sp, w, ksq);
            }
            const double sgn = gamma(l).y != 0.0 ? 1.0 : -1.0;
            A_omega[static_cast<size_t>(l) * M + i] = omega_coefficient(l) + sgn * G2 * resp;
        }
        if (electrostatics()) {
            double resp = 0.0;
            for (const Species& sp : species) resp += static_cast<double>(sp.n) / V * pair_sum(sp, sp.charge, ksq);
            double ion_resp = 0.0;
            for (size_t j = 0; j < ions.size(); ++j) {
                const double g = ion_kernels[j * M + i];
                ion_resp += ions[j].z * ions[j].z * static_cast<double>(ions[j].n) / V * g * g;
            }
            A_psi[static_cast<size_t>(i)] = ksq / (4.0 * Pi * lB) + G2 * resp + ion_resp;
        }
    }
}

// ---- polymer: polymer.cu:341-478 --------------------------------------------
void fts_engine::build_smoothed_sources(const Cx* om, const Cx* ps)
{
    for (int l = 0; l < n_omega; ++l) apply_kernel(om + static_cast<size_t>(l) * M, fld(omega_s, l), Gamma.data());
    for (int al = 0; al < nT; ++al) {
        Cx* Wa = fld(W_type_s, al);
        clear(Wa);
        for (int l = 0; l < n_omega; ++l) caxpy(Wa, fld(omega_s, l), cscale(gamma(l), Ocoef(l, al)));
    }
    if (electrostatics()) apply_kernel(ps, psi_s.data(), Gamma.data());
    else clear(psi_s.data());
}

Cx fts_engine::solve_species_prepared(int idx)
{
    const Species& sp = species[static_cast<size_t>(idx)];
    last_species = idx;
    // build_segment_fields + build_exponentials (polymer.cu:365-384)
    for (int s = 0; s < sp.N; ++s) {
        Cx* Ws = fld(W, s);
        copy(Ws, fld(W_type_s, sp.type[static_cast<size_t>(s)]));
        const double z = electrostatics() ? sp.charge[static_cast<size_t>(s)] : 0.0;
        if (z != 0.0) caxpy(Ws, psi_s.data(), cx(0.0, z));
        Cx* em = fld(exp_mW, s);
        Cx* ep = fld(exp_W, s);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) {
            em[i] = cexp(cx(-Ws[i].x, -Ws[i].y));
            ep[i] = cexp(Ws[i]);
        }
    }
    // propagate_chain (polymer.cu:402-416)
    copy(fld(qF, 0), fld(exp_mW, 0));
    copy(fld(qB, sp.N - 1), fld(exp_mW, sp.N - 1));
    for (int s = 0; s < sp.N - 1; ++s) {
        const double* phi = PHI.data() + (sp.bond_offset + s) * static_cast<size_t>(M);
        apply_kernel(fld(qF, s), fld(qF, s + 1), phi);
        mul(fld(qF, s + 1), fld(exp_mW, s + 1));
    }
    for (int s = sp.N - 1; s > 0; --s) {
        const double* phi = PHI.data() + (sp.bond_offset + s - 1) * static_cast<size_t>(M);
        apply_kernel(fld(qB, s), fld(qB, s - 1), phi);
        mul(fld(qB, s - 1), fld(exp_mW, s - 1));
    }
    return mean(fld(qF, sp.N - 1));
}

void fts_engine::accumulate_species_densities(int idx, Cx Q)
{
    const Species& sp = species[static_cast<size_t>(idx)];
    const double chain_density = static_cast<double>(sp.n) / V;
    for (int s = 0; s < sp.N; ++s) {
        const Cx* ep = fld(exp_W, s);
        const Cx* f = fld(qF, s);
        const Cx* b = fld(qB, s);
        auto accumulate = [&](Cx* rho, double weight) {
            if (weight == 0.0) return;
            const Cx factor = cscale(cinv(Q), chain_density * weight);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < M; ++i) rho[i] = cadd(rho[i], cmul(cmul(cmul(f[i], b[i]), ep[i]), factor));
        };
        accumulate(fld(rho_type, sp.type[static_cast<size_t>(s)]), 1.0);
        if (electrostatics()) accumulate(rho_c_poly.data(), sp.charge[static_cast<size_t>(s)]);
    }
}

std::vector<Cx> fts_engine::solve_and_compute_densities(const Cx* om, const Cx* ps)
{
    for (int al = 0; al < nT; ++al) clear(fld(rho_type, al));
    clear(rho_c_poly.data());
    clear(rho_c_ion.data());
    clear(rho_c.data());
    build_smoothed_sources(om, ps);
    std::vector<Cx> Qs;
    for (int idx = 0; idx < static_cast<int>(species.size()); ++idx) {
        const Cx Q = solve_species_prepared(idx);
        Qs.push_back(Q);
        accumulate_species_densities(idx, Q);
    }
    clear(rho_p.data());
    for (int al = 0; al < nT; ++al) {
        apply_kernel(fld(rho_type, al), fld(rho_type, al), Gamma.data());
        axpy(rho_p.data(), fld(rho_type, al), 1.0);
    }
    if (electrostatics()) {
        apply_kernel(rho_c_poly.data(), rho_c_poly.data(), Gamma.data());
        axpy(rho_c.data(), rho_c_poly.data(), 1.0);
    }
    return Qs;
}

// ---- ions: ion.cu:54-72, common.h:1104-1130 --------------------------------
Cx fts_engine::ion_partition_function(Cx* smoothed, Cx* boltz, const Cx* ps, const double* kernel, double z)
{
    apply_kernel(ps, smoothed, kernel);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) boltz[i] = cexp(cx(z * smoothed[i].y, -z * smoothed[i].x));
    return mean(boltz);
}

void fts_engine::ion_compute_densities(const Cx* ps)
{
    Q_ion.assign(ions.size(), cx(1.0, 0.0));
    for (size_t j = 0; j < ions.size(); ++j) {
        const double* kernel = ion_kernels.data() + j * M;
        Cx* rho = fld(ion_rho, j);
        const double nd = static_cast<double>(ions[j].n) / V;
        if (!electrostatics()) {
            std::fill(rho, rho + M, cx(nd, 0.0));
            continue;
        }
        const Cx Q = ion_partition_function(ion_psi.data(), ion_boltzmann.data(), ps, kernel, ions[j].z);
        Q_ion[j] = Q;
        const Cx factor = cscale(cinv(Q), nd);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) rho[i] = cmul(ion_boltzmann[static_cast<size_t>(i)], factor);
        apply_kernel(rho, ion_charge.data(), kernel);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) ion_charge[static_cast<size_t>(i)] = cscale(ion_charge[static_cast<size_t>(i)], ions[j].z);
        axpy(rho_c_ion.data(), ion_charge.data(), 1.0);
        axpy(rho_c.data(), ion_charge.data(), 1.0);
    }
}

// ---- drifts and noise: complexLangevin.cu:272-317 --------------------------
void fts_engine::compute_drifts(const Cx* om, Cx* u)
{
    for (int l = 0; l < n_omega; ++l) {
        Cx* ul = u + static_cast<size_t>(l) * M;
        clear(ul);
        axpy(ul, om + static_cast<size_t>(l) * M, omega_coefficient(l));
        for (int al = 0; al < nT; ++al) caxpy(ul, fld(rho_type, al), cscale(gamma(l), Ocoef(l, al)));
    }
}

void fts_engine::remove_psi_zero_mode(Cx* ps)
{
    if (electrostatics()) {
        const Cx m = cscale(mean(ps), -1.0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) ps[i] = cadd(ps[i], m);
    } else {
        clear(ps);
    }
}

void fts_engine::compute_psi_drift(const Cx* ps, const Cx* rho_src, Cx* out)
{
    if (!electrostatics()) {
        clear(out);
        return;
    }
    apply_kernel(ps, out, k2.data());
    const double s = 1.0 / (4.0 * Pi * lB);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) {
        out[i] = cscale(out[i], s);
        out[i].x -= rho_src[i].y;
        out[i].y += rho_src[i].x;
    }
    remove_psi_zero_mode(out);
}

void fts_engine::generate_noises(long long t)
{
    for (int l = 0; l < n_omega; ++l) {
        const double sigma = (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_omega[static_cast<size_t>(l)] * dt / dV);
        double* eta = eta_omega.data() + static_cast<size_t>(l) * M;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) eta[i] = sigma * oracle::cl_normal(seed, static_cast<uint64_t>(t), static_cast<uint32_t>(l), static_cast<uint32_t>(i));
    }
    const double sp = electrostatics() ? (noise ? 1.0 : 0.0) * std::sqrt(2.0 * lambda_psi * dt / dV) : 0.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; ++i) etaPsi[static_cast<size_t>(i)] = sp * oracle::cl_normal(seed, static_cast<uint64_t>(t), 0x100u, static_cast<uint32_t>(i));
    if (electrostatics()) {
        const double m = sum_real(etaPsi.data()) / static_cast<double>(M);
        for (int i = 0; i < M; ++i) etaPsi[static_cast<size_t>(i)] += -m;
    }
}

// ---- observables: observable.cu:53-317, stress.cu:161-321 ------------------
void fts_engine::stress_measure(const std::vector<Cx>& Qp, const Cx* om, const Cx* ps, Cx* out6)
{
    long long ion_count = 0;
    for (const Ion& ion : ions) ion_count += ion.n;
    long long chains = 0;
    for (const Species& sp : species) chains += sp.n;
    const double ideal = (static_cast<double>(chains) + static_cast<double>(ion_count)) / V;
    double P_sr = 0.0; // param.cu:135-149
    {
        long double wd = 0.0L;
        for (int al = 0; al < nT; ++al) wd += static_cast<long double>(total_type_segments[static_cast<size_t>(al)]) * K[static_cast<size_t>(al) * nT + al];
        if (a > 0.0 && total_segments > 0) {
            const long double md = wd / static_cast<long double>(total_segments);
            const long double den = 16.0L * static_cast<long double>(Pi) * std::sqrt(static_cast<long double>(Pi)) * a * a * a;
            P_sr = static_cast<double>(md / den);
        }
    }
    const int nc = 6;
    for (int c = 0; c < nc; ++c) {
        const double delta = c < 3 ? 1.0 : 0.0;
        out6[c] = cx((ideal + P_sr) * delta + mode_term[c], 0.0);
        if (electrostatics()) {
            apply_kernel(ps, sa.data(), k2.data());
            const Cx g2 = mean_product(ps, sa.data());
            apply_kernel(ps, sa.data(), kab.data() + static_cast<size_t>(c) * M);
            const Cx gab = mean_product(ps, sa.data());
            out6[c] = cadd(out6[c], cscale(csub(gab, cscale(g2, delta / 3.0)), 1.0 / (4.0 * Pi * lB)));
        }
    }
    build_smoothed_sources(om, ps);
    const double a2 = a * a;
    for (int idx = 0; idx < static_cast<int>(species.size()); ++idx) {
        const Cx solved = solve_species_prepared(idx);
        const Cx Q = idx < static_cast<int>(Qp.size()) ? Qp[static_cast<size_t>(idx)] : solved;
        const Species& sp = species[static_cast<size_t>(idx)];
        for (int c = 0; c < nc; ++c) {
            const double delta = c < 3 ? 1.0 : 0.0;
            const double* mk = minus_kab.data() + static_cast<size_t>(c) * M;
            // chain_connectivity_term (stress.cu:224-251)
            Cx total = cx(0, 0);
            for (int s = 0; s < sp.N - 1; ++s) {
                copy(sa.data(), fld(exp_W, s));
                mul(sa.data(), fld(qB, s));
                apply_kernel(sa.data(), sb.data(), mk);
                copy(sc.data(), fld(qF, s));
                mul(sc.data(), sb.data());
                const Cx avg = mean(sc.data());
                const double b = sp.bond[static_cast<size_t>(s)];
                total = cadd(total, cscale(avg, b * b / 3.0));
            }
            out6[c] = cadd(out6[c], cmul(cscale(cinv(Q), -static_cast<double>(sp.n) / V), total));
            // polymer_field_coupling_term (stress.cu:253-295)
            if (electrostatics()) apply_kernel(psi_s.data(), dpsi_s.data(), mk);
            else clear(dpsi_s.data());
            total = cx(0, 0);
            for (int s = 0; s < sp.N; ++s) {
                const Cx* Wa = fld(W_type_s, sp.type[static_cast<size_t>(s)]);
                apply_kernel(Wa, dfield.data(), mk);
                clear(sb.data());
                axpy(sb.data(), dfield.data(), a2);
                if (c < 3) axpy(sb.data(), Wa, 0.5);
                if (electrostatics()) {
                    caxpy(sb.data(), dpsi_s.data(), cx(0.0, sp.charge[static_cast<size_t>(s)] * a2));
                    if (c < 3) caxpy(sb.data(), psi_s.data(), cx(0.0, sp.charge[static_cast<size_t>(s)] / 6.0));
                }
                copy(sa.data(), fld(qF, s));
                mul(sa.data(), fld(qB, s));
                mul(sa.data(), fld(exp_W, s));
                copy(sc.data(), sa.data());
                mul(sc.data(), sb.data());
                total = cadd(total, mean(sc.data()));
            }
            out6[c] = cadd(out6[c], cmul(cscale(cinv(Q), static_cast<double>(sp.n) / V), total));
            (void)delta;
        }
    }
    if (electrostatics()) {
        for (int c = 0; c < nc; ++c) {
            const double* mk = minus_kab.data() + static_cast<size_t>(c) * M;
            Cx total = cx(0, 0);
            for (size_t j = 0; j < ions.size(); ++j) {
                const Cx Qj = ion_partition_function(ion_psi.data(), ion_boltzmann.data(), ps, ion_kernels.data() + j * M, ions[j].z);
                apply_kernel(ion_psi.data(), sa.data(), mk);
                const double aj2 = ions[j].a * ions[j].a;
#pragma omp parallel for schedule(static)
                for (int i = 0; i < M; ++i) sa[static_cast<size_t>(i)] = cscale(sa[static_cast<size_t>(i)], aj2);
                if (c < 3) axpy(sa.data(), ion_psi.data(), 1.0 / 6.0);
                const Cx avg = mean_product(ion_boltzmann.data(), sa.data());
                Cx pref = cscale(cinv(Qj), static_cast<double>(ions[j].n) * ions[j].z / V);
                pref = cmul(cx(0.0, 1.0), pref);
                total = cadd(total, cmul(pref, avg));
            }
            out6[c] = cadd(out6[c], total);
        }
    }
}

Cx fts_engine::isotropic_pressure(const std::vector<Cx>& Qp, const Cx* om, const Cx* ps)
{
    long long ion_count = 0;
    for (const Ion& ion : ions) ion_count += ion.n;
    long long chains = 0;
    for (const Species& sp : species) chains += sp.n;
    double P_sr = 0.0;
    {
        long double wd = 0.0L;
        for (int al = 0; al < nT; ++al) wd += static_cast<long double>(total_type_segments[static_cast<size_t>(al)]) * K[static_cast<size_t>(al) * nT + al];
        if (a > 0.0 && total_segments > 0) {
            const long double md = wd / static_cast<long double>(total_segments);
            P_sr = static_cast<double>(md / (16.0L * static_cast<long double>(Pi) * std::sqrt(static_cast<long double>(Pi)) * a * a * a));
        }
    }
    Cx value = cx((static_cast<double>(chains) + static_cast<double>(ion_count)) / V + P_sr, 0.0);
    build_smoothed_sources(om, ps);
    const double a2 = a * a;
    // build_isotropic_trace_block (observable.cu:184-195): -(l^2/3) S_{k^2} f + c f
    auto trace_block = [&](Cx* out, const Cx* f, double l2, double c) {
        apply_kernel(f, out, k2.data());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; ++i) out[i] = cscale(out[i], -l2 / 3.0);
        axpy(out, f, c);
    };
    for (int idx = 0; idx < static_cast<int>(species.size()); ++idx) {
        const Cx solved = solve_species_prepared(idx);
        const Cx Q = idx < static_cast<int>(Qp.size()) ? Qp[static_cast<size_t>(idx)] : solved;
        const Species& sp = species[static_cast<size_t>(idx)];
        Cx total = cx(0, 0);
        for (int s = 0; s < sp.N; ++s) {
            trace_block(sb.data(), fld(W_type_s, sp.type[static_cast<size_t>(s)]), a2, 0.5);
            if (electrostatics()) {
                trace_block(sc.data(), psi_s.data(), a2, 1.0 / 6.0);
                caxpy(sb.data(), sc.data(), cx(0.0, sp.charge[static_cast<size_t>(s)]));
            }
            copy(sa.data(), fld(qF, s));
            mul(sa.data(), fld(qB, s));
            mul(sa.data(), fld(exp_W, s));
            mul(sa.data(), sb.data());
            total = cadd(total, mean(sa.data()));
        }
        value = cadd(value, cmul(cscale(cinv(Q), static_cast<double>(sp.n) / V), total));
        total = cx(0, 0);
        for (int s = 0; s < sp.N - 1; ++s) {
            copy(sa.data(), fld(exp_W, s));
            mul(sa.data(), fld(qB, s));
            apply_kernel(sa.data(), sb.data(), k2.data());
            copy(sc.data(), fld(qF, s));
            mul(sc.data(), sb.data());
            const double b = sp.bond[static_cast<size_t>(s)];
            total = cadd(total, cscale(mean(sc.data()), b * b / 9.0));
        }
        value = cadd(value, cmul(cscale(cinv(Q), static_cast<double>(sp.n) / V), total));
    }
    if (electrostatics()) {
        Cx total = cx(0, 0);
        for (size_t j = 0; j < ions.size(); ++j) {
            const Cx Qj = ion_partition_function(ion_psi.data(), ion_boltzmann.data(), ps, ion_kernels.data() + j * M, ions[j].z);
            trace_block(sb.data(), ion_psi.data(), ions[j].a * ions[j].a, 1.0 / 6.0);
            copy(sc.data(), ion_boltzmann.data());
            mul(sc.data(), sb.data());
            const Cx avg = mean(sc.data());
            Cx pref = cscale(cinv(Qj), static_cast<double>(ions[j].n) * ions[j].z / V);
            pref = cmul(cx(0.0, 1.0), pref);
            total = cadd(total, cmul(pref, avg));
        }
        value = cadd(value, total);
    }
    return value;
}

void fts_engine::measure(const std::vector<Cx>& Qp, const Cx* om, const Cx* ps, fts_observables* rec)
{
    std::memset(rec, 0, sizeof(*rec));
    rec->step = step_counter;
    rec->n_species = static_cast<int>(species.size());
    rec->n_ions = static_cast<int>(ions.size());
    rec->stress_mask = stress_mask;
    std::vector<Cx> chem;
    for (size_t p = 0; p < species.size(); ++p) {
        const Cx mu = cneglog(Qp[p]);
        chem.push_back(mu);
        rec->Q_poly[p][0] = Qp[p].x;
        rec->Q_poly[p][1] = Qp[p].y;
        rec->mu_ex_poly[p][0] = mu.x;
        rec->mu_ex_poly[p][1] = mu.y;
        rec->mu_id_poly[p] = std::log(static_cast<double>(species[p].n) / V);
        rec->mu_self_poly[p] = -lB * species[p].charge_square_sum() / (2.0 * std::sqrt(Pi) * a);
    }
    for (size_t j = 0; j < ions.size(); ++j) {
        Cx mu = cx(0, 0);
        Cx Qj = cx(1, 0);
        if (electrostatics()) {
            Qj = ion_partition_function(obs_a.data(), obs_b.data(), ps, ion_kernels.data() + j * M, ions[j].z);
            mu = cneglog(Qj);
        }
        chem.push_back(mu);
        rec->Q_ion[j][0] = Qj.x;
        rec->Q_ion[j][1] = Qj.y;
        rec->mu_ex_ion[j][0] = mu.x;
        rec->mu_ex_ion[j][1] = mu.y;
        rec->mu_id_ion[j] = std::log(static_cast<double>(ions[j].n) / V);
        rec->mu_self_ion[j] = -lB * ions[j].z * ions[j].z / (2.0 * std::sqrt(Pi) * ions[j].a);
    }
    // mean_field_hamiltonian_density (observable.cu:141-182)
    Cx h = cx(0, 0);
    for (int l = 0; l < n_omega; ++l) {
        const Cx* w = om + static_cast<size_t>(l) * M;
        h = cadd(h, cscale(sum_sq(w), 0.5 * omega_coefficient(l) / static_cast<double>(M)));
    }
    if (electrostatics()) {
        apply_kernel(ps, obs_a.data(), k2.data());
        h = cadd(h, cscale(sum_product(ps, obs_a.data()), 1.0 / (8.0 * Pi * lB * static_cast<double>(M))));
    }
    size_t off = 0;
    for (const Species& sp : species) h = cadd(h, cscale(chem[off++], static_cast<double>(sp.n) / V));
    for (size_t j = 0; j < ions.size(); ++j) h = cadd(h, cscale(chem[off + j], static_cast<double>(ions[j].n) / V));
    rec->H[0] = h.x;
    rec->H[1] = h.y;
    if (stress_mask & 0x3F) {
        Cx s6[6];
        stress_measure(Qp, om, ps, s6);
        for (int c = 0; c < 6; ++c)
            if (stress_mask & (1 << c)) {
                rec->stress[c][0] = s6[c].x;
                rec->stress[c][1] = s6[c].y;
            }
    }
    if (stress_mask & FTS_STRESS_PISO) {
        const Cx p = isotropic_pressure(Qp, om, ps);
        rec->p_iso[0] = p.x;
        rec->p_iso[1] = p.y;
    }
}

bool fts_engine::fields_finite() const
{
    for (const Cx& z : omega)
        if (!finite(z)) return false;
    for (const Cx& z : psi)
        if (!finite(z)) return false;
    return true;
}

// run_step (complexLangevin.cu:669-698)
int fts_engine::step_pc(bool record, fts_observables* rec)
{
    std::vector<Cx> Qp = compute_densities(omega.data(), psi.data());
    if (!partition_finite(Qp)) return fail(FTS_NONFINITE, "non-finite Q");
    copy(rho_c_old.data(), rho_c.data());
    generate_noises(step_counter);
    compute_drifts(omega.data(), drift.data());
    // predictor_step (complexLangevin.cu:328-347)
    compute_psi_drift(psi.data(), rho_c_old.data(), dpsi.data());
    for (int l = 0; l < n_omega; ++l) {
        Cx* pred = fld(omega_bar, l);
        copy(pred, fld(omega, l));
        axpy(pred, fld(drift, l), -dt * lambda_omega[static_cast<size_t>(l)]);
        const double* eta = eta_omega.data() + static_cast<size_t>(l) * M;
        for (int i = 0; i < M; ++i) pred[i].x += 1.0 * eta[i];
    }
    copy(psi_bar.data(), psi.data());
    if (electrostatics()) {
        axpy(psi_bar.data(), dpsi.data(), -dt * lambda_psi);
        for (int i = 0; i < M; ++i) psi_bar[static_cast<size_t>(i)].x += 1.0 * etaPsi[static_cast<size_t>(i)];
    } else {
        clear(psi_bar.data());
    }
    remove_psi_zero_mode(psi_bar.data());
    Qp = compute_densities(omega_bar.data(), psi_bar.data());
    if (!partition_finite(Qp)) return fail(FTS_NONFINITE, "non-finite Qbar");
    compute_psi_drift(psi_bar.data(), rho_c.data(), dpsi_bar.data());
    if (record) measure(Qp, omega_bar.data(), psi_bar.data(), rec);
    compute_drifts(omega_bar.data(), drift_bar.data());
    drift_statistics(true); // write_drift_max before corrector_step (complexLangevin.cu:690-692)
    // corrector_step (complexLangevin.cu:349-368)
    for (int l = 0; l < n_omega; ++l) {
        Cx* w = fld(omega, l);
        const double lam = lambda_omega[static_cast<size_t>(l)];
        axpy(w, fld(drift, l), -0.5 * dt * lam);
        axpy(w, fld(drift_bar, l), -0.5 * dt * lam);
        const double* eta = eta_omega.data() + static_cast<size_t>(l) * M;
        for (int i = 0; i < M; ++i) w[i].x += 1.0 * eta[i];
    }
    if (electrostatics()) {
        axpy(psi.data(), dpsi.data(), -0.5 * dt * lambda_psi);
        axpy(psi.data(), dpsi_bar.data(), -0.5 * dt * lambda_psi);
        for (int i = 0; i < M; ++i) psi[static_cast<size_t>(i)].x += 1.0 * etaPsi[static_cast<size_t>(i)];
    }
    remove_psi_zero_mode(psi.data());
    ++step_counter;
    inject();
    if (!fields_finite()) return fail(FTS_NONFINITE, "updated fields have non-finite values");
    return FTS_OK;
}

// Plain Euler-Maruyama and the semi-implicit variant (no reference counterpart):
//   EM: w <- w - dt lam u(w) + eta
//   SI: w <- w + S_P[-dt lam u(w) + eta],  P(k) = 1 / (1 + dt lam A(k))
// psi likewise with its own kernel; psi's zero mode is removed afterwards.
// The record, when asked for, is taken at the state the step starts from.
int fts_engine::step_em(bool record, fts_observables* rec, bool semi_implicit)
{
    std::vector<Cx> Qp = compute_densities(omega.data(), psi.data());
    if (!partition_finite(Qp)) return fail(FTS_NONFINITE, "non-finite Q");
    if (record) measure(Qp, omega.data(), psi.data(), rec);
    generate_noises(step_counter);
    compute_drifts(omega.data(), drift.data());
    compute_psi_drift(psi.data(), rho_c.data(), dpsi.data());
    drift_statistics(false);
    std::vector<double> mult[2];
    if (semi_implicit && lrk_active()) lrk_step(mult);
    // Semi-implicit: w += S_P[-dt lam u] + S_Pn[eta] with P = 1 / (1 + a), a = dt lam A(k),
    // and the noise factor Pn = sqrt(P (1 + P) / 2) = P sqrt(1 + a / 2): for a linear drift
    // u = A w the stationary variance of every mode is then exactly 1 / A at any dt
    // (eta transformed by P would give 2 / (A (2 + a))).  The noise is drawn in k-space:
    // FFT[eta] = sigma sqrt(M) g(k), g Hermitian (the pair {k, -k} takes the two
    // Box-Muller normals a, b of one Philox draw, field F, site = the pair's smaller index:
    // g = (a +- i b) / sqrt 2, + at the smaller index; a self-conjugate k takes a alone),
    // psi's without k = 0.  EM adds the
    // real-space eta to the drift.
    const bool split = semi_implicit && noise;
    std::vector<double> P(static_cast<size_t>(M));
    auto add_noise = [&](Cx* w, double sigma, unsigned field, bool drop_zero) { add_kspace_noise(w, sigma, field, drop_zero, P, false); };
    for (int l = 0; l < n_omega; ++l) {
        const double lam = lambda_omega[static_cast<size_t>(l)];
        Cx* w = fld(omega, l);
        Cx* r = fld(drift, l);
        const double* eta = eta_omega.data() + static_cast<size_t>(l) * M;
        for (int i = 0; i < M; ++i) {
            r[i] = cscale(r[i], -dt * lam);
            if (!semi_implicit) r[i].x += eta[i];
        }
        if (semi_implicit) {
            for (int i = 0; i < M; ++i) {
                const double A = A_omega[static_cast<size_t>(l) * M + i] * (l == 0 && !mult[0].empty() ? mult[0][static_cast<size_t>(i)] : 1.0);
                P[static_cast<size_t>(i)] = 1.0 / (1.0 + dt * lam * A);
            }
            apply_kernel(r, r, P.data());
        }
        axpy(w, r, 1.0);
        if (split) add_noise(w, std::sqrt(2.0 * lam * dt / dV), 0x400u + 2u * static_cast<unsigned>(l), false);
    }
    if (electrostatics()) {
        Cx* r = dpsi.data();
        for (int i = 0; i < M; ++i) {
            r[i] = cscale(r[i], -dt * lambda_psi);
            if (!semi_implicit) r[i].x += etaPsi[static_cast<size_t>(i)];
        }
        if (semi_implicit) {
            for (int i = 0; i < M; ++i) {
                const double A = A_psi[static_cast<size_t>(i)] * (!mult[1].empty() ? mult[1][static_cast<size_t>(i)] : 1.0);
                P[static_cast<size_t>(i)] = 1.0 / (1.0 + dt * lambda_psi * A);
            }
            apply_kernel(r, r, P.data());
        }
        axpy(psi.data(), r, 1.0);
        if (split) add_noise(psi.data(), std::sqrt(2.0 * lambda_psi * dt / dV), 0x500u, true);
    }
    remove_psi_zero_mode(psi.data());
    ++step_counter;
    inject();
    if (!fields_finite()) return fail(FTS_NONFINITE, "updated fields have non-finite values");
    return FTS_OK;
}

// The k-space noise of a semi-implicit step added to w: IFFT[sigma sqrt(M) f(P) g(k)] / M
// with g the Hermitian spectrum of include/fts.h (one Philox draw per pair {k, -k}; field
// `field`, the step's counter), f(P) = sqrt(P (1 + P) / 2) for the semi-implicit step and
// f(P) = P2 = 2 P / (1 + P) for the predictor-corrector's corrector (cn); k = 0 dropped
// when drop_zero.
void fts_engine::add_kspace_noise(Cx* w, double sigma, unsigned field, bool drop_zero, const std::vector<double>& P, bool cn)
{
    std::vector<Cx> nz(static_cast<size_t>(M));
    const double amp0 = sigma / std::sqrt(static_cast<double>(M));
    for (int x = 0; x < m1; ++x)
        for (int y = 0; y < m2; ++y)
            for (int z = 0; z < m3; ++z) {
                const int gi = (x * m2 + y) * m3 + z;
                const int pi = (((m1 - x) % m1) * m2 + (m2 - y) % m2) * m3 + (m3 - z) % m3;
                const double t = P[static_cast<size_t>(gi)];
                const double amp = amp0 * (cn ? 2.0 * t / (1.0 + t) : std::sqrt(0.5 * t * (1.0 + t)));
                Cx v{0.0, 0.0};
                if (drop_zero && gi == 0) {
                } else if (pi == gi) {
                    v.x = amp * oracle::cl_normal(seed, static_cast<uint64_t>(step_counter), field, static_cast<uint32_t>(gi));
                } else {
                    double ab[2];
                    oracle::cl_normal_pair(seed, static_cast<uint64_t>(step_counter), field, static_cast<uint32_t>(std::min(gi, pi)), ab);
                    const double h = amp * 0.70710678118654752440;
                    v = Cx{h * ab[0], gi < pi ? h * ab[1] : -h * ab[1]};
                }
                nz[static_cast<size_t>(gi)] = v;
            }
    fft->run(nz.data(), +1);
    axpy(w, nz.data(), 1.0);
}

// Semi-implicit predictor-corrector (FTS_SCHEME_SIPC / _SIPC_LEARNED, no reference
// counterpart; include/fts.h states it):
//   predictor  w_bar = w + S_P[-dt lam u(w)] + S_Pn[eta]                      (the SI step)
//   corrector  w    <- w + S_P2[-(dt lam / 2)(u(w) + u(w_bar))] + S_P2[eta] + S_{1-P2}[w_bar - w]
// P = 1 / (1 + a), P2 = 2 P / (1 + P) = 1 / (1 + a / 2), a = dt lam A(k) (A_D, times the
// learned factor when the judge keeps it), one k-space draw of eta for both.  The record
// is taken at the predictor state with its Q (as step_pc, complexLangevin.cu:680-684).
int fts_engine::step_sipc(bool record, fts_observables* rec)
{
    std::vector<Cx> Qp = compute_densities(omega.data(), psi.data());
    if (!partition_finite(Qp)) return fail(FTS_NONFINITE, "non-finite Q");
    compute_drifts(omega.data(), drift.data());
    compute_psi_drift(psi.data(), rho_c.data(), dpsi.data());
    drift_statistics(false);
    std::vector<double> mult[2];
    if (lrk_active()) lrk_step(mult);
    const size_t Ms = static_cast<size_t>(M);
    // per field: P, -dt lam u(w) (u0) and the predictor's increment (d)
    const int nf = n_omega + (electrostatics() ? 1 : 0);
    std::vector<std::vector<double>> P(static_cast<size_t>(nf), std::vector<double>(Ms));
    std::vector<Field> u0(static_cast<size_t>(nf), Field(Ms)), d(static_cast<size_t>(nf), Field(Ms));
    for (int f = 0; f < nf; ++f) {
        const bool is_psi = f == n_omega;
        const double lam = is_psi ? lambda_psi : lambda_omega[static_cast<size_t>(f)];
        const Cx* u = is_psi ? dpsi.data() : fld(drift, f);
        for (size_t i = 0; i < Ms; ++i) {
            const double A = is_psi ? A_psi[i] * (!mult[1].empty() ? mult[1][i] : 1.0)
                                    : A_omega[static_cast<size_t>(f) * Ms + i] * (f == 0 && !mult[0].empty() ? mult[0][i] : 1.0);
            P[static_cast<size_t>(f)][i] = 1.0 / (1.0 + dt * lam * A);
            u0[static_cast<size_t>(f)][i] = cscale(u[i], -dt * lam);
        }
        apply_kernel(u0[static_cast<size_t>(f)].data(), d[static_cast<size_t>(f)].data(), P[static_cast<size_t>(f)].data());
        if (noise)
            add_kspace_noise(d[static_cast<size_t>(f)].data(), std::sqrt(2.0 * lam * dt / dV), is_psi ? 0x500u : 0x400u + 2u * static_cast<unsigned>(f),
                             is_psi, P[static_cast<size_t>(f)], false);
    }
    for (int l = 0; l < n_omega; ++l) {
        copy(fld(omega_bar, l), fld(omega, l));
        axpy(fld(omega_bar, l), d[static_cast<size_t>(l)].data(), 1.0);
    }
    copy(psi_bar.data(), psi.data());
    if (electrostatics()) axpy(psi_bar.data(), d[static_cast<size_t>(n_omega)].data(), 1.0);
    else clear(psi_bar.data());
    remove_psi_zero_mode(psi_bar.data());
    Qp = compute_densities(omega_bar.data(), psi_bar.data());
    if (!partition_finite(Qp)) return fail(FTS_NONFINITE, "non-finite Qbar");
    compute_psi_drift(psi_bar.data(), rho_c.data(), dpsi_bar.data());
    if (record) measure(Qp, omega_bar.data(), psi_bar.data(), rec);
    compute_drifts(omega_bar.data(), drift_bar.data());
    std::vector<double> P2(Ms), Q2(Ms);
    Field r(Ms);
    for (int f = 0; f < nf; ++f) {
        const bool is_psi = f == n_omega;
        const double lam = is_psi ? lambda_psi : lambda_omega[static_cast<size_t>(f)];
        const Cx* ub = is_psi ? dpsi_bar.data() : fld(drift_bar, f);
        for (size_t i = 0; i < Ms; ++i) {
            const double t = P[static_cast<size_t>(f)][i];
            P2[i] = 2.0 * t / (1.0 + t);
            Q2[i] = (1.0 - t) / (1.0 + t);
            r[i] = cscale(cadd(u0[static_cast<size_t>(f)][i], cscale(ub[i], -dt * lam)), 0.5);
        }
        Cx* w = is_psi ? psi.data() : fld(omega, f);
        apply_kernel(r.data(), r.data(), P2.data());
        if (noise) add_kspace_noise(r.data(), std::sqrt(2.0 * lam * dt / dV), is_psi ? 0x500u : 0x400u + 2u * static_cast<unsigned>(f), is_psi, P[static_cast<size_t>(f)], true);
        apply_kernel(d[static_cast<size_t>(f)].data(), d[static_cast<size_t>(f)].data(), Q2.data());
        axpy(w, r.data(), 1.0);
        axpy(w, d[static_cast<size_t>(f)].data(), 1.0);
    }
    remove_psi_zero_mode(psi.data());
    ++step_counter;
    inject();
    if (!fields_finite()) return fail(FTS_NONFINITE, "updated fields have non-finite values");
    return FTS_OK;
}

// ---------------------------------------------------------------------------
// C ABI (prefix fto_)
// ---------------------------------------------------------------------------
namespace {

bool supported_length(int n)
{
    if (n <= 0) return false;
    int m = n;
    for (int p = 2; p <= 61 && m > 1; ++p)
        while (m % p == 0) m /= p;
    return m == 1;
}

} // namespace

extern "C" {

const char* fto_last_error(void) { return g_error.c_str(); }

int fto_create(const fts_system* s, fts_engine** out)
{
    if (!s || !out) return fail(FTS_INVALID, "null argument");
    *out = nullptr;
    if (s->m1 <= 0 || s->m2 <= 0 || s->m3 <= 0) return fail(FTS_INVALID, "mesh");
    const long long msize = static_cast<long long>(s->m1) * s->m2 * s->m3;
    if (msize > 2147483647LL) return fail(FTS_INVALID, "mesh size");
    if (!supported_length(s->m1) || !supported_length(s->m2) || !supported_length(s->m3)) return fail(FTS_INVALID, "mesh length has a prime factor above 61");
    if (s->a <= 0.0 || s->dL <= 0.0 || s->lB < 0.0) return fail(FTS_INVALID, "a, dL, or lB");
    if (s->dt < 0.0 || s->lambda_psi < 0.0) return fail(FTS_INVALID, "dt or lambdaPsi");
    if (s->n_species <= 0 || s->n_species > FTS_MAX_SPECIES || !s->species) return fail(FTS_INVALID, "polymer species_count");
    if (s->scheme < 0 || s->scheme > 5) return fail(FTS_INVALID, "scheme");
    fts_engine* e = new fts_engine();
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
    e->seed = s->seed;
    e->seed0 = s->seed;
    e->stress_mask = s->stress_mask;
    int rc = e->build_modes(s);
    if (rc != FTS_OK) {
        delete e;
        return rc;
    }
    // species (polymer.cu:123-160, 293-300)
    e->total_type_segments.assign(static_cast<size_t>(e->nT), 0);
    for (int p = 0; p < s->n_species; ++p) {
        const fts_species_desc& d = s->species[p];
        if (d.N <= 0 || d.num_chains <= 0 || !d.types || !d.charges) {
            delete e;
            return fail(FTS_INVALID, "polymer species");
        }
        Species sp;
        sp.n = d.num_chains;
        sp.N = d.N;
        sp.type.assign(d.types, d.types + d.N);
        sp.charge.assign(d.charges, d.charges + d.N);
        sp.bond.assign(static_cast<size_t>(std::max(0, d.N - 1)), 1.0);
        for (int b = 0; b + 1 < d.N; ++b)
            if (d.bonds) sp.bond[static_cast<size_t>(b)] = d.bonds[b];
        for (int b = 0; b + 1 < d.N; ++b)
            if (!(sp.bond[static_cast<size_t>(b)] > 0.0)) {
                delete e;
                return fail(FTS_INVALID, "sequence bond length");
            }
        for (int t : sp.type)
            if (t < 0 || t >= e->nT) {
                delete e;
                return fail(FTS_INVALID, "sequence monomer type");
            }
        sp.bond_offset = e->total_bonds;
        e->total_bonds += static_cast<size_t>(std::max(0, d.N - 1));
        e->Nmax = std::max(e->Nmax, d.N);
        e->total_segments += sp.n * d.N;
        for (int t : sp.type) e->total_type_segments[static_cast<size_t>(t)] += sp.n;
        e->species.push_back(sp);
    }
    e->rho0 = static_cast<double>(e->total_segments) / e->V;
    // mobilities (param.cu:390-410)
    e->lambda_omega.assign(static_cast<size_t>(e->n_omega), 0.0);
    if (s->n_lambda_omega == 1 && s->lambda_omega) e->lambda_omega.assign(static_cast<size_t>(e->n_omega), s->lambda_omega[0]);
    else if (s->n_lambda_omega == e->n_omega && s->lambda_omega) e->lambda_omega.assign(s->lambda_omega, s->lambda_omega + e->n_omega);
    else if (s->n_lambda_omega != 0) {
        delete e;
        return fail(FTS_INVALID, "lambda_omega length must be 1 or the number of active interaction modes");
    }
    for (double x : e->lambda_omega)
        if (!std::isfinite(x) || x < 0.0) {
            delete e;
            return fail(FTS_INVALID, "lambda_omega");
        }
    rc = e->build_ions(s);
    if (rc != FTS_OK) {
        delete e;
        return rc;
    }
    e->fft = new oracle::Fft3d(e->m1, e->m2, e->m3);
    e->build_kernels();
    if (e->learned_scheme() && e->n_omega > 1) {
        delete e;
        return fail(FTS_INVALID, "the learned kernel covers one interaction mode and psi (n_omega <= 1)");
    }
    if (e->si_scheme()) {
        e->build_si_kernels();
        e->lrk_build_bins();
    }
    const size_t M = static_cast<size_t>(e->M);
    const size_t nO = static_cast<size_t>(e->n_omega), nT = static_cast<size_t>(e->nT), nI = e->ions.size(), Nm = static_cast<size_t>(e->Nmax);
    const Cx zero{0.0, 0.0};
    e->omega.assign(nO * M, zero);
    e->omega_bar.assign(nO * M, zero);
    e->omega_s.assign(nO * M, zero);
    e->drift.assign(nO * M, zero);
    e->drift_bar.assign(nO * M, zero);
    e->eta_omega.assign(nO * M, 0.0);
    e->etaPsi.assign(M, 0.0);
    for (Field* f : {&e->psi, &e->psi_bar, &e->psi_s, &e->dpsi, &e->dpsi_bar, &e->rho_c, &e->rho_c_poly, &e->rho_c_ion,
                     &e->rho_c_old, &e->rho_p, &e->ion_psi, &e->ion_boltzmann, &e->ion_charge, &e->obs_a, &e->obs_b,
                     &e->sa, &e->sb, &e->sc, &e->dfield, &e->dpsi_s})
        f->assign(M, zero);
    e->W_type_s.assign(nT * M, zero);
    e->rho_type.assign(nT * M, zero);
    for (Field* f : {&e->W, &e->exp_mW, &e->exp_W, &e->qF, &e->qB}) f->assign(Nm * M, zero);
    e->ion_rho.assign(std::max<size_t>(1, nI) * M, zero);
    *out = e;
    return FTS_OK;
}

void fto_destroy(fts_engine* h) { delete h; }

int fto_get_info(fts_engine* h, fts_info* o)
{
    if (!h || !o) return fail(FTS_INVALID, "null argument");
    std::memset(o, 0, sizeof(*o));
    o->grid_size = h->M;
    o->step = h->step_counter;
    o->restarts = h->restarts;
    o->noise_epoch = h->noise_epoch;
    o->checkpoint_step = h->ckpt_step;
    o->failed_step = h->last_failed_step;
    o->n_omega = h->n_omega;
    o->monomer_types = h->nT;
    o->n_species = static_cast<int>(h->species.size());
    o->n_ions = static_cast<int>(h->ions.size());
    o->precision = FTS_PREC_FP64;
    o->scheme = h->scheme;
    o->volume = h->V;
    o->rho0 = h->rho0;
    {
        long double wd = 0.0L;
        for (int al = 0; al < h->nT; ++al) wd += static_cast<long double>(h->total_type_segments[static_cast<size_t>(al)]) * h->K[static_cast<size_t>(al) * h->nT + al];
        const long double md = wd / static_cast<long double>(h->total_segments);
        o->short_range_self_pressure = static_cast<double>(md / (16.0L * static_cast<long double>(Pi) * std::sqrt(static_cast<long double>(Pi)) * h->a * h->a * h->a));
    }
    for (int l = 0; l < h->n_omega; ++l) {
        o->mode_eigenvalue[l] = h->eig[static_cast<size_t>(h->active[static_cast<size_t>(l)])];
        o->omega_coefficient[l] = h->omega_coefficient(l);
        o->gamma_is_imag[l] = h->gamma(l).y != 0.0 ? 1.0 : 0.0;
        for (int al = 0; al < h->nT; ++al) o->mode_O[l][al] = h->Ocoef(l, al);
    }
    for (size_t j = 0; j < h->ions.size(); ++j) {
        o->ion_valence[j] = h->ions[j].z;
        o->ion_count[j] = h->ions[j].n;
        o->ion_smearing[j] = h->ions[j].a;
    }
    return FTS_OK;
}

int fto_set_fields(fts_engine* h, const double* om, const double* ps)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    const size_t M = static_cast<size_t>(h->M);
    if (om) std::memcpy(h->omega.data(), om, sizeof(Cx) * static_cast<size_t>(h->n_omega) * M);
    else std::fill(h->omega.begin(), h->omega.end(), Cx{0, 0});
    if (ps) std::memcpy(h->psi.data(), ps, sizeof(Cx) * M);
    else std::fill(h->psi.begin(), h->psi.end(), Cx{0, 0});
    h->remove_psi_zero_mode(h->psi.data());
    h->ckpt_step = -1;
    h->lrk_have_prev = false;
    return FTS_OK;
}

int fto_get_field(fts_engine* h, int id, double* out, size_t n_doubles)
{
    if (!h || !out) return fail(FTS_INVALID, "null argument");
    const Field* f = nullptr;
    size_t count = 0;
    const size_t M = static_cast<size_t>(h->M);
    switch (id) {
    case FTS_FIELD_OMEGA: f = &h->omega; count = static_cast<size_t>(h->n_omega) * M; break;
    case FTS_FIELD_PSI: f = &h->psi; count = M; break;
    case FTS_FIELD_RHO_TYPE: f = &h->rho_type; count = static_cast<size_t>(h->nT) * M; break;
    case FTS_FIELD_RHO_P: f = &h->rho_p; count = M; break;
    case FTS_FIELD_RHO_C: f = &h->rho_c; count = M; break;
    case FTS_FIELD_RHO_C_POLY: f = &h->rho_c_poly; count = M; break;
    case FTS_FIELD_RHO_C_ION: f = &h->rho_c_ion; count = M; break;
    case FTS_FIELD_ION_RHO: f = &h->ion_rho; count = h->ions.size() * M; break;
    case FTS_FIELD_QF: f = &h->qF; count = static_cast<size_t>(h->species[static_cast<size_t>(h->last_species)].N) * M; break;
    case FTS_FIELD_QB: f = &h->qB; count = static_cast<size_t>(h->species[static_cast<size_t>(h->last_species)].N) * M; break;
    default: return fail(FTS_INVALID, "field id");
    }
    if (n_doubles != 2 * count) return fail(FTS_INVALID, "field size mismatch");
    std::memcpy(out, f->data(), sizeof(double) * n_doubles);
    return FTS_OK;
}

int fto_compute_densities(fts_engine* h, double* Q_out)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    const std::vector<Cx> Q = h->compute_densities(h->omega.data(), h->psi.data());
    bool ok = true;
    for (size_t p = 0; p < Q.size(); ++p) {
        if (Q_out) {
            Q_out[2 * p] = Q[p].x;
            Q_out[2 * p + 1] = Q[p].y;
        }
        ok = ok && finite(Q[p]);
    }
    return ok ? FTS_OK : fail(FTS_NONFINITE, "non-finite Q");
}

int fto_measure(fts_engine* h, fts_observables* out)
{
    if (!h || !out) return fail(FTS_INVALID, "null argument");
    const std::vector<Cx> Q = h->compute_densities(h->omega.data(), h->psi.data());
    h->measure(Q, h->omega.data(), h->psi.data(), out);
    for (const Cx& q : Q)
        if (!finite(q)) return fail(FTS_NONFINITE, "non-finite Q");
    return FTS_OK;
}

int fto_step(fts_engine* h, int n, int record_every, fts_observables* records, int capacity, int* n_records)
{
    if (!h || n < 0) return fail(FTS_INVALID, "null engine or negative step count");
    int nrec = 0;
    if (n_records) *n_records = 0;
    const long long target = h->step_counter + n;
    while (h->step_counter < target) {
        // checkpoint at the first step of the call and at multiples of ckpt_every
        if (h->ckpt_every > 0 && (h->ckpt_step < 0 || (h->step_counter % h->ckpt_every == 0 && h->step_counter != h->ckpt_step))) {
            h->ckpt_omega = h->omega;
            h->ckpt_psi = h->psi;
            h->ckpt_step = h->step_counter;
        }
        const long long step = h->step_counter + 1; // 1-based step label (complexLangevin.cu:206)
        const bool record = record_every > 0 && step % record_every == 0 && records && nrec < capacity;
        fts_observables* rec = record ? &records[nrec] : nullptr;
        int rc;
        if (h->scheme == FTS_SCHEME_PC_REFERENCE) rc = h->step_pc(record, rec);
        else if (h->scheme == FTS_SCHEME_SIPC || h->scheme == FTS_SCHEME_SIPC_LEARNED) rc = h->step_sipc(record, rec);
        else rc = h->step_em(record, rec, h->scheme == FTS_SCHEME_SEMI_IMPLICIT || h->scheme == FTS_SCHEME_SI_LEARNED);
        if (record) {
            rec->step = step;
            ++nrec;
        }
        if (rc == FTS_OK && h->drift_stats) {
            h->drift_rows.push_back(static_cast<double>(step));
            h->drift_rows.insert(h->drift_rows.end(), h->drift_cur.begin(), h->drift_cur.end());
        }
        if (rc == FTS_NONFINITE && h->ckpt_every > 0) {
            if (h->ckpt_step < 0) return rc;
            const size_t w = static_cast<size_t>(h->n_omega) + 2;
            while (!h->drift_rows.empty() && h->drift_rows[h->drift_rows.size() - w] > static_cast<double>(h->ckpt_step))
                h->drift_rows.resize(h->drift_rows.size() - w);
            if (h->restarts >= h->max_restarts) return fail(FTS_NONFINITE, "maximum non-finite restarts exceeded");
            h->omega = h->ckpt_omega;
            h->psi = h->ckpt_psi;
            ++h->restarts;
            ++h->noise_epoch;
            h->last_failed_step = step;
            h->seed = h->seed0 + static_cast<unsigned long long>(h->noise_epoch) * 0x9E3779B97F4A7C15ull;
            h->step_counter = h->ckpt_step;
            h->lrk_have_prev = false;
            while (nrec > 0 && records[nrec - 1].step > h->ckpt_step) --nrec;
            rc = FTS_OK;
        }
        if (n_records) *n_records = nrec;
        if (rc != FTS_OK) return rc;
    }
    return FTS_OK;
}

int fto_set_restart(fts_engine* h, int checkpoint_every, int max_restarts)
{
    if (!h || checkpoint_every < 0 || max_restarts < 0) return fail(FTS_INVALID, "bad restart settings");
    h->ckpt_every = checkpoint_every;
    h->max_restarts = max_restarts;
    h->restarts = 0;
    h->ckpt_step = -1;
    return FTS_OK;
}

// Binary state file, the engine's format: "FTSSTAT1", int32 m1 m2 m3 n_omega, int64 step,
// uint64 seed0, int32 noise_epoch restarts, omega [n_omega][M] and psi [M] (interleaved doubles)
int fto_save_state(fts_engine* h, const char* path)
{
    if (!h || !path) return fail(FTS_INVALID, "null argument");
    const std::string tmp = std::string(path) + ".new";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return fail(FTS_INVALID, "cannot open " + tmp);
    const int32_t hdr[4] = {h->m1, h->m2, h->m3, h->n_omega};
    const int64_t step = h->step_counter;
    const uint64_t seed0 = h->seed0;
    const int32_t epoch[2] = {h->noise_epoch, h->restarts};
    const size_t no = static_cast<size_t>(h->n_omega) * static_cast<size_t>(h->M);
    bool ok = std::fwrite("FTSSTAT1", 1, 8, fp) == 8 && std::fwrite(hdr, sizeof(hdr), 1, fp) == 1 && std::fwrite(&step, sizeof(step), 1, fp) == 1 &&
              std::fwrite(&seed0, sizeof(seed0), 1, fp) == 1 && std::fwrite(epoch, sizeof(epoch), 1, fp) == 1 &&
              (no == 0 || std::fwrite(h->omega.data(), sizeof(Cx), no, fp) == no) &&
              std::fwrite(h->psi.data(), sizeof(Cx), static_cast<size_t>(h->M), fp) == static_cast<size_t>(h->M);
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) return fail(FTS_INVALID, std::string("cannot write ") + path);
    return FTS_OK;
}

int fto_load_state(fts_engine* h, const char* path)
{
    if (!h || !path) return fail(FTS_INVALID, "null argument");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(FTS_INVALID, std::string("cannot open ") + path);
    char magic[8];
    int32_t hdr[4];
    int64_t step = 0;
    uint64_t seed0 = 0;
    int32_t epoch[2];
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, "FTSSTAT1", 8) == 0 && std::fread(hdr, sizeof(hdr), 1, fp) == 1 &&
              std::fread(&step, sizeof(step), 1, fp) == 1 && std::fread(&seed0, sizeof(seed0), 1, fp) == 1 && std::fread(epoch, sizeof(epoch), 1, fp) == 1;
    if (!ok || hdr[0] != h->m1 || hdr[1] != h->m2 || hdr[2] != h->m3 || hdr[3] != h->n_omega || step < 0 || epoch[0] < 0 || epoch[1] < 0) {
        std::fclose(fp);
        return fail(FTS_INVALID, std::string(path) + ": not a state file of this mesh and mode count");
    }
    const size_t no = static_cast<size_t>(h->n_omega) * static_cast<size_t>(h->M);
    std::vector<Cx> w(no), p(static_cast<size_t>(h->M));
    ok = (no == 0 || std::fread(w.data(), sizeof(Cx), no, fp) == no) && std::fread(p.data(), sizeof(Cx), p.size(), fp) == p.size();
    std::fclose(fp);
    if (!ok) return fail(FTS_INVALID, std::string(path) + ": truncated state file");
    std::copy(w.begin(), w.end(), h->omega.begin());
    h->psi = p;
    h->step_counter = step;
    h->seed0 = seed0;
    h->noise_epoch = epoch[0];
    h->restarts = epoch[1];
    h->seed = seed0 + static_cast<unsigned long long>(epoch[0]) * 0x9E3779B97F4A7C15ull;
    h->ckpt_step = -1;
    h->lrk_have_prev = false;
    return FTS_OK;
}

int fto_set_drift_stats(fts_engine* h, int enabled)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    h->drift_stats = enabled != 0;
    h->drift_rows.clear();
    return FTS_OK;
}

int fto_get_drift_rows(fts_engine* h, double* out, int capacity, int* n_rows)
{
    if (!h || !n_rows || capacity < 0 || (capacity > 0 && !out)) return fail(FTS_INVALID, "bad argument");
    const size_t w = static_cast<size_t>(h->n_omega) + 2;
    const size_t n = std::min(h->drift_rows.size() / w, static_cast<size_t>(capacity));
    if (n) std::memcpy(out, h->drift_rows.data(), sizeof(double) * n * w);
    h->drift_rows.erase(h->drift_rows.begin(), h->drift_rows.begin() + static_cast<std::ptrdiff_t>(n * w));
    *n_rows = static_cast<int>(n);
    return FTS_OK;
}

int fto_inject_nonfinite(fts_engine* h, long long step)
{
    if (!h) return fail(FTS_INVALID, "null engine");
    h->inject_step = step;
    return FTS_OK;
}

int fto_set_step(fts_engine* h, long long step)
{
    if (!h || step < 0) return fail(FTS_INVALID, "bad step");
    h->step_counter = step;
    h->ckpt_step = -1;
    return FTS_OK;
}

int fto_fft3d(fts_engine* h, const double* in, double* out, int sign)
{
    if (!h || !in || !out || (sign != 1 && sign != -1)) return fail(FTS_INVALID, "bad argument");
    std::memcpy(out, in, sizeof(Cx) * static_cast<size_t>(h->M));
    h->fft->run(reinterpret_cast<Cx*>(out), sign);
    return FTS_OK;
}

int fto_bead_step(fts_engine* h, const double* q, const double* emw, double bond, double* q_out)
{
    if (!h || !q || !emw || !q_out || !(bond > 0.0)) return fail(FTS_INVALID, "bad argument");
    std::vector<double> phi(static_cast<size_t>(h->M));
    for (int i = 0; i < h->M; ++i) phi[static_cast<size_t>(i)] = std::exp(-bond * bond * h->k2[static_cast<size_t>(i)] / 6.0);
    Cx* o = reinterpret_cast<Cx*>(q_out);
    h->apply_kernel(reinterpret_cast<const Cx*>(q), o, phi.data());
    h->mul(o, reinterpret_cast<const Cx*>(emw));
    return FTS_OK;
}

// ---- learned linear-response kernel ------------------------------------------
int fto_lrk_set_network(fts_engine* h, const float* params, double threshold)
{
    if (!h || !params) return fail(FTS_INVALID, "null argument");
    for (int j = 0; j < FTS_LRK_NPARAMS; ++j) h->lrk_P[static_cast<size_t>(j)] = params[j];
    h->lrk_threshold = threshold;
    return FTS_OK;
}

int fto_lrk_set_recording(fts_engine* h, int capacity)
{
    if (!h || capacity < 0) return fail(FTS_INVALID, "bad argument");
    if (capacity > 0 && !h->si_scheme())
        return fail(FTS_INVALID, "training records need a semi-implicit scheme");
    if (capacity > 0 && (h->n_omega > 1 || (h->n_omega == 0 && !h->electrostatics())))
        return fail(FTS_INVALID, "training records need n_omega <= 1 and at least one field");
    h->lrk_capacity = capacity;
    h->lrk_ring.clear();
    h->lrk_st.records = 0;
    return FTS_OK;
}

int fto_lrk_get_records(fts_engine* h, double* out, int max_records, int* n_out)
{
    if (!h || !n_out || max_records < 0 || (max_records > 0 && !out)) return fail(FTS_INVALID, "bad argument");
    const int n = std::min(max_records, static_cast<int>(h->lrk_ring.size()));
    for (int k = 0; k < n; ++k) std::memcpy(out + static_cast<size_t>(k) * FTS_LRK_RECORD, h->lrk_ring[static_cast<size_t>(k)].data(), sizeof(double) * FTS_LRK_RECORD);
    h->lrk_ring.clear();
    *n_out = n;
    return FTS_OK;
}

int fto_lrk_get_stats(fts_engine* h, fts_lrk_stats* out)
{
    if (!h || !out) return fail(FTS_INVALID, "null argument");
    *out = h->lrk_st;
    return FTS_OK;
}

int fto_lrk_forward64(const double* params, const double* features, int n, double* r_out)
{
    if (!params || n < 0 || (n > 0 && (!features || !r_out))) return fail(FTS_INVALID, "bad argument");
    for (int k = 0; k < n; ++k) {
        oracle::LrkActs a;
        oracle::lrk_forward(params, features + static_cast<size_t>(k) * FTS_LRK_FEATURES, a);
        for (int c = 0; c < FTS_LRK_OUT; ++c) r_out[static_cast<size_t>(k) * FTS_LRK_OUT + c] = a.r[c];
    }
    return FTS_OK;
}

int fto_lrk_forward(const float* params, const float* features, int n, float* r_out)
{
    if (!params || n < 0 || (n > 0 && (!features || !r_out))) return fail(FTS_INVALID, "bad argument");
    std::vector<double> P(params, params + FTS_LRK_NPARAMS), f(features, features + static_cast<size_t>(n) * FTS_LRK_FEATURES),
        r(static_cast<size_t>(n) * FTS_LRK_OUT);
    fto_lrk_forward64(P.data(), f.data(), n, r.data());
    for (size_t j = 0; j < r.size(); ++j) r_out[j] = static_cast<float>(r[j]);
    return FTS_OK;
}

int fto_lrk_loss_grad64(const double* params, const double* records, int n_records, double* loss, double* grad)
{
    if (!params || !loss || n_records < 0 || (n_records > 0 && !records)) return fail(FTS_INVALID, "bad argument");
    *loss = oracle::lrk_loss_grad(params, records, n_records, grad);
    return FTS_OK;
}

int fto_lrk_loss_grad(const float* params, const double* records, int n_records, double* loss, float* grad)
{
    if (!params || !loss || n_records < 0 || (n_records > 0 && !records)) return fail(FTS_INVALID, "bad argument");
    std::vector<double> P(params, params + FTS_LRK_NPARAMS), G(FTS_LRK_NPARAMS, 0.0);
    *loss = oracle::lrk_loss_grad(P.data(), records, n_records, grad ? G.data() : nullptr);
    if (grad)
        for (int j = 0; j < FTS_LRK_NPARAMS; ++j) grad[j] = static_cast<float>(G[static_cast<size_t>(j)]);
    return FTS_OK;
}

// Adam on minibatches, as fts_lrk_trainer_run, in FP64: the same per-epoch shuffle
// (std::mt19937_64 + std::shuffle of the record order), minibatch loss = mean over its
// records, learning rate lr0 (lr1/lr0)^(step/(total-1)), beta 0.9/0.999, eps 1e-8.
struct fts_lrk_trainer {
    std::vector<double> records;
    int n = 0;
};

int fto_lrk_trainer_create(const double* records, int n_records, fts_lrk_trainer** out)
{
    if (!records || n_records <= 0 || !out) return fail(FTS_INVALID, "lrk_trainer_create: bad argument");
    fts_lrk_trainer* t = new fts_lrk_trainer();
    t->records.assign(records, records + static_cast<size_t>(n_records) * FTS_LRK_RECORD);
    t->n = n_records;
    *out = t;
    return FTS_OK;
}

void fto_lrk_trainer_destroy(fts_lrk_trainer* t) { delete t; }

int fto_lrk_trainer_run(fts_lrk_trainer* t, float* params, int epochs, int batch_records, double lr0, double lr1,
                        unsigned long long seed, double* loss_history)
{
    if (!t || !params || epochs < 0 || batch_records <= 0 || !(lr0 > 0.0) || !(lr1 > 0.0)) return fail(FTS_INVALID, "lrk_trainer_run: bad argument");
    const int B = std::min(batch_records, t->n);
    const int per_epoch = (t->n + B - 1) / B;
    const long long total = static_cast<long long>(per_epoch) * epochs;
    std::vector<double> P(params, params + FTS_LRK_NPARAMS), m1(FTS_LRK_NPARAMS, 0.0), m2(FTS_LRK_NPARAMS, 0.0), G(FTS_LRK_NPARAMS);
    std::vector<double> batch;
    std::mt19937_64 rng(seed);
    std::vector<int> order(static_cast<size_t>(t->n));
    std::iota(order.begin(), order.end(), 0);
    long long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        double el = 0.0;
        for (int b0 = 0; b0 < t->n; b0 += B) {
            const int nb = std::min(B, t->n - b0);
            batch.resize(static_cast<size_t>(nb) * FTS_LRK_RECORD);
            for (int k = 0; k < nb; ++k)
                std::memcpy(batch.data() + static_cast<size_t>(k) * FTS_LRK_RECORD,
                            t->records.data() + static_cast<size_t>(order[static_cast<size_t>(b0 + k)]) * FTS_LRK_RECORD, sizeof(double) * FTS_LRK_RECORD);
            el += oracle::lrk_loss_grad(P.data(), batch.data(), nb, G.data()) / per_epoch;
            ++step;
            const double frac = total > 1 ? static_cast<double>(step - 1) / static_cast<double>(total - 1) : 0.0;
            const double lr = lr0 * std::pow(lr1 / lr0, frac);
            const double bc1 = 1.0 - std::pow(0.9, static_cast<double>(step)), bc2 = 1.0 - std::pow(0.999, static_cast<double>(step));
            for (int j = 16; j < FTS_LRK_NPARAMS; ++j) {
                m1[static_cast<size_t>(j)] = 0.9 * m1[static_cast<size_t>(j)] + 0.1 * G[static_cast<size_t>(j)];
                m2[static_cast<size_t>(j)] = 0.999 * m2[static_cast<size_t>(j)] + 0.001 * G[static_cast<size_t>(j)] * G[static_cast<size_t>(j)];
                P[static_cast<size_t>(j)] -= lr * (m1[static_cast<size_t>(j)] / bc1) / (std::sqrt(m2[static_cast<size_t>(j)] / bc2) + 1e-8);
            }
        }
        if (loss_history) loss_history[ep] = el;
    }
    for (int j = 0; j < FTS_LRK_NPARAMS; ++j) params[j] = static_cast<float>(P[static_cast<size_t>(j)]);
    return FTS_OK;
}

} // extern "C"

// ---- test exports: the noise generator --------------------------------------
extern "C" {

void fto_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out)
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    uint32_t o[4];
    oracle::philox4x32_10(c, k, o);
    for (int i = 0; i < 4; ++i) out[i] = o[i];
}

double fto_cl_normal(unsigned long long seed, unsigned long long step, unsigned field, unsigned site)
{
    return oracle::cl_normal(seed, step, field, site);
}

void fto_cl_normal_pair(unsigned long long seed, unsigned long long step, unsigned field, unsigned site, double* out)
{
    oracle::cl_normal_pair(seed, step, field, site, out);
}

} // extern "C"
