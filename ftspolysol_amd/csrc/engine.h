// engine.h -- the host engine object shared by the single-system C ABI
// (engine.hip) and the replica batch (batch.hip).
//
// One fts_engine owns every device buffer of one system (the role of
// deviceState, deviceState.cu:8-105) and runs on one HIP stream of default
// priority, its own or one a replica batch shares between its replicas.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fts.h"
#include "fft.h"
#include "kernel_params.h"
#include "lrk.h"

using namespace fts;

namespace fts_host {

constexpr double Pi = 3.141592653589793238462643383279502884;
inline thread_local std::string g_error;

inline int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

#define HIP_TRY(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return fail(FTS_DEVICE_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int ew_blocks(size_t n) { return static_cast<int>((n + EW_THREADS - 1) / EW_THREADS); }

template <typename X> struct DevBuf {
    X* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc(reinterpret_cast<void**>(&p), sizeof(X) * count);
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct HostSpecies {
    long long n = 0;
    int N = 0;
    std::vector<int> type;
    std::vector<double> charge, bond;
    std::vector<int> bond_cls;   // sep-table class of each bond
    // bead classes: distinct (type, charge)
    std::vector<int> cls_of_bead, cls_type, cls_start, cls_beads;
    std::vector<double> cls_z;
    DevBuf<int> d_cls_type, d_cls_start, d_cls_beads;
    DevBuf<double> d_cls_z, d_cls_shift;
    std::vector<double> cls_shift; // host copy of d_cls_shift (min Re W_c + BOLTZ_CAP per class, FP32 only)
    // mirrored sweep (palindromic chain): qB_s = qF_{N-1-s}; half_* list, per class,
    // the beads s <= N-1-s the density contraction visits
    bool mirror = false;
    std::vector<int> half_start, half_beads;
    DevBuf<int> d_half_start, d_half_beads;
    double charge_square_sum() const
    {
        double t = 0;
        for (double z : charge) t += z * z;
        return t;
    }
    double net_charge() const
    {
        double t = 0;
        for (double z : charge) t += z;
        return t;
    }
};

struct HostIon {
    double z = 0, a = 0;
    long long n = 0;
};

// scalar slots in the device scalar buffer
enum {
    SC_Q = 0,        // [8] polymer partition functions (of the stored, possibly shifted, propagators)
    SC_QION = 8,     // [8] ion partition functions
    Q_SLOTS = 16,    // SC_Q and SC_QION: the partition functions a step checks
    SC_PSI_MEAN = 16,
    SC_ETA_MEAN = 17,
    SC_DRIFT_MEAN = 18,
    SC_TMP = 19,
    SC_CONTRACT = 24, // [7]
    SC_LF = 32,       // [8] per species: log-gain LF(N-1) of the renormalised forward sweep
    SC_H = 40,        // [8] per interaction mode: (c_l / 2) <omega_l^2>
    SC_PSI_S = 48,    // [7] k-space contraction of psi with itself
    SC_LRK = 56,      // [4] (mean, mean square) of Re w0, Im w0, Re psi, Im psi (learned kernel's moments)
    SC_DRIFT = 60,    // [FTS_MAX_TYPES + 1] drift statistics of the last step (.x): modes, then psi
    SC_COUNT = 60 + FTS_MAX_TYPES + 1
};

} // namespace fts_host

using namespace fts_host;

struct fts_engine {
    // ---- parameters ---------------------------------------------------------
    int m1 = 0, m2 = 0, m3 = 0, M = 0;
    double L1 = 0, L2 = 0, L3 = 0, V = 0, dV = 0, a = 0, lB = 0, dt = 0, lambda_psi = 0;
    bool noise = true;
    int scheme = 0, precision = 0, stress_mask = 0;
    unsigned long long seed = 0;
    long long step_counter = 0;
    int nT = 0, n_omega = 0;
    double xi = 0, rho0 = 0, P_sr = 0;
    std::vector<double> K, eig, O;
    std::vector<int> active;
    std::vector<double> lambda_omega;
    long long total_segments = 0;
    std::vector<long long> total_type_segments;
    std::vector<HostSpecies> species;
    std::vector<HostIon> ions;
    int Nmax = 0, max_cls = 0, last_species = 0;
    double mode_term[6] = {0, 0, 0, 0, 0, 0};
    ModeCoefs mc{};
    std::vector<double2> mode_coef_host; // gamma O [nO][nT]

    // ---- device ---------------------------------------------------------------
    hipStream_t stream = nullptr;
    int device = 0;                  // the GPU current at creation
    void* noalloc_scratch = nullptr; // this GPU's address of the no-alloc scratch array (noalloc.hip)
    bool own_stream = true; // false: the stream belongs to a replica batch
    FftGeom g64{}, g32{};
    DevBuf<double> kvec, sep_gamma, sep_bond, sep_ion, si_table_omega, si_table_psi;
    DevBuf<double2> tw64_x, tw64_yz, tw64_z;
    DevBuf<float2> tw32_x, tw32_yz, tw32_z;
    int n_bond_cls = 0;
    std::vector<double> bond_cls_len;
    DevBuf<double2> omega, omega_bar, omega_s, drift, psi, psi_bar, psi_s, dpsi, dpsi_bar;
    DevBuf<double2> W_type_s, rho_hat, rho_type, rho_c_hat, rho_c, rho_c_poly, rho_c_ion, rho_p;
    DevBuf<double2> ion_psi, ion_b, ion_rho, t1, t2, t3, mode_coef;
    // spectrum of the psi field the last densities_begin smoothed (plane schedule,
    // electrostatics): the ion kernels, the psi drift, the learned kernel and the
    // measurement start from it instead of transforming psi again.  psi_hat_of names
    // the field it belongs to and is cleared wherever psi changes.
    DevBuf<double2> psi_hat;
    const double2* psi_hat_of = nullptr;
    bool psi_hat_valid(const double2* ps) const { return psi_hat_of != nullptr && psi_hat_of == ps; }
    DevBuf<double2> partials, scal;
    DevBuf<int> flag;
    // propagators (double or float complex)
    void* qF = nullptr;
    void* qB = nullptr;
    void* E = nullptr;
    void* tmpF = nullptr;
    void* tmpB = nullptr;
    void* kA = nullptr;
    void* kB = nullptr;
    std::vector<double> q_shift_total; // per species: sum of class shifts over the chain
    DevBuf<double> norms, logc, rfac, rfac2; // FP32 per-bead renormalisation (see k_scale_prefix)
    std::vector<double> q_log_gain;          // per species: LF(N-1) of the last solve (host copy)

    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> event_pool;
    fts_timing tstat{};

    ~fts_engine() { release(); }
    void release();

    size_t csize() const { return precision == FTS_PREC_FP32 ? sizeof(float2) : sizeof(double2); }
    double2* f(DevBuf<double2>& b, size_t k) { return b.p + k * static_cast<size_t>(M); }
    bool electro() const { return lB > 0.0; }
    double omega_coefficient(int l) const { return rho0 * (1.0 / std::fabs(eig[static_cast<size_t>(active[static_cast<size_t>(l)])])); }
    bool gamma_imag(int l) const { return eig[static_cast<size_t>(active[static_cast<size_t>(l)])] > 0.0; }
    double Ocoef(int l, int al) const { return O[static_cast<size_t>(active[static_cast<size_t>(l)]) * nT + al]; }

    // setup
    int build_modes(const fts_system* s);
    int build_ions(const fts_system* s);
    int setup_device();
    int build_si_tables();

    // primitives
    KOp kop_sep(const double* sep, double c0) const { return KOp{sep, nullptr, c0, 0.0, 0.0, 0, 0}; }
    int fft_apply(const double2* src, double2* dst, const KOp& op);
    int fft_apply_hat(const double2* hat, double2* dst, const KOp& op); // dst = IFFT[K hat] (plane schedule)
    int fft_forward(const double2* src, double2* dst);
    // Independent FFT jobs of one replica, two to a launch on the plane schedule (twice the
    // workgroups per pass, a third of the launches); the same arithmetic per job as the
    // single forms, which the line schedule runs one job at a time.
    struct FftJob {
        const double2* src; // field (apply, forward) or spectrum (apply_hat)
        double2* dst;
        KOp op;
    };
    int fft_apply_n(const FftJob* jobs, int n);
    // An FP32 engine on the plane passes transforms its FP64 fields in complex FP32
    // (the passes read and write the FP64 fields, the spectra in between live in
    // tmpF / tmpB, free outside the sweeps); an FP64 engine in FP64.
    bool fields_fp32() const { return precision == FTS_PREC_FP32 && plane_ok; }
    int fft32_n(int kind, const FftJob* jobs, int n); // kind: 0 apply, 1 apply from a spectrum, 2 forward
    int fft_apply_hat_n(const FftJob* jobs, int n);
    int fft_forward_n(const FftJob* jobs, int n);
    template <typename T> int fft_forward_t(const V2<T>* src, V2<T>* dst, const V2<T>* src2 = nullptr, V2<T>* dst2 = nullptr);
    int reduce_finalize(int K, int slot, double scale);
    int sum_field(const double2* a, int slot, double scale);
    int remove_mean(double2* a);
    int read_scal(int slot, int count, double2* out);
    int read_log_gain();

    // hot path
    template <typename T> int solve_species(int p, bool accumulate = true);
    template <typename T> int run_sweeps(int p, int schedule, int nsteps, bool timed);
    int autotune_schedule();
    template <typename T> int bead_hook(V2<T>* q, const V2<T>* emw, V2<T>* qout, const KOp& op, hipStream_t s);
    int sched = FTS_SCHED_CONV;
    bool plane_ok = true;
    size_t norm_stride = 0;
    double tune_ms[2] = {0.0, 0.0};
    int compute_densities(const double2* om, const double2* ps);
    // the phases of compute_densities, so that a replica batch can run the
    // propagator sweeps of all its replicas in shared launches:
    //   densities_begin; per species p: solve_begin(p), sweeps, solve_end(p); densities_end
    int densities_begin(const double2* om, const double2* ps);
    template <typename T> int solve_begin(int p);
    template <typename T> int solve_end(int p, bool accumulate);
    int densities_end();
    const double2* dens_psi = nullptr; // the psi field densities_begin was given
    int check_Q(const char* what);
    bool partition_finite(const double2* Q) const; // Q: Q_SLOTS values from SC_Q
    int finite_queue(int* host_out);      // queue the field finiteness check; *host_out valid after a stream sync
    int q_queue(double2* host_out);       // queue the copy of the Q_SLOTS partition functions; valid after a stream sync
    int psi_drift(const double2* ps, const double2* rho_src, double2* out);
    int check_finite(bool& ok);
    // measure = measure_queue (device work and copies into pinned memory), a stream
    // sync, measure_finish (host arithmetic).  With defer_measure set, measure()
    // only queues and leaves the sync and measure_finish to the caller, so that a
    // batch can queue the work of all its replicas before it waits.
    int measure(const double2* om, const double2* ps, fts_observables* rec);
    int measure_queue(const double2* om, const double2* ps, fts_observables* rec);
    int measure_finish(fts_observables* rec);
    void measure_header(fts_observables* rec) const;
    bool defer_measure = false;
    template <typename T> int stress_chain(int p);
    int contract_to(const void* A, const void* B, bool fp32, const double* sep, double2* dst);
    // measure(): device results gathered in stress_acc and copied, with the scalar
    // slots, the FP32 shifts and rfac2, into pinned host memory behind one sync
    DevBuf<double2> stress_acc;
    size_t stress_chain_base(int p) const; // first stress_acc row of species p's chain
    void* pin = nullptr;                   // pinned staging
    // the noise key {seed, step} of the step being taken: kernels that draw CL noise read
    // it from nkey, which key_queue refreshes from the pinned hkey on the stream (a copy a
    // captured step graph repeats on every replay, reading hkey as the host left it)
    DevBuf<unsigned long long> nkey;
    unsigned long long* hkey = nullptr;
    int key_queue();
    // fill kernels for the clears of the step (no memset nodes in captured graphs)
    int zero_async(double2* p, size_t n);
    // bumped by every call that changes values a captured step graph holds (the learned
    // kernel's threshold and ring, drift statistics, the test hook): the batch drops its
    // step graphs of this replica when it changes
    unsigned long long config_epoch = 0;
    size_t pin_bytes = 0;
    int step_pc(bool record, fts_observables* rec);
    int step_em(bool record, fts_observables* rec, bool si);
    int step_sipc(bool record, fts_observables* rec);
    // the parts of a CL step between its density computations
    int pc_predict();                                          // after the densities at (w, psi)
    int pc_correct(bool record, fts_observables* rec);         // after the densities at (w_bar, psi_bar)
    int em_update(bool record, fts_observables* rec, bool si); // after the densities at (w, psi)
    int sipc_predict();                                        // after the densities at (w, psi)
    int sipc_correct(bool record, fts_observables* rec);       // after the densities at (w_bar, psi_bar)
    // the semi-implicit predictor-corrector's per-step fields: the predictor's increments
    // w_bar - w (omega modes, then psi), kept for the corrector's (1 - P2) share
    DevBuf<double2> sipc_d, sipc_dpsi;
    int finish_step();
    // drift statistics (complexLangevin.cu:318-324, 466-497): with drift_stats set, every
    // step queues max over sites of |lam_l/2 (u_l + u_bar_l)| per mode and |lam_psi/2
    // (u_psi + u_psi_bar)| (predictor-corrector; EM and semi-implicit: |lam u| at the state
    // the step starts from) into SC_DRIFT; fts_step appends [step, values] to drift_rows
    bool drift_stats = false;
    std::vector<double> drift_rows;
    int drift_queue(bool pc);

    // timing
    hipEvent_t get_event();
    int resolve_timing();

    // Non-finite restart (complexLangevin.cu:518-667, 640-650): a device copy of the
    // fields every ckpt_every steps; a non-finite step restores it and moves the
    // noise to a new Philox key (seed0 + epoch * 0x9E3779B97F4A7C15), so that the
    // retried steps do not replay the noise that led to the failure.
    int ckpt_every = 0, max_restarts = 0, restarts = 0, noise_epoch = 0;
    long long last_failed_step = -1;
    long long ckpt_step = -1;
    long long inject_step = -1; // test hook: the step ending at this count leaves a NaN in psi (once)
    unsigned long long seed0 = 0;
    DevBuf<double2> ckpt_omega, ckpt_psi;
    int maybe_checkpoint();
    int restart_from_checkpoint(long long failed_step);

    // Learned linear-response kernel (FTS_SCHEME_SI_LEARNED; the Debye scheme runs the
    // same judge while it records training data).  A_D tables, |k| bins, the wave
    // vectors sorted by bin (with A_D in that order) and their partial-sum chunks;
    // k-space w0, u0, psi, u_psi of the current and previous step (ping-pong).
    DevBuf<double> si_A_omega, si_A_psi;
    DevBuf<unsigned char> lrk_bin;
    DevBuf<int> lrk_perm, lrk_chunk_start, lrk_bin_chunk;
    DevBuf<double> lrk_ad_perm, lrk_chunk_sums, lrk_ring;
    DevBuf<double2> lrk_u, lrk_hat[2][2][2]; // [field][w, u][parity]
    DevBuf<float> lrk_params;
    DevBuf<LrkDevState> lrk_state;
    int lrk_nperm = 0, lrk_nchunks = 0, lrk_par = 0, lrk_capacity = 0;
    long long lrk_read_idx = 0;
    bool lrk_ready = false, lrk_have_prev = false;
    double lrk_threshold = 1.0e-3;
    bool lrk_active() const { return learned_scheme() || lrk_capacity > 0; }
    bool learned_scheme() const { return scheme == FTS_SCHEME_SI_LEARNED || scheme == FTS_SCHEME_SIPC_LEARNED; }
    bool si_scheme() const { return scheme == FTS_SCHEME_SEMI_IMPLICIT || scheme == FTS_SCHEME_SIPC || learned_scheme(); }
    int lrk_setup();
    int lrk_step();
};

// Schemes whose step computes the densities twice (at w and at the predictor w_bar)
inline bool scheme_two_stage(int scheme)
{
    return scheme == FTS_SCHEME_PC_REFERENCE || scheme == FTS_SCHEME_SIPC || scheme == FTS_SCHEME_SIPC_LEARNED;
}

// Builds an engine for system s.  shared: a stream the engine uses instead of
// creating its own (a replica batch's); tune: run the schedule autotuner when
// s->fft_schedule is FTS_SCHED_AUTO.
int engine_create(const fts_system* s, hipStream_t shared, bool tune, fts_engine** out);

// The no-alloc scratch array of the current GPU (noalloc.hip): looks its address up
// once per device (hipGetSymbolAddress with that device current) and returns it.
int noalloc_setup(void** scratch_out);
