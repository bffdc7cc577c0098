/*
 * fts.h -- C-ABI boundary of the MI355X complex-Langevin field-theoretic
 * simulation engine (libfts_hip.so).
 *
 * The reference exposes its hot path only as in-process C++ class methods that
 * take a cufftHandle and a deviceState& (no plugin or FFI layer).  Each entry
 * point below replaces one of those methods; the reference interface it stands
 * for is cited beside it.  Conventions:
 *   - every call returns an int status: FTS_OK, FTS_NONFINITE, FTS_INVALID or
 *     FTS_DEVICE_ERROR; fts_last_error() gives the text of the last failure
 *     (the reference prints "error: ..." and returns false, errHandle.cu:5-9;
 *     it calls exit() on CUDA errors, errHandle.cu:65-84 -- this library never
 *     exits);
 *   - complex arrays are interleaved doubles (re, im), grid index
 *     i = (x*m2 + y)*m3 + z (common.h:54-58), multi-field arrays field-major
 *     [count][M] (deviceState.h:7-13);
 *   - the engine owns device memory; the caller owns host arrays, which are
 *     copied synchronously at the boundary;
 *   - one engine per host thread; an engine handle is not thread-safe.
 */
#ifndef FTS_H
#define FTS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTS_OK 0
#define FTS_NONFINITE 1
#define FTS_INVALID 2
#define FTS_DEVICE_ERROR 3

#define FTS_SCHEME_PC_REFERENCE 0 /* predictor-corrector, complexLangevin.cu:328-368 */
#define FTS_SCHEME_EM 1           /* plain explicit Euler-Maruyama (no reference counterpart) */
#define FTS_SCHEME_SEMI_IMPLICIT 2 /* k-space Debye-preconditioned semi-implicit EM (no reference counterpart) */
#define FTS_SCHEME_SI_LEARNED 3    /* semi-implicit EM with the learned linear-response kernel (no reference counterpart) */
#define FTS_SCHEME_SIPC 4          /* semi-implicit predictor-corrector, Debye kernel (no reference counterpart; see below) */
#define FTS_SCHEME_SIPC_LEARNED 5  /* the same with the learned linear-response kernel (no reference counterpart) */

#define FTS_PREC_FP64 0 /* complex FP64 everywhere, as the reference (common.h:23) */
#define FTS_PREC_FP32 1 /* complex FP32 propagators and FFTs (the FP64 fields are transformed in FP32 on the plane
                           schedules), per-bead log-renormalisation, FP64 reductions and field updates */

#define FTS_MAX_SPECIES 8
#define FTS_MAX_IONS 8
#define FTS_MAX_TYPES 8

/* chain symmetry (fts_system.chain_symmetry).  A chain whose bead sequence and bonds
 * read the same from both ends has qB_s = qF_{N-1-s} exactly (the backward recursion
 * repeats the forward one operation for operation), so the engine runs the forward
 * sweep only and reads qB from the stored qF. */
#define FTS_SYMMETRY_AUTO 0 /* mirror every palindromic species */
#define FTS_SYMMETRY_OFF 1  /* always run both sweeps */

/* bead-step FFT schedules (fts_system.fft_schedule, fts_info.fft_schedule) */
#define FTS_SCHED_AUTO 0  /* time conv and separable at creation, keep the faster (separable when conv does not fit) */
#define FTS_SCHED_PLANE 1 /* pencil(x: FFT, x phi/M, IFFT) + plane(yz: IFFT, x e^{-W}, store q, FFT); needs m2 == m3, plane in LDS;
                             used only when asked for */
#define FTS_SCHED_LINE 2  /* four line passes x, y, z, y; any supported lengths */
#define FTS_SCHED_CONV 3  /* pencil(x: FFT of q_s, x phi_x/M, IFFT) + plane(yz: FFT, x phi_y phi_z, IFFT, x e^{-W},
                             store q_{s+1}); separable bond kernels; needs m2 == m3, plane in LDS */
#define FTS_SCHED_SEPARABLE 4 /* three line passes, one bond-kernel factor each: x (FFT of q_s, x phi_x/M, IFFT),
                                 z (FFT, x phi_z, IFFT), y (FFT, x phi_y, IFFT, x e^{-W}, store q_{s+1}); any supported lengths */
#define FTS_SCHED_CONV_MFMA 5 /* FTS_SCHED_CONV with its x pass as DFT-as-GEMM on the matrix cores (DFT-16 stage on
                                 v_mfma_f32_16x16x4f32); FP32, m1 in {32, 48, 64, 96}; used only when asked for */

/* stress_mask bits: components in the reference's fixed order (stress.cu:56-61) */
#define FTS_STRESS_XX 0x01
#define FTS_STRESS_YY 0x02
#define FTS_STRESS_ZZ 0x04
#define FTS_STRESS_XY 0x08
#define FTS_STRESS_XZ 0x10
#define FTS_STRESS_YZ 0x20
#define FTS_STRESS_PISO 0x40 /* isotropic pressure, observable.cu:297-317 */

/* field ids for fts_get_field (complexLangevin.cu:236-253 plus the propagators) */
#define FTS_FIELD_OMEGA 0      /* [n_omega][M] */
#define FTS_FIELD_PSI 1        /* [M] */
#define FTS_FIELD_RHO_TYPE 2   /* [monomer_types][M] */
#define FTS_FIELD_RHO_P 3      /* [M] */
#define FTS_FIELD_RHO_C 4      /* [M] */
#define FTS_FIELD_RHO_C_POLY 5 /* [M] */
#define FTS_FIELD_RHO_C_ION 6  /* [M] */
#define FTS_FIELD_ION_RHO 7    /* [n_ions][M] */
#define FTS_FIELD_QF 8         /* [N][M] forward propagator of the last species solved */
#define FTS_FIELD_QB 9         /* [N][M] backward propagator of the last species solved */

/* One polymer species: param.h:8-17 plus its sequence block (polymer.cu:123-160). */
typedef struct fts_species_desc {
    long long num_chains;  /* n_p */
    int N;                 /* beads per chain */
    const int* types;      /* [N] monomer type of each bead */
    const double* charges; /* [N] valence of each bead */
    const double* bonds;   /* [N-1] bond lengths b_s, or NULL for all 1.0 (polymer.cu:151) */
} fts_species_desc;

/* [ions.counterions] (param.h:19-26) */
typedef struct fts_counterion_desc {
    int enabled;
    int auto_number; /* neutralise the polymer charge (ion.cu:11-19) */
    long long number;
    double valence;
    double smearing_length; /* <= 0: default smearing length */
} fts_counterion_desc;

/* [ions.salt] (param.h:28-39) */
typedef struct fts_salt_desc {
    int enabled;
    long long num_salt;
    double cation_valence;
    double anion_valence;
    int cation_stoich;
    int anion_stoich;
    double smearing_length; /* <= 0: default smearing length */
    int combine_identical_species; /* merge equal (z, a_j) species (ion.cu:117-127) */
} fts_salt_desc;

/* Everything the reference derives from its parameter and sequence files that the
 * hot path reads (param.cu:342-436, interactionMode.cu:9-298). */
typedef struct fts_system {
    int m1, m2, m3;  /* grid */
    double dL;       /* grid spacing; L_d = dL * m_d, dV = dL^3 (param.cu:348-351) */
    double a;        /* Gaussian smearing length */
    int monomer_types;  /* M */
    double xi;          /* K = xi + chi (1 - delta) */
    const double* chi;  /* [M*M] Flory matrix or NULL (all zero) */
    double eigen_tol_rel; /* active-mode tolerance; < 0 selects the reference default 1e-12 */
    double lB;          /* Bjerrum length; 0 disables electrostatics */
    int n_species;
    const fts_species_desc* species;
    int ions_enabled;
    double default_smearing_length; /* <= 0: a (param.cu:353) */
    fts_counterion_desc counterions;
    fts_salt_desc salt;
    double dt;
    int n_lambda_omega;          /* 1 (broadcast) or n_omega values (param.cu:390-410) */
    const double* lambda_omega;
    double lambda_psi;
    int noise;                   /* 0 disables the CL noise (param.h:100) */
    int scheme;                  /* FTS_SCHEME_* */
    int precision;               /* FTS_PREC_* */
    unsigned long long seed;     /* Philox4x32-10 key */
    int stress_mask;             /* FTS_STRESS_* bits; 0 = no stress observable */
    int fft_schedule;            /* FTS_SCHED_*: bead-step pass schedule; 0 = timed at creation */
    int chain_symmetry;          /* FTS_SYMMETRY_*: mirrored sweep for palindromic chains */
} fts_system;

/* observableRecord (observable.h:7-13), with the partition functions beside it. */
typedef struct fts_observables {
    long long step;
    int n_species;
    int n_ions;
    int stress_mask;
    int pad_;
    double Q_poly[FTS_MAX_SPECIES][2];
    double Q_ion[FTS_MAX_IONS][2];
    double mu_ex_poly[FTS_MAX_SPECIES][2]; /* -ln Q, principal branch (common.h:197-201) */
    double mu_id_poly[FTS_MAX_SPECIES];
    double mu_self_poly[FTS_MAX_SPECIES];
    double mu_ex_ion[FTS_MAX_IONS][2];
    double mu_id_ion[FTS_MAX_IONS];
    double mu_self_ion[FTS_MAX_IONS];
    double H[2];          /* mean-field Hamiltonian density H/V (observable.cu:160-182) */
    double stress[6][2];  /* xx yy zz xy xz yz (stress.cu:161-196); zero if not in mask */
    double p_iso[2];      /* observable.cu:297-317; zero if not in mask */
} fts_observables;

typedef struct fts_info {
    long long grid_size;
    long long step;     /* CL steps taken (noise counter) */
    int n_omega;        /* active interaction modes */
    int monomer_types;
    int n_species;
    int n_ions;
    int n_segment_classes; /* distinct (type, charge) bead classes: e^{-W} fields kept per species */
    int precision;
    int scheme;
    int fft_schedule; /* FTS_SCHED_PLANE, _LINE, _CONV, _SEPARABLE or _CONV_MFMA in use */
    double volume;
    double rho0;
    double short_range_self_pressure; /* param.cu:135-149 */
    double mode_eigenvalue[FTS_MAX_TYPES];   /* active modes, descending (interactionMode.cu:214-263) */
    double omega_coefficient[FTS_MAX_TYPES]; /* rho0/|d_l| (param.cu:115-118) */
    double gamma_is_imag[FTS_MAX_TYPES];     /* 1 if gamma_l = i else 0 (interactionMode.cu:328-332) */
    double mode_O[FTS_MAX_TYPES][FTS_MAX_TYPES]; /* O[l][alpha] of the active modes */
    double ion_valence[FTS_MAX_IONS];
    long long ion_count[FTS_MAX_IONS];
    double ion_smearing[FTS_MAX_IONS];
    double tune_ms[2]; /* autotuner: time of 8 bead steps, conv and separable schedule (0 if not timed) */
    int mirrored_species; /* bit p set: species p runs the mirrored (forward-only) sweep */
    int restarts;         /* non-finite restarts taken (fts_set_restart) */
    int noise_epoch;      /* Philox key in use: seed + noise_epoch * 0x9E3779B97F4A7C15 */
    long long checkpoint_step; /* step counter of the device checkpoint, -1 if none */
    long long failed_step;     /* label of the last non-finite step a restart recovered from, -1 if none */
} fts_info;

/* Accumulated device timing of the propagator sweeps (HIP events on the engine's
 * stream), enabled by fts_set_timing. */
typedef struct fts_timing {
    double sweep_ms;          /* total event time of all propagator sweeps */
    long long bead_steps;     /* bead steps (q_s -> q_{s+1}) executed inside them, both directions */
    long long sweeps;         /* number of timed sweep launches */
    double algorithmic_bytes; /* 2*c*M per bead step, summed */
    long long step_launches;  /* bead-step launch groups: one per bead index, carrying every job (sweep
                               * direction, replica) of that index; sweep_ms / step_launches is the
                               * time of one group (pencil + plane pass, or the four line passes) */
} fts_timing;

typedef struct fts_engine fts_engine;

/* Build an engine: param::derive + interactionMode::build + ion::build +
 * spectral::build_kernels + complexLangevin::initialize (param.cu:342-436,
 * ion.cu:3-46, spectral.cu:3-35, complexLangevin.cu:166-191).  Fields start at 0. */
int fts_create(const fts_system* sys, fts_engine** out);
void fts_destroy(fts_engine* h);
const char* fts_last_error(void);
int fts_get_info(fts_engine* h, fts_info* out);

/* Set omega [n_omega][M] and psi [M] (interleaved complex); removes the psi zero
 * mode as complexLangevin.cu:183 does on load.  psi may be NULL (zero). */
int fts_set_fields(fts_engine* h, const double* omega, const double* psi);

/* Copy a field out: n_doubles must equal the field's size in doubles. */
int fts_get_field(fts_engine* h, int field_id, double* out, size_t n_doubles);

/* complexLangevin::compute_densities at the current fields (complexLangevin.cu:255-260):
 * polymerSpecies::solve_and_compute_densities (polymer.cu:449-478) then
 * ion::compute_densities (ion.cu:54-72).  Q_out: [n_species][2] or NULL. */
int fts_compute_densities(fts_engine* h, double* Q_out);

/* compute_densities then observable::measure (observable.cu:53-67) at the
 * current fields. */
int fts_measure(fts_engine* h, fts_observables* out);

/* Run n CL steps (complexLangevin.cu:669-698 for PC_REFERENCE).  When
 * record_every > 0, every step with step % record_every == 0 stores a record
 * (PC: at the predictor state, as complexLangevin.cu:684; EM/SI: at the state the
 * step starts from) into records[0..capacity); *n_records receives the count.
 * Returns FTS_NONFINITE (fields left at the failing step) on a non-finite Q or field. */
int fts_step(fts_engine* h, int n, int record_every, fts_observables* records, int capacity, int* n_records);

/* Noise counter: the next step draws Philox counter `step`. */
int fts_set_step(fts_engine* h, long long step);

/* Binary state (no reference counterpart: the reference's checkpoint is the tmp_omega /
 * tmp_psi .rf pair written every Ntmp steps with %.10f, complexLangevin.cu:518-556).
 * The file holds the fields exactly with the step counter and the noise key, so a run
 * resumed from it repeats the uninterrupted run bit for bit (FTS_SCHEME_SI_LEARNED:
 * the first resumed step is not judged, as after fts_set_fields).  Layout, host byte
 * order: "FTSSTAT1", int32 m1 m2 m3 n_omega, int64 step, uint64 seed, int32 noise_epoch,
 * int32 restarts, omega [n_omega][M] and psi [M] as interleaved doubles.  Saving writes
 * path.new and renames it into place; loading checks the mesh and mode count and takes
 * the file's seed and noise epoch as the engine's. */
int fts_save_state(fts_engine* h, const char* path);
int fts_load_state(fts_engine* h, const char* path);

/* Checkpoint and non-finite restart (the reference's Ntmp checkpoint and
 * max_nan_restarts, complexLangevin.cu:518-667).  With checkpoint_every > 0 the
 * engine keeps a device copy of (w, psi) taken at the first step of every
 * fts_step call and whenever the step counter is a multiple of checkpoint_every.
 * A step that fails with FTS_NONFINITE restores that copy, rewinds the step
 * counter to it, moves the noise to a new Philox key (seed + epoch *
 * 0x9E3779B97F4A7C15; the reference draws new wall-clock seeds) and drops the
 * records taken after the checkpoint; fts_step still advances the counter by n.
 * After max_restarts restarts a failure is returned.  0, 0 turns it off. */
int fts_set_restart(fts_engine* h, int checkpoint_every, int max_restarts);
/* Test hook (the proposal's fts_inject_nonfinite): the step that brings the step
 * counter to `step` ends with a NaN in psi, once, so that the failure path and the
 * restart can be tested without a physical blow-up. */
int fts_inject_nonfinite(fts_engine* h, long long step);

/* Drift statistics (complexLangevin.cu:318-324, 466-497, the rows of drift_max.dat).
 * While enabled, every step fts_step takes appends one row of n_omega + 2 doubles:
 * the step label, then per interaction mode max over sites |lambda_l / 2 (u_l + u_bar_l)|
 * and last max |lambda_psi / 2 (u_psi + u_psi_bar)| (predictor-corrector: the drifts at
 * (w, psi) and at the predicted fields, as corrector_drift_max; a non-finite value is
 * +inf).  EM and semi-implicit steps report max |lambda u| at the state the step starts
 * from.  A restart from the checkpoint drops the rows after it.  fts_set_drift_stats
 * clears the rows; fts_get_drift_rows moves up to `capacity` of them, oldest first, into
 * out[capacity][n_omega + 2].  Replica batches do not collect them. */
int fts_set_drift_stats(fts_engine* h, int enabled);
int fts_get_drift_rows(fts_engine* h, double* out, int capacity, int* n_rows);

/* Test hooks on the hot path. */
/* 3D DFT of one field on the engine's FFT, unnormalised; sign -1 forward (e^{-ikx}), +1 inverse
 * (the cuFFT convention of common.h:1063-1077). */
int fts_fft3d(fts_engine* h, const double* in, double* out, int sign);
/* One bead step q' = e^{-W} * IFFT[phi_b(k) FFT[q]/M] (polymer.cu:396-400) with bond length b. */
int fts_bead_step(fts_engine* h, const double* q, const double* exp_minus_w, double bond, double* q_out);

/* fts_bead_step on device arrays without allocating.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  q, exp_minus_w and q_out are device arrays of
 * M interleaved complex doubles on the engine's GPU (q_out may not alias the inputs); the
 * work is enqueued on `stream` (a hipStream_t; NULL: the engine's stream) and the call
 * returns without waiting, never allocating or freeing device memory and never
 * synchronising.  The result equals fts_bead_step's bit for bit.  The intermediates live
 * in one scratch array per GPU (about 24 MiB, held until the process ends), so grids are
 * limited to 2^20 points (larger ones fail with FTS_INVALID), and calls that overlap on
 * one GPU (two streams or host threads) race on it.  The engine's GPU must be the current
 * device.  The engine's own buffers and state are not touched. */
int fts_bead_step_noalloc(fts_engine* h, const double* q, const double* exp_minus_w, double bond, double* q_out, void* stream);

/* Device timing of the propagator sweeps (engine only; the oracle ignores it). */
int fts_set_timing(fts_engine* h, int enabled);
int fts_get_timing(fts_engine* h, fts_timing* out);

/* ---- Replica batches ------------------------------------------------------
 * No counterpart in the reference, which runs one system per process
 * (main.cu:9-93); this is the batched form: many independent CL replicas of
 * uneven mesh, chain length and salt stepping together.  Replicas that share
 * mesh, precision and grid spacing form a group: every bead-step launch of a
 * group carries the propagator jobs of all its replicas, and groups run
 * concurrently on their own streams.  Each replica is an fts_engine owned by
 * the batch: fts_set_fields, fts_get_field, fts_measure, fts_get_info and the
 * other single-engine calls work on it (never fts_destroy it, and never call
 * fts_step on it while a batch step runs).  All replicas must use one scheme. */
#define FTS_MAX_BATCH_GROUPS 16

typedef struct fts_batch fts_batch;

typedef struct fts_batch_info {
    int n_replicas;
    int n_groups;
    int group_size[FTS_MAX_BATCH_GROUPS];
    int group_mesh[FTS_MAX_BATCH_GROUPS][3];
    int group_precision[FTS_MAX_BATCH_GROUPS];
    int group_schedule[FTS_MAX_BATCH_GROUPS];        /* FTS_SCHED_PLANE, _LINE, _CONV or _SEPARABLE */
    double group_tune_ms[FTS_MAX_BATCH_GROUPS][2];   /* autotuner: 8 batched bead steps, conv and separable */
} fts_batch_info;

/* Build n replicas from systems[0..n-1] (each with its own seed); fields start at 0.
 * A group whose first system asks for FTS_SCHED_AUTO has its schedule timed with all
 * of its replicas in each launch. */
int fts_batch_create(const fts_system* systems, int n, fts_batch** out);
void fts_batch_destroy(fts_batch* b);
int fts_batch_get_info(fts_batch* b, fts_batch_info* out);
/* Borrowed engine handle of replica r and the index of its group. */
int fts_batch_replica(fts_batch* b, int r, fts_engine** engine, int* group);
/* n CL steps of every replica, each exactly as fts_step would take it (same
 * noise stream, same records).  records: [n_replicas][capacity], replica r's
 * records from records + r * capacity; n_records: [n_replicas].  On a failure
 * the call returns its status, and fts_last_error names the replica. */
int fts_batch_step(fts_batch* b, int n, int record_every, fts_observables* records, int capacity, int* n_records);
/* Graph level of the batch's steps (same kernels, same results at every level):
 *   2 (default) each CL step of a group -- its sweeps, every replica's own work on the
 *     side streams and the copies of Q and the finiteness flags -- is captured once as
 *     a hipGraph and replayed; a step is captured again when the replicas taking part or
 *     recording, the learned kernel's state or a replica's configuration change.  The
 *     noise key {seed, step} reaches the kernels through device memory that a captured
 *     copy refreshes from pinned host memory on every replay.  Steps of replicas with
 *     drift statistics or an armed fts_inject_nonfinite run as at level 1;
 *   1 the bead-step launches of each batched sweep are captured once and replayed, the
 *     rest is launched one by one;
 *   0 everything is launched one by one. */
int fts_batch_set_graph(fts_batch* b, int level);
/* The work of single replicas between the batched sweeps (smoothing FFTs, ions,
 * contractions, field updates, records) runs on n side streams per group, forked
 * from the group's stream and joined back into it (replica r on side stream
 * r mod n), so the small launches of several replicas share the GPU.  n = 0 runs
 * it all on the group's stream.  Default min(group size, FTS_BATCH_DEFAULT_STREAMS);
 * the results do not depend on n. */
#define FTS_BATCH_DEFAULT_STREAMS 8
#define FTS_BATCH_MAX_STREAMS 16
int fts_batch_set_streams(fts_batch* b, int n);
/* Device timing of the batched sweeps of one group (fts_timing as for an engine). */
int fts_batch_set_timing(fts_batch* b, int enabled);
int fts_batch_get_timing(fts_batch* b, int group, fts_timing* out);

/* ---- Front end: input decks and field files -------------------------------
 * The reference reads an INI-like parameter file (param.cu:165-436) and a block
 * sequence file (polymer.cu:219-273) into its param object, and reads and writes
 * fields as .rf text tables (common.h:229-471, fieldIO.cu:213-726).  A deck holds
 * both files, parsed with the reference's keys, aliases, defaults and checks;
 * fts_deck_system fills the engine's system description from it.  Keys this
 * engine adds, all in [dynamics]: scheme = pc | em | si | si_learned (default pc,
 * the reference's), precision = fp64 | fp32 (default fp64) and seed (Philox key,
 * default 0x5EED; the reference draws wall-clock seeds, common.h:922-927). */
typedef struct fts_deck fts_deck;

typedef struct fts_run_settings {
    int Nt, Nss, Ns, Ntmp, Nso, N_stat, max_nan_restarts; /* [dynamics] / [output] (param.h:62-70) */
    int legacy_ab;   /* no [monomers] and no chi: omega files hold the legacy A/B columns (param.cu:145-148) */
    int n_ion_names; /* ion species after merging; names via fts_deck_string(d, "ion.<j>") */
} fts_run_settings;

/* param::read + polymerSpecies::setup + param::validate (main.cu:15-24);
 * sequence_file overrides [polymer] sequence_file when not NULL (--seq). */
int fts_deck_read(const char* param_file, const char* sequence_file, fts_deck** out);
void fts_deck_destroy(fts_deck* d);
/* The system description of the deck; its pointers stay valid while the deck lives. */
int fts_deck_system(fts_deck* d, fts_system* out);
int fts_deck_settings(fts_deck* d, fts_run_settings* out);
/* [output] file names ("omega_file", "psi_file", "observable_file", ...), "sequence_file",
 * and the labels "species.<p>", "ion.<j>", "monomer.<a>"; NULL if unknown. */
const char* fts_deck_string(fts_deck* d, const char* key);

/* Read an .rf table: m holds the mesh to assume when the file has no header and
 * receives the file's mesh; *n_fields receives the complex columns per row.  With
 * data == NULL only m and n_fields are filled; otherwise data receives the fields
 * field-major [n_fields][M] interleaved complex (capacity in doubles). */
int fts_rf_read(const char* path, int* m, int* n_fields, double* data, size_t capacity_doubles);
/* Periodic trilinear resize of `count` complex fields from mesh n to mesh m
 * (trilinear::resize, trilinear.cu:11-64); in [count][n1 n2 n3], out [count][m1 m2 m3]. */
int fts_field_resize(const double* in, const int* n, int count, double* out, const int* m);
/* Write an .rf table (header "m1 m2 m3" and "L1 L2 L3", rows of "%.10f  %.10f" per field). */
int fts_rf_write(const char* path, const int* m, const double* L, int n_fields, const double* fields);

/* ---- Learned linear-response kernel (LRK) ---------------------------------
 * No counterpart in the reference, which has no semi-implicit scheme at all.
 *
 * The semi-implicit update (FTS_SCHEME_SEMI_IMPLICIT) is
 *     w <- w + S_P[-dt lam u(w)] + S_Pn[eta],   P(k) = 1 / (1 + dt lam A(k)),
 *     Pn(k) = sqrt(P (1 + P) / 2) = P sqrt(1 + dt lam A / 2),
 * with A = A_D, the Debye (discrete-chain RPA) linear response of the
 * homogeneous state.  The noise factor Pn makes the stationary variance of a
 * linear mode (u = A w) exactly 1 / A at any dt; P alone would give
 * 2 / (A (2 + dt lam A)), an O(dt) deficit of the short-wavelength fluctuations.
 * eta, real white noise of variance 2 lam dt / dV per site, is drawn in k-space:
 * FFT[eta](k) = sqrt(2 lam dt M / dV) g(k), g Hermitian; the pair {k, -k} takes
 * the two Box-Muller normals a (cosine) and b (sine) of one Philox draw (seed,
 * step, field F, the pair's smaller grid index), g = (a + i b) / sqrt 2 at the
 * smaller index and its conjugate at the other, a self-conjugate k takes a alone;
 * F = 0x400 + 2 l for mode l and 0x500 for psi, whose k = 0 term is dropped
 * (mean-free noise).  The
 * increment's transform adds it in its multiply pass, so the noise costs no
 * transform of its own.  FP32 transforms (an FP32 engine on the plane schedules)
 * evaluate the Box-Muller logarithm, root and sine/cosine in FP32 (the draws
 * themselves are the same).  FTS_SCHEME_SI_LEARNED replaces it by the field-adapted
 *     A(k) = A_D(k) exp(r_c(bin(|k|))),   r_c = 4 tanh(z_c / 4),
 * where z = MLP(|k|, moments of the current fields): field c = 0 is interaction
 * mode 0, c = 1 is psi.  The MLP has FTS_LRK_FEATURES inputs (padded to
 * FTS_LRK_IN), three tanh layers of FTS_LRK_HIDDEN units and FTS_LRK_OUT linear
 * outputs, FP32 weights.  Features: x0 = (bin + 0.5) / FTS_LRK_BINS, the centre of
 * the |k| bin on the scale k_ref = sqrt(3) pi / dL (bin = floor(BINS |k| / k_ref)
 * clamped to BINS - 1), then the moments std Re w0, std Im w0, mean Re w0,
 * mean Im w0, std Re psi, std Im psi over the grid; feature j enters the first
 * layer as (f_j - shift_j) * scale_j.
 *
 * The exact force judges the kernel on every step: with dW = FFT[w_t - w_{t-1}]
 * and dU = FFT[u(w_t) - u(w_{t-1})], the residual of a kernel A on that step is
 *     R_c(A) = sum_{k != 0} |dU - A dW|^2
 *            = E_c - 2 sum_b e^{r_b} S1_{c,b} + sum_b e^{2 r_b} S2_{c,b}
 * with E_c = sum |dU|^2, S1_{c,b} = sum_{k in b} A_D Re(dU conj dW) and
 * S2_{c,b} = sum_{k in b} A_D^2 |dW|^2.  The learned kernel is used for the next
 * step while its excess force residual over the Debye kernel on the same step,
 * max_c (R_c(learned) - R_c(A_D)) / E_c, stays at or below the threshold; otherwise
 * the step falls back to A_D (the kernel is still judged, and comes back when it
 * does well again).  The first step after creation, fts_set_fields or a restart has
 * nothing judged yet and uses A_D.
 *
 * The network minimises the same residual: a training record is one judged step
 * (moments at its start, E, S1, S2), and the loss of a record is
 *     sum_c sum_b (S2_{c,b} / E_c) (e^{r_{c,b}} - S1_{c,b} / S2_{c,b})^2
 *     = sum_c (R_c(learned) - min_r R_c) / E_c.
 * The batch loss is the mean over its records.
 *
 * The semi-implicit predictor-corrector (FTS_SCHEME_SIPC with A = A_D,
 * FTS_SCHEME_SIPC_LEARNED with the learned A, judged as above) takes the
 * semi-implicit step above as its predictor, w_bar = w + S_P[-dt lam u(w)] + S_Pn[eta],
 * and then, with the densities at w_bar,
 *     w <- w + S_P2[-(dt lam / 2) (u(w) + u(w_bar))] + S_P2[eta] + S_{1-P2}[w_bar - w],
 *     P2(k) = 1 / (1 + dt lam A(k) / 2) = 2 P / (1 + P),
 * with the same eta (the same k-space draw).  For a linear drift u = A w this is the
 * trapezoidal (Crank-Nicolson) step w <- (1 - a/2) / (1 + a/2) w + P2 eta, a = dt lam A:
 * second order, stable at every dt, and the stationary variance of every mode is
 * exactly 1 / A.  psi likewise (k = 0 dropped, mean removed).
 * Records are taken at the predictor state with its Q, as the reference's
 * predictor-corrector takes them. */
#define FTS_LRK_FEATURES 7
#define FTS_LRK_IN 8
#define FTS_LRK_HIDDEN 64
#define FTS_LRK_OUT 2
#define FTS_LRK_BINS 128
#define FTS_LRK_MOMENTS 6
/* parameter block (floats): shift[8] scale[8] W1[64][8] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[2][64] b4[2]; W[out][in] row-major */
#define FTS_LRK_OFF_SHIFT 0
#define FTS_LRK_OFF_SCALE 8
#define FTS_LRK_OFF_W1 16
#define FTS_LRK_OFF_B1 528
#define FTS_LRK_OFF_W2 592
#define FTS_LRK_OFF_B2 4688
#define FTS_LRK_OFF_W3 4752
#define FTS_LRK_OFF_B3 8848
#define FTS_LRK_OFF_W4 8912
#define FTS_LRK_OFF_B4 9040
#define FTS_LRK_NPARAMS 9042
/* training record (doubles): moments[6] E[2] S1[2][BINS] S2[2][BINS] */
#define FTS_LRK_REC_MOMENTS 0
#define FTS_LRK_REC_E 6
#define FTS_LRK_REC_S1 8
#define FTS_LRK_REC_S2 (8 + 2 * FTS_LRK_BINS)
#define FTS_LRK_RECORD (8 + 4 * FTS_LRK_BINS)

typedef struct fts_lrk_stats {
    long long judged;         /* steps judged by the exact force */
    long long learned_steps;  /* steps taken with the learned kernel */
    long long fallback_steps; /* steps taken with A_D (first steps and fallbacks) */
    long long records;        /* training records written since recording started */
    double last_excess;       /* max_c (R_c(learned) - R_c(A_D)) / E_c of the last judged step */
    double sum_rel_residual_learned[2]; /* sum over judged steps of R_c(learned) / E_c */
    double sum_rel_residual_debye[2];   /* sum over judged steps of R_c(A_D) / E_c */
} fts_lrk_stats;

/* Load the network (FTS_LRK_NPARAMS floats, copied) and the fallback threshold
 * into an engine.  Engines start with all-zero parameters (r = 0, i.e. A_D) and
 * threshold 1e-3. */
int fts_lrk_set_network(fts_engine* h, const float* params, double threshold);
/* Write a training record for every judged step into a device ring of `capacity`
 * records (0 stops recording and frees it).  Needs one of the semi-implicit schemes
 * (FTS_SCHEME_SEMI_IMPLICIT, _SI_LEARNED, _SIPC, _SIPC_LEARNED), and
 * 1 <= n_omega + (lB > 0) with n_omega <= 1. */
int fts_lrk_set_recording(fts_engine* h, int capacity);
/* Copy the records written since the last call (at most max_records, the oldest
 * first; older ones are lost if the ring wrapped) and empty the ring. */
int fts_lrk_get_records(fts_engine* h, double* out, int max_records, int* n_out);
int fts_lrk_get_stats(fts_engine* h, fts_lrk_stats* out);

/* The network on the device, outside an engine (training and tests).
 * forward: features [n][FTS_LRK_FEATURES] -> r [n][FTS_LRK_OUT]. */
int fts_lrk_forward(const float* params, const float* features, int n, float* r_out);
/* Loss of n_records training records (mean over records) and, if grad is not
 * NULL, its gradient with respect to every parameter except shift and scale
 * (those entries are set to 0). */
int fts_lrk_loss_grad(const float* params, const double* records, int n_records, double* loss, float* grad);

/* Trainer: Adam on minibatches of records, on the device. */
typedef struct fts_lrk_trainer fts_lrk_trainer;
int fts_lrk_trainer_create(const double* records, int n_records, fts_lrk_trainer** out);
void fts_lrk_trainer_destroy(fts_lrk_trainer* t);
/* `epochs` passes over the records in an order shuffled per epoch from `seed`,
 * batch_records records per Adam step, learning rate decaying geometrically
 * from lr0 to lr1 over the run.  params is updated in place (shift and scale are
 * kept); loss_history[e] (if not NULL) receives the mean minibatch loss of epoch e. */
int fts_lrk_trainer_run(fts_lrk_trainer* t, float* params, int epochs, int batch_records, double lr0, double lr1,
                        unsigned long long seed, double* loss_history);

#ifdef __cplusplus
}
#endif

#endif /* FTS_H */
