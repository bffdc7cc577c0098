// lrk.h -- device state and host launchers of the learned linear-response
// kernel (FTS_SCHEME_SI_LEARNED; the formulas are the ones include/fts.h states).
//
// Per semi-implicit step the engine runs, on its stream:
//   lrk_moment_sums   4 complex partial sums -> (mean, mean square) of Re/Im w0, Re/Im psi
//   lrk_infer         network at the FTS_LRK_BINS bin centres -> r_cur (FP32)
//   lrk_bin_chunks    per-chunk sums of S1, S2, E over the bin-sorted wave vectors
//   lrk_judge         residuals of r_prev and of A_D, decision, counters, record;
//                     then r_prev <- r_cur, m_prev <- m_cur
//   lrk_ptable        P_c(k) = 1 / (1 + dt lam_c A_D,c(k) e^{use ? r_c(bin(k)) : 0})
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/fts.h"

namespace fts {

constexpr int LRK_NB = FTS_LRK_BINS;
constexpr int LRK_CHUNK = 2048; // wave vectors per partial-sum chunk of lrk_bin_chunks

struct LrkDevState {
    float r_cur[LRK_NB * 2];  // [bin][field] network output at the current state
    float r_prev[LRK_NB * 2]; // the output the previous step was taken with (judged now)
    double m_cur[8];          // moments of the current state
    double m_prev[8];         // moments at the start of the judged step
    int use;                  // 1: this step uses the learned kernel
    int pad_;
    long long judged, learned, fallback, records, write_idx;
    double last_excess;
    double sum_rl[2], sum_rd[2];
};

// Partial sums for the moments: partials[k * RED_BLOCKS + b], k = 0..3:
// (Re w0, Re w0^2), (Im w0, Im w0^2), (Re psi, Re psi^2), (Im psi, Im psi^2).
// w0 or psi may be NULL (zero field).
hipError_t lrk_moment_sums(const double2* w0, const double2* psi, int M, double2* partials, hipStream_t s);
// means: the 4 finalized (mean, mean square) pairs.
hipError_t lrk_infer(const float* params, const double2* means, LrkDevState* st, hipStream_t s);
// chunk partials: out[chunk][6] = S1_0, S2_0, E_0, S1_1, S2_1, E_1.  The spectra are
// complex FP64, or complex FP32 (fp32 != 0: the pointers hold float2 values).
hipError_t lrk_bin_chunks(int n_chunks, const int* chunk_start, const int* perm, const double* ad_perm, int n_perm,
                          const void* Wc0, const void* Wp0, const void* Uc0, const void* Up0,
                          const void* Wc1, const void* Wp1, const void* Uc1, const void* Up1, int fp32, double* out,
                          hipStream_t s);
// bin_chunk[b] .. bin_chunk[b + 1]: the chunks of bin b.  have_prev = 0 skips the judgement.
hipError_t lrk_judge(LrkDevState* st, int have_prev, const int* bin_chunk, const double* chunk_sums, double threshold,
                     int learned_scheme, double* ring, int capacity, hipStream_t s);
hipError_t lrk_ptable(const LrkDevState* st, const unsigned char* bin, const double* A0, const double* A1, double lamdt0,
                      double lamdt1, int M, double* P0, double* P1, hipStream_t s);

} // namespace fts
