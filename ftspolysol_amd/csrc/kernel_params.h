// kernel_params.h -- launch constants and kernel argument types shared by the
// kernels (kernels.h) and the host code that launches them.
#pragma once

#include <hip/hip_runtime.h>

namespace fts {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 256;

// Philox field ids of the semi-implicit k-space noise: mode l draws from field
// SI_NOISE_FIELD + 2l, psi from SI_NOISE_FIELD_PSI (the real-space noise of EM and PC
// uses fields l and 0x100)
constexpr unsigned SI_NOISE_FIELD = 0x400u;
constexpr unsigned SI_NOISE_FIELD_PSI = 0x500u;
constexpr int EW_THREADS = 256;
constexpr int MAX_CLASSES = 64;

// FP32 Boltzmann factors are capped at exp(BOLTZ_CAP) by a per-class shift of W.
constexpr double BOLTZ_CAP = 40.0;

// Interaction-mode coefficients of the CL drift and update kernels.
struct ModeCoefs {
    int nO, nT;
    double c[8];        // rho0 / |d_l|
    double2 g[8 * 8];   // gamma_l O[l][alpha]
    double lam_dt[8];   // dt * lambda_l
    double sigma[8];    // noise standard deviation of mode l
};

} // namespace fts
