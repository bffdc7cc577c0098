// kernels.h -- elementwise, contraction and reduction kernels of the engine.
//
// Fields are complex FP64 (double2), grid index i = (x*m2 + y)*m3 + z.  Propagator
// storage and the hot FFTs use T = double or float.  Every reduction is FP64 and
// deterministic: a fixed grid of RED_BLOCKS blocks writes per-block partials,
// and a one-block finalize kernel sums them in a fixed order.
#pragma once

#include "fts_device.h"
#include "kernel_params.h"

namespace fts {

// Sum over the block of K complex values; the result is valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum(double2 (&v)[K])
{
    __shared__ double2 sh[K][RED_THREADS / 64];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            v[k].x += __shfl_down(v[k].x, off, 64);
            v[k].y += __shfl_down(v[k].y, off, 64);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) sh[k][w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double2 s = sh[k][0];
            for (int j = 1; j < RED_THREADS / 64; ++j) s = cadd(s, sh[k][j]);
            v[k] = s;
        }
    }
}

// partials[k * RED_BLOCKS + b] -> out[k] = scale * sum_b partials
__global__ void k_finalize(const double2* __restrict__ partials, int K, double2* __restrict__ out, double scale)
{
    for (int k = 0; k < K; ++k) {
        double2 v[1] = {partials[k * RED_BLOCKS + threadIdx.x]};
        block_sum<1>(v);
        if (threadIdx.x == 0) out[k] = cscale(v[0], scale);
        __syncthreads();
    }
}

// dst = scale * src, converting the complex precision
template <typename S, typename D>
__global__ void k_cast(const V2<S>* __restrict__ src, V2<D>* __restrict__ dst, double scale, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * EW_THREADS + threadIdx.x;
    if (i < n) dst[i] = cmk<V2<D>>(static_cast<D>(scale * src[i].x), static_cast<D>(scale * src[i].y));
}

// separable bond kernel of one bond length: out[i] = exp(-b^2 k_i^2 / 6) for the
// m1 + m2 + m3 axis wave numbers k (x, then y, then z), the factors of
// phi_b(k) = exp(-b^2 |k|^2 / 6) (spectral.cu:22-30)
__global__ void k_bond_table(const double* __restrict__ kvec, double bond, double* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = exp(-bond * bond * (kvec[i] * kvec[i]) / 6.0);
}

// ---- plain sums -------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(RED_THREADS) k_sum(const V2<T>* __restrict__ a, int M, double2* __restrict__ partials)
{
    double2 v[1] = {d2(0.0, 0.0)};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        const V2<T> x = a[i];
        v[0].x += static_cast<double>(x.x);
        v[0].y += static_cast<double>(x.y);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// sum of a * b (complex product)
__global__ void __launch_bounds__(RED_THREADS) k_sum_product(const double2* __restrict__ a, const double2* __restrict__ b, int M,
                                                             double2* __restrict__ partials)
{
    double2 v[1] = {d2(0.0, 0.0)};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) v[0] = cadd(v[0], cmul(a[i], b[i]));
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// sum of the real CL noise sigma * n(seed, step, field, i); key = {seed, step} on the device
// (the noise key of the step, refreshed from the host before every step)
__global__ void __launch_bounds__(RED_THREADS) k_sum_noise(const unsigned long long* __restrict__ key, unsigned field, double sigma, int M,
                                                           double2* __restrict__ partials)
{
    const unsigned long long seed = key[0], step = key[1];
    double2 v[1] = {d2(0.0, 0.0)};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS)
        v[0].x += sigma * cl_normal(seed, step, field, static_cast<uint32_t>(i));
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// ion Boltzmann factor b = exp(-i z psi~) = exp(z psi~.y - i z psi~.x), stored and summed
// (minusIValenceBoltzmannKernel, common.h:547-552, and its mean, common.h:1104-1115)
__global__ void __launch_bounds__(RED_THREADS) k_ion_boltzmann(const double2* __restrict__ ps, double z, int M, double2* __restrict__ b,
                                                               double2* __restrict__ partials)
{
    double2 v[1] = {d2(0.0, 0.0)};
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        const double2 p = ps[i];
        const double2 e = cexp(d2(z * p.y, -z * p.x));
        b[i] = e;
        v[0] = cadd(v[0], e);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// Minimum of Re W over a bead class (partials in .x): the FP32 shift that caps
// the class's Boltzmann factor at exp(BOLTZ_CAP).
__global__ void __launch_bounds__(RED_THREADS) k_min_class_w(const double2* __restrict__ Wt, const double2* __restrict__ ps, double z, int M,
                                                             double2* __restrict__ partials)
{
    __shared__ double sh[RED_THREADS];
    double v = 1.0e300;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        const double2 w = cadd(Wt[i], cmul(ps[i], d2(0.0, z)));
        v = fmin(v, w.x);
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = d2(sh[0], 0.0);
}

// out[0] = add + min over the RED_BLOCKS partials (.x)
__global__ void k_finalize_min(const double2* __restrict__ partials, double add, double* __restrict__ out)
{
    __shared__ double sh[RED_THREADS];
    sh[threadIdx.x] = partials[threadIdx.x].x;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] + add;
}

// k-space contraction: out_c = sum_k w_c(k) S(k) A(k) B(-k), c = xx yy zz xy xz yz 1
template <typename T>
__global__ void __launch_bounds__(RED_THREADS) k_contract_kspace(const V2<T>* __restrict__ A, const V2<T>* __restrict__ B, const double* __restrict__ kvec,
                                                                 const double* __restrict__ sep, int m1, int m2, int m3, double2* __restrict__ partials)
{
    double2 v[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] = d2(0.0, 0.0);
    const int M = m1 * m2 * m3;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        const int z = i % m3, y = (i / m3) % m2, x = i / (m2 * m3);
        const int xr = x == 0 ? 0 : m1 - x, yr = y == 0 ? 0 : m2 - y, zr = z == 0 ? 0 : m3 - z;
        const V2<T> a = A[i];
        const V2<T> b = B[(xr * m2 + yr) * m3 + zr];
        double2 p = cmul(d2(a.x, a.y), d2(b.x, b.y));
        if (sep) p = cscale(p, sep[x] * sep[m1 + y] * sep[m1 + m2 + z]);
        const double kx = kvec[x], ky = kvec[m1 + y], kz = kvec[m1 + m2 + z];
        v[0] = cadd(v[0], cscale(p, kx * kx));
        v[1] = cadd(v[1], cscale(p, ky * ky));
        v[2] = cadd(v[2], cscale(p, kz * kz));
        v[3] = cadd(v[3], cscale(p, kx * ky));
        v[4] = cadd(v[4], cscale(p, kx * kz));
        v[5] = cadd(v[5], cscale(p, ky * kz));
        v[6] = cadd(v[6], p);
    }
    block_sum<7>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < 7; ++c) partials[c * RED_BLOCKS + blockIdx.x] = v[c];
    }
}

// ---- elementwise ------------------------------------------------------------
__global__ void k_fill(double2* __restrict__ a, double2 v, int M)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i < M) a[i] = v;
}

// a[i] += v[0] * sign  (remove_complex_mean with sign = -1, common.h:1008-1012)
__global__ void k_add_scalar(double2* __restrict__ a, const double2* __restrict__ v, double sign, int M)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i < M) a[i] = cadd(a[i], cscale(v[0], sign));
}

// W_type[alpha] = sum_l coef[l][alpha] * omega_s[l]   (polymer.cu:352-359)
__global__ void k_combine_modes(const double2* __restrict__ om_s, const double2* __restrict__ coef, int nO, int nT, int M,
                                double2* __restrict__ Wt)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    for (int al = 0; al < nT; ++al) {
        double2 w = d2(0.0, 0.0);
        for (int l = 0; l < nO; ++l) w = cadd(w, cmul(om_s[static_cast<size_t>(l) * M + i], coef[l * nT + al]));
        Wt[static_cast<size_t>(al) * M + i] = w;
    }
}

// Boltzmann factor of a bead class c: E_c = exp(-(W~_alpha + i z psi~) + shift_c)
// (build_segment_fields + build_exponentials, polymer.cu:365-384)
template <typename T>
__global__ void k_boltzmann(const double2* __restrict__ Wt, const double2* __restrict__ ps, const int* __restrict__ cls_type,
                            const double* __restrict__ cls_z, const double* __restrict__ cls_shift, int ncls, int M, V2<T>* __restrict__ E)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    for (int c = 0; c < ncls; ++c) {
        double2 w = Wt[static_cast<size_t>(cls_type[c]) * M + i];
        const double z = cls_z[c];
        if (z != 0.0) w = cadd(w, cmul(ps[i], d2(0.0, z)));
        const double2 e = cexp(d2(-w.x + cls_shift[c], -w.y));
        E[static_cast<size_t>(c) * M + i] = cmk<V2<T>>(static_cast<T>(e.x), static_cast<T>(e.y));
    }
}

// Density contraction for one species (accumulatePolymerDensityByChainDensityKernel,
// common.h:560-580, over all beads at once):
//   rho_hat[alpha] += f * sum_{s in alpha} qF_s qB_s e^{W_s - shift},  rho_c_hat += f * sum_s z_s (...)
// with f = chain_density / Q.  Beads are visited class by class (cls_start / cls_beads).
template <typename T>
__global__ void k_contract_density(const V2<T>* __restrict__ qF, const V2<T>* __restrict__ qB, const int* __restrict__ cls_start,
                                   const int* __restrict__ cls_beads, const int* __restrict__ cls_type, const double* __restrict__ cls_z,
                                   const double* __restrict__ cls_shift, int ncls, const double2* __restrict__ Wt,
                                   const double2* __restrict__ ps, const double2* __restrict__ Qp, double chain_density, int electro,
                                   const double* __restrict__ rfac, int M, int mirrorN, double2* __restrict__ rho_hat,
                                   double2* __restrict__ rho_c_hat)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const double2 f = cscale(cinv(Qp[0]), chain_density);
    double2 charge = d2(0.0, 0.0);
    for (int c = 0; c < ncls; ++c) {
        double2 acc = d2(0.0, 0.0);
        for (int t = cls_start[c]; t < cls_start[c + 1]; ++t) {
            const int bead = cls_beads[t];
            const V2<T> a = qF[static_cast<size_t>(bead) * M + i];
            // mirrored species: qB_s = qF_{N-1-s}; the lists hold beads s <= N-1-s and the
            // pair (s, N-1-s) contributes twice the product (rfac is symmetric under s -> N-1-s)
            const V2<T> b = mirrorN ? qF[static_cast<size_t>(mirrorN - 1 - bead) * M + i] : qB[static_cast<size_t>(bead) * M + i];
            double2 p = cmul(d2(a.x, a.y), d2(b.x, b.y));
            if (rfac) p = cscale(p, rfac[bead]);
            if (mirrorN && 2 * bead != mirrorN - 1) p = cscale(p, 2.0);
            acc = cadd(acc, p);
        }
        double2 w = Wt[static_cast<size_t>(cls_type[c]) * M + i];
        const double z = cls_z[c];
        if (z != 0.0) w = cadd(w, cmul(ps[i], d2(0.0, z)));
        const double2 v = cmul(cmul(acc, cexp(d2(w.x - cls_shift[c], w.y))), f);
        double2* r = rho_hat + static_cast<size_t>(cls_type[c]) * M + i;
        *r = cadd(*r, v);
        if (electro && z != 0.0) charge = cadd(charge, cscale(v, z));
    }
    if (electro) rho_c_hat[i] = cadd(rho_c_hat[i], charge);
}

// Renormalisation bookkeeping of one species' sweeps (one thread).  logc holds
// the base-2 exponents of the per-bead scales: with LF(s) = sum_{t<=s} logcF[t]
// and LB(s) = sum_{t>=s} logcB[t] the stored propagators are qF_s 2^{LF(s)} and
// qB_s 2^{LB(s)}, so
//   rfac[s]  = 2^{LF(N-1) - LF(s) - LB(s)}     (density term qF_s qB_s / Q)
//   rfac2[s] = 2^{LF(N-1) - LF(s) - LB(s+1)}   (stress term qF_s qB_{s+1} / Q)
// and out->x = LF(N-1) ln 2, the log-gain of Q.
__global__ void k_scale_prefix(const double* __restrict__ logcF, const double* __restrict__ logcB, int N, int mirror, double* __restrict__ rfac,
                               double* __restrict__ rfac2, double2* __restrict__ out)
{
    // one block; dynamic LDS holds LF(s) = sum_{t<=s} logcF[t] and LB(s) = sum_{t>=s} logcB[t]
    // (mirrored species: the backward exponents are the forward ones reversed)
    extern __shared__ double sh_prefix[];
    double* LF = sh_prefix;
    double* LB = sh_prefix + N;
    for (int s = threadIdx.x; s < N; s += blockDim.x) {
        LF[s] = logcF[s];
        LB[s] = mirror ? logcF[N - 1 - s] : logcB[s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int s = 1; s < N; ++s) LF[s] += LF[s - 1];
        for (int s = N - 2; s >= 0; --s) LB[s] += LB[s + 1];
        out->x = LF[N - 1] * 0.69314718055994530942;
        out->y = 0.0;
    }
    __syncthreads();
    const double LFtot = LF[N - 1];
    for (int s = threadIdx.x; s < N; s += blockDim.x) {
        rfac[s] = exp2(LFtot - LF[s] - LB[s]);
        if (s < N - 1) rfac2[s] = exp2(LFtot - LF[s] - LB[s + 1]);
    }
}

// rho = (n/V) b / Q   (normalizedDensityKernel, common.h:554-558)
__global__ void k_ion_density(const double2* __restrict__ b, const double2* __restrict__ Q, double density, int M, double2* __restrict__ rho)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    rho[i] = cmul(b[i], cscale(cinv(Q[0]), density));
}

// out = sum of nsrc fields (rho_p = sum_alpha rho_alpha)
__global__ void k_sum_fields(const double2* __restrict__ src, int nsrc, int M, double2* __restrict__ out)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    double2 s = d2(0.0, 0.0);
    for (int k = 0; k < nsrc; ++k) s = cadd(s, src[static_cast<size_t>(k) * M + i]);
    out[i] = s;
}

// a += c; b += c
__global__ void k_add2(double2* __restrict__ a, double2* __restrict__ b, const double2* __restrict__ c, int M)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const double2 v = c[i];
    a[i] = cadd(a[i], v);
    b[i] = cadd(b[i], v);
}

// psi drift before its mean is removed: out = lap / (4 pi lB) + i rho   (complexLangevin.cu:299-311)
__global__ void k_psi_drift(const double2* __restrict__ lap, const double2* __restrict__ rho, double s, int M, double2* __restrict__ out)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    double2 v = cscale(lap[i], s);
    v.x -= rho[i].y;
    v.y += rho[i].x;
    out[i] = v;
}

__device__ __forceinline__ double2 mode_drift(const ModeCoefs& mc, int l, const double2* __restrict__ om, const double2* __restrict__ rho, int M, int i)
{
    double2 u = cscale(om[static_cast<size_t>(l) * M + i], mc.c[l]);
    for (int al = 0; al < mc.nT; ++al) u = cadd(u, cmul(rho[static_cast<size_t>(al) * M + i], mc.g[l * mc.nT + al]));
    return u;
}

// predictor (complexLangevin.cu:272-285, 328-340): u = drift(omega, rho); omega_bar = omega - dt lam u + eta
__global__ void k_omega_predict(ModeCoefs mc, const double2* __restrict__ om, const double2* __restrict__ rho,
                                const unsigned long long* __restrict__ key, int M, double2* __restrict__ drift, double2* __restrict__ om_bar)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const unsigned long long seed = key[0], step = key[1];
    for (int l = 0; l < mc.nO; ++l) {
        const size_t o = static_cast<size_t>(l) * M + i;
        const double2 u = mode_drift(mc, l, om, rho, M, i);
        drift[o] = u;
        double2 w = cadd(om[o], cscale(u, -mc.lam_dt[l]));
        w.x += 1.0 * (mc.sigma[l] * cl_normal(seed, step, static_cast<uint32_t>(l), static_cast<uint32_t>(i)));
        om_bar[o] = w;
    }
}

// corrector (complexLangevin.cu:349-360): omega += -dt lam/2 (u + u_bar) + eta (the same eta)
__global__ void k_omega_correct(ModeCoefs mc, const double2* __restrict__ drift, const double2* __restrict__ om_bar,
                                const double2* __restrict__ rho_bar, const unsigned long long* __restrict__ key, int M,
                                double2* __restrict__ om)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const unsigned long long seed = key[0], step = key[1];
    for (int l = 0; l < mc.nO; ++l) {
        const size_t o = static_cast<size_t>(l) * M + i;
        const double2 ub = mode_drift(mc, l, om_bar, rho_bar, M, i);
        double2 w = cadd(om[o], cscale(drift[o], -0.5 * mc.lam_dt[l]));
        w = cadd(w, cscale(ub, -0.5 * mc.lam_dt[l]));
        w.x += 1.0 * (mc.sigma[l] * cl_normal(seed, step, static_cast<uint32_t>(l), static_cast<uint32_t>(i)));
        om[o] = w;
    }
}

// EM / semi-implicit increment: r = -dt lam u(omega, rho) + eta.  EM adds it to omega
// directly (add = 1); the semi-implicit scheme filters r in k-space first (add = 0).
__global__ void k_omega_increment(ModeCoefs mc, const double2* __restrict__ rho, const unsigned long long* __restrict__ key, int M,
                                  int add, double2* __restrict__ om, double2* __restrict__ r_out, int with_noise)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const unsigned long long seed = key[0], step = key[1];
    for (int l = 0; l < mc.nO; ++l) {
        const size_t o = static_cast<size_t>(l) * M + i;
        const double2 u = mode_drift(mc, l, om, rho, M, i);
        double2 r = cscale(u, -mc.lam_dt[l]);
        // without noise: the semi-implicit step adds its noise in k-space
        if (with_noise) r.x += mc.sigma[l] * cl_normal(seed, step, static_cast<uint32_t>(l), static_cast<uint32_t>(i));
        if (add) om[o] = cadd(om[o], r);
        else r_out[o] = r;
    }
}

// Drift statistics (complexLangevin.cu:318-324; common.h:725-751, 1020-1050): per-block
// max over sites of |s|^2, s = u0 + u(om, rho) of mode l (u0 the stored predictor drift, or
// none), a non-finite |s|^2 counting as +inf.  Maxima do not depend on the order of the sites.
__global__ void __launch_bounds__(RED_THREADS) k_mode_drift_max2(ModeCoefs mc, int l, const double2* __restrict__ u0, const double2* __restrict__ om,
                                                                 const double2* __restrict__ rho, int M, double2* __restrict__ partials)
{
    __shared__ double sh[RED_THREADS];
    double m = 0.0;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        double2 s = mode_drift(mc, l, om, rho, M, i);
        if (u0) s = cadd(u0[static_cast<size_t>(l) * M + i], s);
        double v = s.x * s.x + s.y * s.y;
        if (!isfinite(v)) v = INFINITY;
        m = fmax(m, v);
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = d2(sh[0], 0.0);
}

// the same for s = a + b (b may be null): the psi drifts
__global__ void __launch_bounds__(RED_THREADS) k_sum_max2(const double2* __restrict__ a, const double2* __restrict__ b, int M,
                                                          double2* __restrict__ partials)
{
    __shared__ double sh[RED_THREADS];
    double m = 0.0;
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        const double2 s = b ? cadd(a[i], b[i]) : a[i];
        double v = s.x * s.x + s.y * s.y;
        if (!isfinite(v)) v = INFINITY;
        m = fmax(m, v);
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = d2(sh[0], 0.0);
}

// out->x = |scale| sqrt(max over the RED_BLOCKS partials)
__global__ void k_finalize_max(const double2* __restrict__ partials, double scale, double2* __restrict__ out)
{
    __shared__ double sh[RED_THREADS];
    sh[threadIdx.x] = partials[threadIdx.x].x;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = d2(fabs(scale) * sqrt(fmax(0.0, sh[0])), 0.0);
}

// u of interaction mode 0 (the force the learned kernel is judged against)
__global__ void k_mode0_drift(ModeCoefs mc, const double2* __restrict__ om, const double2* __restrict__ rho, int M, double2* __restrict__ u)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    u[i] = mode_drift(mc, 0, om, rho, M, i);
}

// psi update: out = base + c1 d1 + c2 d2 + (sigma n_i - eta_mean) on the real part
__global__ void k_psi_update(const double2* __restrict__ base, const double2* __restrict__ d1, double c1, const double2* __restrict__ dd2,
                             double c2, double sigma, const double2* __restrict__ eta_mean, const unsigned long long* __restrict__ key,
                             int M, double2* __restrict__ out)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const unsigned long long seed = key[0], step = key[1];
    double2 v = cadd(base[i], cscale(d1[i], c1));
    if (dd2) v = cadd(v, cscale(dd2[i], c2));
    const double eta = sigma * cl_normal(seed, step, 0x100u, static_cast<uint32_t>(i)) + (-eta_mean[0].x);
    v.x += 1.0 * eta;
    out[i] = v;
}

// psi increment for EM / SI: r = -dt lamPsi dpsi + eta
__global__ void k_psi_increment(const double2* __restrict__ dpsi, double c1, double sigma, const double2* __restrict__ eta_mean,
                                const unsigned long long* __restrict__ key, int M, double2* __restrict__ r, int with_noise)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const unsigned long long seed = key[0], step = key[1];
    double2 v = cscale(dpsi[i], c1);
    // without noise: the semi-implicit step adds its (mean-free) noise in k-space
    if (with_noise) v.x += sigma * cl_normal(seed, step, 0x100u, static_cast<uint32_t>(i)) + (-eta_mean[0].x);
    r[i] = v;
}

// out = a + b
__global__ void k_add_out(const double2* __restrict__ a, const double2* __restrict__ b, size_t n, double2* __restrict__ out)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * EW_THREADS + threadIdx.x;
    if (i < n) out[i] = cadd(a[i], b[i]);
}

// a = (a + b) + c
__global__ void k_axpy2(double2* __restrict__ a, const double2* __restrict__ b, const double2* __restrict__ c, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * EW_THREADS + threadIdx.x;
    if (i < n) a[i] = cadd(cadd(a[i], b[i]), c[i]);
}

// r = (r + c u) / 2: the trapezoid of the semi-implicit predictor-corrector, r holding
// -dt lam u(w) and u = u(w_bar), c = -dt lam
__global__ void k_midpoint(const double2* __restrict__ u, double c, size_t n, double2* __restrict__ r)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * EW_THREADS + threadIdx.x;
    if (i < n) r[i] = cscale(cadd(r[i], cscale(u[i], c)), 0.5);
}

// the same for every omega mode, u(w_bar) from the densities at the predictor
__global__ void k_omega_midpoint(ModeCoefs mc, const double2* __restrict__ om_bar, const double2* __restrict__ rho_bar, int M,
                                 double2* __restrict__ r)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    for (int l = 0; l < mc.nO; ++l) {
        const size_t o = static_cast<size_t>(l) * M + i;
        const double2 ub = mode_drift(mc, l, om_bar, rho_bar, M, i);
        r[o] = cscale(cadd(r[o], cscale(ub, -mc.lam_dt[l])), 0.5);
    }
}

// a += b (count fields)
__global__ void k_axpy1(double2* __restrict__ a, const double2* __restrict__ b, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * EW_THREADS + threadIdx.x;
    if (i < n) a[i] = cadd(a[i], b[i]);
}

// flag[0] = 0 if any value is not finite
__global__ void k_set_flag(int* __restrict__ flag, int v)
{
    if (threadIdx.x == 0) flag[0] = v;
}

__global__ void k_finite(const double2* __restrict__ a, size_t n, int* __restrict__ flag)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * EW_THREADS + threadIdx.x;
    if (i < n) {
        const double2 v = a[i];
        if (!isfinite(v.x) || !isfinite(v.y)) flag[0] = 0;
    }
}

template <typename T>
__global__ void k_convert(const double2* __restrict__ src, V2<T>* __restrict__ dst, int M)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i < M) dst[i] = cmk<V2<T>>(static_cast<T>(src[i].x), static_cast<T>(src[i].y));
}

} // namespace fts
