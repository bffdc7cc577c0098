// lrk.hip -- the learned linear-response kernel on gfx950: the network's forward
// and backward passes, the Adam trainer, and the per-step kernels that judge the
// kernel with the exact force (FTS_SCHEME_SI_LEARNED; formulas in include/fts.h).
// No reference counterpart: the reference has no semi-implicit scheme.
//
// The network is small (7 -> 64 -> 64 -> 64 -> 2, tanh, 9042 FP32 parameters), so
// a workgroup of 256 threads holds all of it in LDS and pushes tiles of 32 points
// (32 |k| bins of one record) through it:
//   forward   three 32x64 GEMM-shaped layers, each thread a 1x8 strip of outputs,
//             weights read as float4 rows of a transposed LDS copy;
//   backward  the deltas (the same strip shape against the natural layout) and the
//             weight gradients dW = D^T H, each thread owning 16 entries of each
//             64x64 layer and accumulating them in registers over all its tiles;
//   reduce    per-workgroup gradient partials summed in a fixed order (FP64) and,
//             for training, Adam on the same pass.
// Every sum has a fixed order, so gradients and training runs are reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/fts.h"
#include "engine.h"
#include "lrk.h"

namespace fts {
namespace {

constexpr int NP = FTS_LRK_NPARAMS;
constexpr int HD = FTS_LRK_HIDDEN;
constexpr int NB = FTS_LRK_BINS;
constexpr int TP = 32;  // points per tile
constexpr int LS = 68;  // LDS row stride of a tile's activations (floats; keeps float4 rows aligned)
constexpr int NT = 256; // threads per workgroup
constexpr int OS = FTS_LRK_OFF_SHIFT, OC = FTS_LRK_OFF_SCALE;
constexpr int OW1 = FTS_LRK_OFF_W1, OB1 = FTS_LRK_OFF_B1, OW2 = FTS_LRK_OFF_W2, OB2 = FTS_LRK_OFF_B2;
constexpr int OW3 = FTS_LRK_OFF_W3, OB3 = FTS_LRK_OFF_B3, OW4 = FTS_LRK_OFF_W4, OB4 = FTS_LRK_OFF_B4;
constexpr int MAX_GRAD_BLOCKS = 512;
static_assert(FTS_LRK_IN == 8 && HD == 64 && FTS_LRK_OUT == 2 && NB % TP == 0, "tile shapes assume the documented network");
static_assert(OW2 % 4 == 0 && OW3 % 4 == 0, "float4 rows of W2 and W3");

struct Smem {
    float P[NP + 2];    // parameters, natural layout
    float WT1[8 * HD];  // W1 transposed: [in][out]
    float WT2[HD * HD]; // W2 transposed
    float WT3[HD * HD]; // W3 transposed
    float X[TP * 8];
    float H1[TP * LS], H2[TP * LS], H3[TP * LS];
    float D1[TP * LS], D2[TP * LS], D3[TP * LS];
    float R[TP * 2];
    float DZ[TP * 2];
};

__device__ void load_params(Smem& s, const float* __restrict__ params)
{
    const int t = threadIdx.x;
    for (int j = t; j < NP; j += NT) s.P[j] = params[j];
    __syncthreads();
    for (int e = t; e < HD * HD; e += NT) {
        const int o = e / HD, i = e % HD;
        s.WT2[i * HD + o] = s.P[OW2 + e];
        s.WT3[i * HD + o] = s.P[OW3 + e];
    }
    for (int e = t; e < HD * 8; e += NT) {
        const int o = e / 8, i = e % 8;
        s.WT1[i * HD + o] = s.P[OW1 + e];
    }
    __syncthreads();
}

// out[p][j] = tanh(b[j] + sum_i in[p][i] W[j][i]); thread: point p = t / 8, outputs 8 (t % 8) .. + 8
template <int K>
__device__ __forceinline__ void fwd_layer(const float* __restrict__ WT, const float* __restrict__ b, const float* __restrict__ in,
                                          int in_stride, float* __restrict__ out)
{
    const int t = threadIdx.x, p = t >> 3, j0 = (t & 7) * 8;
    float acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = b[j0 + q];
#pragma unroll 8
    for (int i = 0; i < K; ++i) {
        const float a = in[p * in_stride + i];
        const float4 w0 = *reinterpret_cast<const float4*>(WT + i * HD + j0);
        const float4 w1 = *reinterpret_cast<const float4*>(WT + i * HD + j0 + 4);
        acc[0] += a * w0.x;
        acc[1] += a * w0.y;
        acc[2] += a * w0.z;
        acc[3] += a * w0.w;
        acc[4] += a * w1.x;
        acc[5] += a * w1.y;
        acc[6] += a * w1.z;
        acc[7] += a * w1.w;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) out[p * LS + j0 + q] = tanhf(acc[q]);
}

// Forward pass of the tile in s.X: hidden activations in H1..H3, outputs r in s.R.
__device__ void forward_tile(Smem& s)
{
    fwd_layer<8>(s.WT1, s.P + OB1, s.X, 8, s.H1);
    __syncthreads();
    fwd_layer<HD>(s.WT2, s.P + OB2, s.H1, LS, s.H2);
    __syncthreads();
    fwd_layer<HD>(s.WT3, s.P + OB3, s.H2, LS, s.H3);
    __syncthreads();
    const int t = threadIdx.x;
    if (t < TP * 2) {
        const int p = t >> 1, c = t & 1;
        float z = s.P[OB4 + c];
        for (int i = 0; i < HD; ++i) z += s.P[OW4 + c * HD + i] * s.H3[p * LS + i];
        s.R[p * 2 + c] = 4.0f * tanhf(0.25f * z);
    }
    __syncthreads();
}

// Features of point p of a tile: bin b, moments m[6]; x = (f - shift) * scale.
__device__ __forceinline__ void set_features(Smem& s, int p, int b, const float* m)
{
    for (int j = 0; j < 8; ++j) {
        const float f = j == 0 ? (static_cast<float>(b) + 0.5f) / static_cast<float>(NB) : (j < 7 ? m[j - 1] : 0.0f);
        s.X[p * 8 + j] = j < 7 ? (f - s.P[OS + j]) * s.P[OC + j] : 0.0f;
    }
}

struct GradAcc {
    float w2[16], w3[16], w1[2], w4, b1, b2, b3, b4;
};

// Backward pass of the tile (dL/dz in s.DZ), accumulating into the thread's entries.
__device__ void backward_tile(Smem& s, GradAcc& g)
{
    const int t = threadIdx.x;
    if (t < 2 * HD) {
        const int c = t >> 6, i = t & 63;
        float a = 0.0f;
        for (int p = 0; p < TP; ++p) a += s.DZ[p * 2 + c] * s.H3[p * LS + i];
        g.w4 += a;
    }
    if (t < 2) {
        float a = 0.0f;
        for (int p = 0; p < TP; ++p) a += s.DZ[p * 2 + t];
        g.b4 += a;
    }
    {
        const int p = t >> 3, i0 = (t & 7) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = i0 + q;
            const float d = s.P[OW4 + i] * s.DZ[p * 2] + s.P[OW4 + HD + i] * s.DZ[p * 2 + 1];
            const float h = s.H3[p * LS + i];
            s.D3[p * LS + i] = d * (1.0f - h * h);
        }
    }
    __syncthreads();
    // layer 3 (weights W3: D3 against H2), then layer 2 (W2: D2 against H1)
    auto hidden = [&](const float* D, const float* Hin, int offW, float (&gw)[16], float& gb, float* Dnext) {
        {
            const int j = t >> 2, i0 = (t & 3) * 16;
            for (int p = 0; p < TP; ++p) {
                const float d = D[p * LS + j];
                const float4* h = reinterpret_cast<const float4*>(Hin + p * LS + i0);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float4 x = h[v];
                    gw[4 * v + 0] += d * x.x;
                    gw[4 * v + 1] += d * x.y;
                    gw[4 * v + 2] += d * x.z;
                    gw[4 * v + 3] += d * x.w;
                }
            }
        }
        if (t < HD) {
            float a = 0.0f;
            for (int p = 0; p < TP; ++p) a += D[p * LS + t];
            gb += a;
        }
        {
            const int p = t >> 3, i0 = (t & 7) * 8;
            float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 8
            for (int j = 0; j < HD; ++j) {
                const float d = D[p * LS + j];
                const float4 w0 = *reinterpret_cast<const float4*>(s.P + offW + j * HD + i0);
                const float4 w1 = *reinterpret_cast<const float4*>(s.P + offW + j * HD + i0 + 4);
                acc[0] += d * w0.x;
                acc[1] += d * w0.y;
                acc[2] += d * w0.z;
                acc[3] += d * w0.w;
                acc[4] += d * w1.x;
                acc[5] += d * w1.y;
                acc[6] += d * w1.z;
                acc[7] += d * w1.w;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float h = Hin[p * LS + i0 + q];
                Dnext[p * LS + i0 + q] = acc[q] * (1.0f - h * h);
            }
        }
        __syncthreads();
    };
    hidden(s.D3, s.H2, OW3, g.w3, g.b3, s.D2);
    hidden(s.D2, s.H1, OW2, g.w2, g.b2, s.D1);
    {
        const int j = t >> 2, i0 = (t & 3) * 2;
        for (int p = 0; p < TP; ++p) {
            const float d = s.D1[p * LS + j];
            g.w1[0] += d * s.X[p * 8 + i0];
            g.w1[1] += d * s.X[p * 8 + i0 + 1];
        }
    }
    if (t < HD) {
        float a = 0.0f;
        for (int p = 0; p < TP; ++p) a += s.D1[p * LS + t];
        g.b1 += a;
    }
    __syncthreads();
}

__device__ double block_sum_d(double v)
{
    __shared__ double sh[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) r += sh[w];
    __syncthreads();
    return r;
}

// Loss and gradient partials over the tiles of `nb` records (record k of the batch is
// idx[k], or k if idx is NULL).  mom: [n][8] moments; Y, Wt: [n][2][NB] targets and weights.
__global__ void __launch_bounds__(NT) k_lrk_grad(const float* __restrict__ params, const float* __restrict__ mom,
                                                 const float* __restrict__ Y, const float* __restrict__ Wt, const int* __restrict__ idx,
                                                 int nb, float inv_nb, float* __restrict__ partials, double* __restrict__ loss_partials)
{
    __shared__ __align__(16) Smem s;
    load_params(s, params);
    const int t = threadIdx.x;
    GradAcc g;
    for (int q = 0; q < 16; ++q) g.w2[q] = g.w3[q] = 0.0f;
    g.w1[0] = g.w1[1] = g.w4 = g.b1 = g.b2 = g.b3 = g.b4 = 0.0f;
    double lacc = 0.0;
    constexpr int TPR = NB / TP; // tiles per record
    const int ntiles = nb * TPR;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int k = tile / TPR, q = tile % TPR;
        const int n = idx ? idx[k] : k;
        if (t < TP) {
            float m[6];
            for (int j = 0; j < 6; ++j) m[j] = mom[static_cast<size_t>(n) * 8 + j];
            set_features(s, t, q * TP + t, m);
        }
        __syncthreads();
        forward_tile(s);
        if (t < TP * 2) {
            const int p = t >> 1, c = t & 1, b = q * TP + p;
            const size_t o = (static_cast<size_t>(n) * 2 + c) * NB + b;
            const float y = Y[o], w = Wt[o], r = s.R[p * 2 + c];
            const float er = expf(r), diff = er - y;
            s.DZ[p * 2 + c] = w > 0.0f ? inv_nb * 2.0f * w * diff * er * (1.0f - 0.0625f * r * r) : 0.0f;
            if (w > 0.0f) lacc += static_cast<double>(inv_nb) * static_cast<double>(w) * static_cast<double>(diff) * static_cast<double>(diff);
        }
        __syncthreads();
        backward_tile(s, g);
    }
    float* out = partials + static_cast<size_t>(blockIdx.x) * NP;
    {
        const int j = t >> 2, i0w = (t & 3) * 16, i01 = (t & 3) * 2;
        for (int q = 0; q < 16; ++q) {
            out[OW2 + j * HD + i0w + q] = g.w2[q];
            out[OW3 + j * HD + i0w + q] = g.w3[q];
        }
        out[OW1 + j * 8 + i01] = g.w1[0];
        out[OW1 + j * 8 + i01 + 1] = g.w1[1];
    }
    if (t < 2 * HD) out[OW4 + t] = g.w4;
    if (t < HD) {
        out[OB1 + t] = g.b1;
        out[OB2 + t] = g.b2;
        out[OB3 + t] = g.b3;
    }
    if (t < 2) out[OB4 + t] = g.b4;
    if (t < 16) out[t] = 0.0f; // shift and scale are not trained
    const double l = block_sum_d(lacc);
    if (t == 0) loss_partials[blockIdx.x] = l;
}

// Sums G gradient partials per parameter (FP64, fixed order) into grad (if not NULL),
// the loss partials into loss_out[0] (+= loss_weight * loss), and with adam != 0 takes
// one Adam step on params (shift and scale excluded).
__global__ void k_lrk_reduce(const float* __restrict__ partials, int G, const double* __restrict__ loss_partials, float* __restrict__ grad,
                             double* __restrict__ loss_out, double loss_weight, int adam, float* __restrict__ params, float* __restrict__ m1,
                             float* __restrict__ m2, float lr, float beta1, float beta2, float eps, float bc1, float bc2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        double l = 0.0;
        for (int k = 0; k < G; ++k) l += loss_partials[k];
        loss_out[0] += loss_weight * l;
    }
    if (j >= NP) return;
    double a = 0.0;
    for (int k = 0; k < G; ++k) a += static_cast<double>(partials[static_cast<size_t>(k) * NP + j]);
    const float gj = static_cast<float>(a);
    if (grad) grad[j] = gj;
    if (adam && j >= 16) {
        const float mm = beta1 * m1[j] + (1.0f - beta1) * gj;
        const float vv = beta2 * m2[j] + (1.0f - beta2) * gj * gj;
        m1[j] = mm;
        m2[j] = vv;
        params[j] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
}

// Forward pass of n feature vectors [n][7] -> r [n][2].
__global__ void __launch_bounds__(NT) k_lrk_forward(const float* __restrict__ params, const float* __restrict__ feat, int n, float* __restrict__ r)
{
    __shared__ __align__(16) Smem s;
    load_params(s, params);
    const int t = threadIdx.x;
    const int ntiles = (n + TP - 1) / TP;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if (t < TP * 8) {
            const int p = t >> 3, j = t & 7, k = tile * TP + p;
            const float f = (k < n && j < 7) ? feat[static_cast<size_t>(k) * FTS_LRK_FEATURES + j] : 0.0f;
            s.X[p * 8 + j] = j < 7 ? (f - s.P[OS + j]) * s.P[OC + j] : 0.0f;
        }
        __syncthreads();
        forward_tile(s);
        if (t < TP * 2) {
            const int k = tile * TP + (t >> 1);
            if (k < n) r[static_cast<size_t>(k) * 2 + (t & 1)] = s.R[t];
        }
        __syncthreads();
    }
}

// ---- per-step kernels of the engine ------------------------------------------

__device__ __forceinline__ void block_sum4(double2 (&v)[4])
{
    __shared__ double2 sh[4][NT / 64];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (int off = 32; off > 0; off >>= 1) {
            v[k].x += __shfl_down(v[k].x, off, 64);
            v[k].y += __shfl_down(v[k].y, off, 64);
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            double2 a = sh[k][0];
            for (int w = 1; w < NT / 64; ++w) {
                a.x += sh[k][w].x;
                a.y += sh[k][w].y;
            }
            v[k] = a;
        }
}

__global__ void __launch_bounds__(RED_THREADS) k_lrk_moment_sums(const double2* __restrict__ w0, const double2* __restrict__ ps, int M,
                                                                 double2* __restrict__ partials)
{
    double2 v[4];
    for (int k = 0; k < 4; ++k) v[k] = make_double2(0.0, 0.0);
    for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < M; i += RED_BLOCKS * RED_THREADS) {
        if (w0) {
            const double2 a = w0[i];
            v[0].x += a.x;
            v[0].y += a.x * a.x;
            v[1].x += a.y;
            v[1].y += a.y * a.y;
        }
        if (ps) {
            const double2 a = ps[i];
            v[2].x += a.x;
            v[2].y += a.x * a.x;
            v[3].x += a.y;
            v[3].y += a.y * a.y;
        }
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partials[k * RED_BLOCKS + blockIdx.x] = v[k];
}

// moments from the (mean, mean square) pairs, then the network at the bin centres:
// workgroup q evaluates bins [q TP, (q + 1) TP); workgroup 0 stores the moments
__global__ void __launch_bounds__(NT) k_lrk_infer(const float* __restrict__ params, const double2* __restrict__ means, LrkDevState* st)
{
    __shared__ __align__(16) Smem s;
    __shared__ float msh[6];
    load_params(s, params);
    const int t = threadIdx.x, q = blockIdx.x;
    if (t == 0) {
        double m[6];
        auto sd = [](double2 a) { return sqrt(fmax(0.0, a.y - a.x * a.x)); };
        m[0] = sd(means[0]);
        m[1] = sd(means[1]);
        m[2] = means[0].x;
        m[3] = means[1].x;
        m[4] = sd(means[2]);
        m[5] = sd(means[3]);
        for (int j = 0; j < 6; ++j) {
            if (q == 0) st->m_cur[j] = m[j];
            msh[j] = static_cast<float>(m[j]);
        }
        if (q == 0) st->m_cur[6] = st->m_cur[7] = 0.0;
    }
    __syncthreads();
    if (t < TP) set_features(s, t, q * TP + t, msh);
    __syncthreads();
    forward_tile(s);
    if (t < TP * 2) st->r_cur[q * TP * 2 + t] = s.R[t];
}

template <typename S>
__global__ void __launch_bounds__(NT) k_lrk_bin_chunks(const int* __restrict__ chunk_start, const int* __restrict__ perm, const double* __restrict__ ad,
                                                       int n_perm, const S* __restrict__ Wc0, const S* __restrict__ Wp0,
                                                       const S* __restrict__ Uc0, const S* __restrict__ Up0,
                                                       const S* __restrict__ Wc1, const S* __restrict__ Wp1,
                                                       const S* __restrict__ Uc1, const S* __restrict__ Up1, double* __restrict__ out)
{
    double2 v[4];
    for (int k = 0; k < 4; ++k) v[k] = make_double2(0.0, 0.0);
    const int j0 = chunk_start[blockIdx.x], j1 = chunk_start[blockIdx.x + 1];
    for (int j = j0 + static_cast<int>(threadIdx.x); j < j1; j += NT) {
        const int i = perm[j];
        if (Wc0) {
            const double a = ad[j];
            const double2 dW = make_double2(static_cast<double>(Wc0[i].x) - Wp0[i].x, static_cast<double>(Wc0[i].y) - Wp0[i].y);
            const double2 dU = make_double2(static_cast<double>(Uc0[i].x) - Up0[i].x, static_cast<double>(Uc0[i].y) - Up0[i].y);
            v[0].x += a * (dU.x * dW.x + dU.y * dW.y);
            v[0].y += a * a * (dW.x * dW.x + dW.y * dW.y);
            v[1].x += dU.x * dU.x + dU.y * dU.y;
        }
        if (Wc1) {
            const double a = ad[n_perm + j];
            const double2 dW = make_double2(static_cast<double>(Wc1[i].x) - Wp1[i].x, static_cast<double>(Wc1[i].y) - Wp1[i].y);
            const double2 dU = make_double2(static_cast<double>(Uc1[i].x) - Up1[i].x, static_cast<double>(Uc1[i].y) - Up1[i].y);
            v[2].x += a * (dU.x * dW.x + dU.y * dW.y);
            v[2].y += a * a * (dW.x * dW.x + dW.y * dW.y);
            v[3].x += dU.x * dU.x + dU.y * dU.y;
        }
    }
    block_sum4(v);
    if (threadIdx.x == 0) {
        double* o = out + static_cast<size_t>(blockIdx.x) * 6;
        o[0] = v[0].x;
        o[1] = v[0].y;
        o[2] = v[1].x;
        o[3] = v[2].x;
        o[4] = v[2].y;
        o[5] = v[3].x;
    }
}

__global__ void __launch_bounds__(NT) k_lrk_judge(LrkDevState* st, int have_prev, const int* __restrict__ bin_chunk,
                                                  const double* __restrict__ chunk_sums, double threshold, int learned_scheme,
                                                  double* __restrict__ ring, int capacity)
{
    __shared__ double S[6][NB]; // S1_0 S2_0 E_0 S1_1 S2_1 E_1 per bin
    __shared__ double Esh[2];
    __shared__ int use_sh, ok_sh;
    const int t = threadIdx.x;
    // a step on non-finite fields (a batch learns of it only after the step, and
    // then restarts the replica) leaves the statistics and the records alone
    if (t == 0) ok_sh = 1;
    if (have_prev) {
        if (t < NB)
            for (int k = 0; k < 6; ++k) {
                double a = 0.0;
                for (int c = bin_chunk[t]; c < bin_chunk[t + 1]; ++c) a += chunk_sums[static_cast<size_t>(c) * 6 + k];
                S[k][t] = a;
            }
        __syncthreads();
        // per-bin terms of the residuals R(learned) and R(A_D), summed over bins in a fixed tree order
        __shared__ double red[5][NB]; // E_0, (RL - E)_0, (RD - E)_0 ... per field: E, learned, Debye
        __shared__ double tot[2][3];
        for (int c = 0; c < 2; ++c) {
            if (t < NB) {
                const double e = exp(static_cast<double>(st->r_prev[t * 2 + c]));
                red[0][t] = S[3 * c + 2][t];
                red[1][t] = -2.0 * e * S[3 * c][t] + e * e * S[3 * c + 1][t];
                red[2][t] = -2.0 * S[3 * c][t] + S[3 * c + 1][t];
            }
            __syncthreads();
            for (int off = NB / 2; off > 0; off >>= 1) {
                if (t < off)
                    for (int k = 0; k < 3; ++k) red[k][t] += red[k][t + off];
                __syncthreads();
            }
            if (t == 0)
                for (int k = 0; k < 3; ++k) tot[c][k] = red[k][0];
            __syncthreads();
        }
        if (t == 0) {
            for (int c = 0; c < 2; ++c)
                for (int k = 0; k < 3; ++k)
                    if (!isfinite(tot[c][k])) ok_sh = 0;
        }
        __syncthreads();
        if (t == 0 && !ok_sh) use_sh = 0;
        if (t == 0 && ok_sh) {
            double ex = -INFINITY;
            bool any = false;
            for (int c = 0; c < 2; ++c) {
                const double E = tot[c][0];
                Esh[c] = E;
                if (!(E > 0.0)) continue;
                const double RL = E + tot[c][1], RD = E + tot[c][2];
                st->sum_rl[c] += RL / E;
                st->sum_rd[c] += RD / E;
                ex = fmax(ex, (RL - RD) / E);
                any = true;
            }
            if (!any) ex = 0.0;
            st->judged += 1;
            st->last_excess = ex;
            use_sh = ex <= threshold ? 1 : 0;
        }
        __syncthreads();
        if (capacity > 0 && ok_sh) {
            double* rec = ring + static_cast<size_t>(st->write_idx % capacity) * FTS_LRK_RECORD;
            if (t < 8) rec[FTS_LRK_REC_MOMENTS + t] = t < 6 ? st->m_prev[t] : Esh[t - 6];
            if (t < NB)
                for (int c = 0; c < 2; ++c) {
                    rec[FTS_LRK_REC_S1 + c * NB + t] = S[3 * c][t];
                    rec[FTS_LRK_REC_S2 + c * NB + t] = S[3 * c + 1][t];
                }
            __syncthreads();
            if (t == 0) {
                st->write_idx += 1;
                st->records += 1;
            }
        }
    } else if (t == 0) {
        use_sh = 0;
    }
    __syncthreads();
    if (t == 0) {
        st->use = use_sh;
        if (learned_scheme && ok_sh) {
            if (use_sh) st->learned += 1;
            else st->fallback += 1;
        }
    }
    for (int j = t; j < 2 * NB; j += NT) st->r_prev[j] = st->r_cur[j];
    if (t < 8) st->m_prev[t] = st->m_cur[t];
}

__global__ void k_lrk_ptable(const LrkDevState* __restrict__ st, const unsigned char* __restrict__ bin, const double* __restrict__ A0,
                             const double* __restrict__ A1, double lamdt0, double lamdt1, int M, double* __restrict__ P0, double* __restrict__ P1)
{
    const int i = blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= M) return;
    const int use = st->use;
    const int b = bin[i];
    if (P0) {
        const double e = use ? exp(static_cast<double>(st->r_cur[b * 2])) : 1.0;
        P0[i] = 1.0 / (1.0 + lamdt0 * (A0[i] * e));
    }
    if (P1) {
        const double e = use ? exp(static_cast<double>(st->r_cur[b * 2 + 1])) : 1.0;
        P1[i] = 1.0 / (1.0 + lamdt1 * (A1[i] * e));
    }
}

} // namespace

hipError_t lrk_moment_sums(const double2* w0, const double2* psi, int M, double2* partials, hipStream_t s)
{
    hipLaunchKernelGGL(k_lrk_moment_sums, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, w0, psi, M, partials);
    return hipGetLastError();
}

hipError_t lrk_infer(const float* params, const double2* means, LrkDevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_lrk_infer, dim3(NB / TP), dim3(NT), 0, s, params, means, st);
    return hipGetLastError();
}

hipError_t lrk_bin_chunks(int n_chunks, const int* chunk_start, const int* perm, const double* ad_perm, int n_perm, const void* Wc0,
                          const void* Wp0, const void* Uc0, const void* Up0, const void* Wc1, const void* Wp1,
                          const void* Uc1, const void* Up1, int fp32, double* out, hipStream_t s)
{
    if (n_chunks <= 0) return hipSuccess;
    if (fp32) {
        using S = float2;
        hipLaunchKernelGGL(k_lrk_bin_chunks<S>, dim3(n_chunks), dim3(NT), 0, s, chunk_start, perm, ad_perm, n_perm, static_cast<const S*>(Wc0),
                           static_cast<const S*>(Wp0), static_cast<const S*>(Uc0), static_cast<const S*>(Up0), static_cast<const S*>(Wc1),
                           static_cast<const S*>(Wp1), static_cast<const S*>(Uc1), static_cast<const S*>(Up1), out);
    } else {
        using S = double2;
        hipLaunchKernelGGL(k_lrk_bin_chunks<S>, dim3(n_chunks), dim3(NT), 0, s, chunk_start, perm, ad_perm, n_perm, static_cast<const S*>(Wc0),
                           static_cast<const S*>(Wp0), static_cast<const S*>(Uc0), static_cast<const S*>(Up0), static_cast<const S*>(Wc1),
                           static_cast<const S*>(Wp1), static_cast<const S*>(Uc1), static_cast<const S*>(Up1), out);
    }
    return hipGetLastError();
}

hipError_t lrk_judge(LrkDevState* st, int have_prev, const int* bin_chunk, const double* chunk_sums, double threshold, int learned_scheme,
                     double* ring, int capacity, hipStream_t s)
{
    hipLaunchKernelGGL(k_lrk_judge, dim3(1), dim3(NT), 0, s, st, have_prev, bin_chunk, chunk_sums, threshold, learned_scheme, ring, capacity);
    return hipGetLastError();
}

hipError_t lrk_ptable(const LrkDevState* st, const unsigned char* bin, const double* A0, const double* A1, double lamdt0, double lamdt1,
                      int M, double* P0, double* P1, hipStream_t s)
{
    hipLaunchKernelGGL(k_lrk_ptable, dim3((M + EW_THREADS - 1) / EW_THREADS), dim3(EW_THREADS), 0, s, st, bin, A0, A1, lamdt0, lamdt1, M, P0,
                       P1);
    return hipGetLastError();
}

} // namespace fts

// ---------------------------------------------------------------------------
// C ABI: the network outside an engine
// ---------------------------------------------------------------------------
namespace {

using fts_host::fail;

template <typename X> struct DBuf {
    X* p = nullptr;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), sizeof(X) * std::max<size_t>(1, n)); }
};

// Compact training data of n records: moments [n][8], targets y = S1 / S2 and
// weights w = S2 / E ([n][2][NB]; w = 0 where E or S2 is not positive).
void compact_records(const double* rec, int n, std::vector<float>& mom, std::vector<float>& Y, std::vector<float>& W)
{
    mom.assign(static_cast<size_t>(n) * 8, 0.0f);
    Y.assign(static_cast<size_t>(n) * 2 * fts::NB, 0.0f);
    W.assign(static_cast<size_t>(n) * 2 * fts::NB, 0.0f);
    for (int k = 0; k < n; ++k) {
        const double* r = rec + static_cast<size_t>(k) * FTS_LRK_RECORD;
        for (int j = 0; j < FTS_LRK_MOMENTS; ++j) mom[static_cast<size_t>(k) * 8 + j] = static_cast<float>(r[FTS_LRK_REC_MOMENTS + j]);
        for (int c = 0; c < 2; ++c) {
            const double E = r[FTS_LRK_REC_E + c];
            for (int b = 0; b < fts::NB; ++b) {
                const double S1 = r[FTS_LRK_REC_S1 + c * fts::NB + b], S2 = r[FTS_LRK_REC_S2 + c * fts::NB + b];
                const size_t o = (static_cast<size_t>(k) * 2 + c) * fts::NB + b;
                if (E > 0.0 && S2 > 0.0 && std::isfinite(S1) && std::isfinite(S2) && std::isfinite(E)) {
                    Y[o] = static_cast<float>(S1 / S2);
                    W[o] = static_cast<float>(S2 / E);
                }
            }
        }
    }
}

int grad_blocks(int nb) { return std::max(1, std::min(fts::MAX_GRAD_BLOCKS, nb * (fts::NB / fts::TP))); }

} // namespace

struct fts_lrk_trainer {
    int n = 0;
    DBuf<float> mom, Y, W, partials, params, m1, m2;
    DBuf<int> idx;
    DBuf<double> lossp, epoch_loss;
};

extern "C" {

int fts_lrk_forward(const float* params, const float* features, int n, float* r_out)
{
    if (!params || n < 0 || (n > 0 && (!features || !r_out))) return fail(FTS_INVALID, "lrk_forward: bad argument");
    if (n == 0) return FTS_OK;
    DBuf<float> P, F, R;
    HIP_TRY(P.alloc(fts::NP));
    HIP_TRY(F.alloc(static_cast<size_t>(n) * FTS_LRK_FEATURES));
    HIP_TRY(R.alloc(static_cast<size_t>(n) * 2));
    HIP_TRY(hipMemcpy(P.p, params, sizeof(float) * fts::NP, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(F.p, features, sizeof(float) * static_cast<size_t>(n) * FTS_LRK_FEATURES, hipMemcpyHostToDevice));
    const int tiles = (n + fts::TP - 1) / fts::TP;
    hipLaunchKernelGGL(fts::k_lrk_forward, dim3(std::min(tiles, 1024)), dim3(fts::NT), 0, nullptr, P.p, F.p, n, R.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(r_out, R.p, sizeof(float) * static_cast<size_t>(n) * 2, hipMemcpyDeviceToHost));
    return FTS_OK;
}

int fts_lrk_loss_grad(const float* params, const double* records, int n_records, double* loss, float* grad)
{
    if (!params || !loss || n_records < 0 || (n_records > 0 && !records)) return fail(FTS_INVALID, "lrk_loss_grad: bad argument");
    *loss = 0.0;
    if (n_records == 0) {
        if (grad) std::fill(grad, grad + fts::NP, 0.0f);
        return FTS_OK;
    }
    std::vector<float> mom, Y, W;
    compact_records(records, n_records, mom, Y, W);
    const int G = grad_blocks(n_records);
    DBuf<float> P, dm, dY, dW, part, g;
    DBuf<double> lp, lo;
    HIP_TRY(P.alloc(fts::NP));
    HIP_TRY(dm.alloc(mom.size()));
    HIP_TRY(dY.alloc(Y.size()));
    HIP_TRY(dW.alloc(W.size()));
    HIP_TRY(part.alloc(static_cast<size_t>(G) * fts::NP));
    HIP_TRY(g.alloc(fts::NP));
    HIP_TRY(lp.alloc(static_cast<size_t>(G)));
    HIP_TRY(lo.alloc(1));
    HIP_TRY(hipMemcpy(P.p, params, sizeof(float) * fts::NP, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dm.p, mom.data(), sizeof(float) * mom.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dY.p, Y.data(), sizeof(float) * Y.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dW.p, W.data(), sizeof(float) * W.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(lo.p, 0, sizeof(double)));
    hipLaunchKernelGGL(fts::k_lrk_grad, dim3(G), dim3(fts::NT), 0, nullptr, P.p, dm.p, dY.p, dW.p, static_cast<const int*>(nullptr), n_records,
                       1.0f / static_cast<float>(n_records), part.p, lp.p);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(fts::k_lrk_reduce, dim3((fts::NP + 255) / 256), dim3(256), 0, nullptr, part.p, G, lp.p, g.p, lo.p, 1.0, 0,
                       static_cast<float*>(nullptr), static_cast<float*>(nullptr), static_cast<float*>(nullptr), 0.0f, 0.0f, 0.0f, 0.0f, 1.0f,
                       1.0f);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(loss, lo.p, sizeof(double), hipMemcpyDeviceToHost));
    if (grad) HIP_TRY(hipMemcpy(grad, g.p, sizeof(float) * fts::NP, hipMemcpyDeviceToHost));
    return FTS_OK;
}

int fts_lrk_trainer_create(const double* records, int n_records, fts_lrk_trainer** out)
{
    if (!records || n_records <= 0 || !out) return fail(FTS_INVALID, "lrk_trainer_create: bad argument");
    *out = nullptr;
    std::vector<float> mom, Y, W;
    compact_records(records, n_records, mom, Y, W);
    fts_lrk_trainer* t = new fts_lrk_trainer();
    t->n = n_records;
    auto bail = [&](int rc) {
        delete t;
        return rc;
    };
    hipError_t e;
    if ((e = t->mom.alloc(mom.size())) != hipSuccess || (e = t->Y.alloc(Y.size())) != hipSuccess || (e = t->W.alloc(W.size())) != hipSuccess ||
        (e = t->partials.alloc(static_cast<size_t>(fts::MAX_GRAD_BLOCKS) * fts::NP)) != hipSuccess || (e = t->params.alloc(fts::NP)) != hipSuccess ||
        (e = t->m1.alloc(fts::NP)) != hipSuccess || (e = t->m2.alloc(fts::NP)) != hipSuccess ||
        (e = t->idx.alloc(static_cast<size_t>(n_records))) != hipSuccess || (e = t->lossp.alloc(fts::MAX_GRAD_BLOCKS)) != hipSuccess ||
        (e = t->epoch_loss.alloc(1)) != hipSuccess)
        return bail(fail(FTS_DEVICE_ERROR, std::string("lrk_trainer_create: ") + hipGetErrorString(e)));
    if ((e = hipMemcpy(t->mom.p, mom.data(), sizeof(float) * mom.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(t->Y.p, Y.data(), sizeof(float) * Y.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(t->W.p, W.data(), sizeof(float) * W.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return bail(fail(FTS_DEVICE_ERROR, std::string("lrk_trainer_create: ") + hipGetErrorString(e)));
    *out = t;
    return FTS_OK;
}

void fts_lrk_trainer_destroy(fts_lrk_trainer* t) { delete t; }

int fts_lrk_trainer_run(fts_lrk_trainer* t, float* params, int epochs, int batch_records, double lr0, double lr1, unsigned long long seed,
                        double* loss_history)
{
    if (!t || !params || epochs < 0 || batch_records <= 0 || !(lr0 > 0.0) || !(lr1 > 0.0)) return fail(FTS_INVALID, "lrk_trainer_run: bad argument");
    const int B = std::min(batch_records, t->n);
    const int per_epoch = (t->n + B - 1) / B;
    const long long total = static_cast<long long>(per_epoch) * epochs;
    HIP_TRY(hipMemcpy(t->params.p, params, sizeof(float) * fts::NP, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(t->m1.p, 0, sizeof(float) * fts::NP));
    HIP_TRY(hipMemset(t->m2.p, 0, sizeof(float) * fts::NP));
    std::mt19937_64 rng(seed);
    std::vector<int> order(static_cast<size_t>(t->n));
    std::iota(order.begin(), order.end(), 0);
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    long long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        HIP_TRY(hipMemcpy(t->idx.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(t->epoch_loss.p, 0, sizeof(double)));
        for (int b0 = 0; b0 < t->n; b0 += B) {
            const int nb = std::min(B, t->n - b0);
            const int G = grad_blocks(nb);
            ++step;
            const double frac = total > 1 ? static_cast<double>(step - 1) / static_cast<double>(total - 1) : 0.0;
            const float lr = static_cast<float>(lr0 * std::pow(lr1 / lr0, frac));
            const float bc1 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta1), static_cast<double>(step)));
            const float bc2 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta2), static_cast<double>(step)));
            hipLaunchKernelGGL(fts::k_lrk_grad, dim3(G), dim3(fts::NT), 0, nullptr, t->params.p, t->mom.p, t->Y.p, t->W.p, t->idx.p + b0, nb,
                               1.0f / static_cast<float>(nb), t->partials.p, t->lossp.p);
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(fts::k_lrk_reduce, dim3((fts::NP + 255) / 256), dim3(256), 0, nullptr, t->partials.p, G, t->lossp.p,
                               static_cast<float*>(nullptr), t->epoch_loss.p, 1.0 / per_epoch, 1, t->params.p, t->m1.p, t->m2.p, lr, beta1, beta2,
                               eps, bc1, bc2);
            HIP_TRY(hipGetLastError());
        }
        double l = 0.0;
        HIP_TRY(hipMemcpy(&l, t->epoch_loss.p, sizeof(double), hipMemcpyDeviceToHost));
        if (loss_history) loss_history[ep] = l;
    }
    HIP_TRY(hipMemcpy(params, t->params.p, sizeof(float) * fts::NP, hipMemcpyDeviceToHost));
    return FTS_OK;
}

} // extern "C"
