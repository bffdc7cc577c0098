// fft_line.h -- the line pass (see fft.h): LINE_TC lines of one axis staged in
// LDS and transformed with the Stockham line FFT.  Instantiated per precision in
// fft_line64.hip / fft_line32.hip.
#pragma once

#include "fft.h"
#include "fft_reg.h"

namespace fts {
namespace line_detail {

template <int N> constexpr int line_threads() { return LINE_TC * (N / plan_e(N)); }
// the separable schedule's passes (CONV, CONV_E) run the register-split transforms
// N = (N / 8) * 8 up to N = 96 (0: the staged Stockham transforms).  Measured on
// the 16-line tiles: at 96 the split is faster (FP32 16-job bead step 247 -> 206 us,
// FP64 one job 53 -> 44 us); at 256 it is slower (a bead step 582 us staged against
// 632 us with a 16 x 16 split and 995 us with 32 x 8)
template <int N> constexpr int line_split() { return (N % 8 == 0 && N >= 16 && N <= 96) ? 8 : 0; }
template <int N, int MODE> constexpr bool line_split_mode() { return (MODE == LINE_CONV || MODE == LINE_CONV_E) && line_split<N>() != 0; }
template <int N, int MODE> constexpr int line_threads_for() { return line_split_mode<N, MODE>() ? LINE_TC * line_split<N>() : line_threads<N>(); }

__device__ __forceinline__ size_t line_base(int axis, int j, int m2, int m3)
{
    if (axis == 0) return static_cast<size_t>(j);
    if (axis == 1) return static_cast<size_t>(j / m3) * m2 * m3 + (j % m3);
    return static_cast<size_t>(j) * m3;
}

template <int NT> __device__ __forceinline__ double block_max(double v, double* red)
{
    constexpr int NRED = NT % 64 == 0 ? NT / 64 : NT;
    if constexpr (NT % 64 == 0) {
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    } else {
        red[threadIdx.x] = v;
    }
    __syncthreads();
    double m = 0.0;
    for (int w = 0; w < NRED; ++w) m = fmax(m, red[w]);
    __syncthreads();
    return m;
}

// k-space kernel along one strided line: the two coordinates off the line axis
// are fixed per thread and folded once into `fixed` (separable part) and k_a/k_b.
struct LineKop {
    double fixed;   // product of the sep factors of the two fixed coordinates
    double kf[3];   // wave numbers (the line axis entry is filled per point)
    int axis;
};

__device__ __forceinline__ LineKop line_kop_setup(const KOp& op, const FftGeom& g, int axis, int colj)
{
    LineKop lk{};
    lk.axis = axis;
    int x = 0, y = 0, z = 0;
    if (axis == 0) { y = colj / g.m3; z = colj % g.m3; }
    else { x = colj / g.m3; z = colj % g.m3; }
    const int off[3] = {0, g.m1, g.m1 + g.m2};
    const int idx[3] = {x, y, z};
    lk.fixed = 1.0;
    for (int d = 0; d < 3; ++d) {
        if (d == axis) continue;
        if (op.sep) lk.fixed *= op.sep[off[d] + idx[d]];
        lk.kf[d] = g.kvec[off[d] + idx[d]];
    }
    return lk;
}

__device__ __forceinline__ double line_kop_value(const KOp& op, const FftGeom& g, const LineKop& lk, int t, size_t gi)
{
    const int off = lk.axis == 0 ? 0 : g.m1;
    double kv = op.sep ? op.sep[off + t] * lk.fixed : 1.0;
    if (op.table) kv *= kop_table(op, op.table[gi]);
    if (op.c1 != 0.0 || op.c2 != 0.0) {
        double k[3] = {lk.kf[0], lk.kf[1], lk.kf[2]};
        k[lk.axis] = g.kvec[off + t];
        double poly = op.c0;
        if (op.c1 != 0.0) poly += op.c1 * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        if (op.c2 != 0.0) poly += op.c2 * (k[op.ka] * k[op.kb]);
        kv *= poly;
    } else {
        kv *= op.c0;
    }
    return kv;
}

template <typename T, int N, int AXIS, int MODE>
__global__ void __launch_bounds__((line_threads_for<N, MODE>())) line_kernel(LineJob<T> j0, LineJob<T> j1, const LineJob<T>* __restrict__ table, FftGeom g, const V2<T>* __restrict__ twg)
{
    using V = V2<T>;
    constexpr int NT = line_threads_for<N, MODE>();
    constexpr bool SPLIT = line_split_mode<N, MODE>();
    constexpr int TC = LINE_TC;
    constexpr bool ZL = AXIS == 2; // z lines: LDS rows are lines ([c][t]); otherwise [t][c]
    constexpr int LS = ZL ? N + 1 : TC + 1;
    constexpr int NRED = NT % 64 == 0 ? NT / 64 : NT;
    constexpr bool PRODUCES_Q = MODE == LINE_INV_E_FWD || MODE == LINE_INV_E || MODE == LINE_E_FWD || MODE == LINE_CONV_E || MODE == LINE_E;
    __shared__ V lds[ZL ? TC * (N + 1) : N * (TC + 1)];
    __shared__ V tw[N];
    __shared__ double red[NRED];
    __shared__ T fk[SPLIT ? N : 1]; // the split path's factor of the line axis times op.c0
    const int tid = threadIdx.x;
    const LineJob<T> job = pick_job<T>(j0, j1, table);
    const int col0 = blockIdx.x * TC;
    const size_t S = AXIS == 0 ? static_cast<size_t>(g.m2) * g.m3 : AXIS == 1 ? static_cast<size_t>(g.m3) : 1;
    for (int t = tid; t < N; t += NT) {
        tw[t] = twg[t];
        if constexpr (SPLIT) fk[t] = static_cast<T>(job.op.sep[(AXIS == 0 ? 0 : AXIS == 1 ? g.m1 : g.m1 + g.m2) + t] * job.op.c0);
    }

    T c = T(1);
    if constexpr (PRODUCES_Q) {
        if (job.norm_in) {
            double m = 0.0;
            for (int k = tid; k < gridDim.x; k += NT) m = fmax(m, job.norm_in[k]);
            m = block_max<NT>(m, red);
            const double kmax = sizeof(T) == 4 ? 120.0 : 1000.0;
            double k = (m > 0.0 && isfinite(m)) ? rint(-0.5 * log2(m)) : 0.0;
            k = fmin(fmax(k, -kmax), kmax);
            c = static_cast<T>(exp2(k));
            if (blockIdx.x == 0 && tid == 0) *job.log_scale = k;
        } else if (job.log_scale && blockIdx.x == 0 && tid == 0) {
            *job.log_scale = 0.0;
        }
    }
    double nrm = 0.0;
    const V* src = (MODE == LINE_E_FWD || MODE == LINE_E) ? job.emw : (job.src ? job.src : job.data);
#pragma unroll
    for (int i = tid; i < N * TC; i += NT) {
        const int cc = ZL ? i / N : i % TC;
        const int t = ZL ? i % N : i / TC;
        const size_t gi = line_base(AXIS, col0 + cc, g.m2, g.m3) + t * S;
        V v = src[gi];
        if constexpr (MODE == LINE_E_FWD || MODE == LINE_E) {
            v = cscale(v, c);
            job.qout[gi] = v;
            nrm = fmax(nrm, static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y);
        }
        lds[ZL ? cc * LS + t : t * LS + cc] = v;
    }
    __syncthreads();

    constexpr bool CONV = MODE == LINE_CONV || MODE == LINE_CONV_E;
    constexpr bool DO_INV = MODE == LINE_INV || MODE == LINE_INV_E_FWD || MODE == LINE_INV_E;
    constexpr bool MUL = MODE == LINE_FWD_MUL_INV || MODE == LINE_FWD_MUL_INV_NOISE;
    constexpr bool DO_FWD_FIRST = MODE == LINE_FWD || MUL || (CONV && !SPLIT);
    if constexpr (SPLIT) {
        // FFT, x the line axis' factor of the separable kernel and op.c0, IFFT: split
        // step A forward, step C (forward, multiply, inverse), step A inverse; line l
        // of the tile at lds[l * RS + t * ES]
        constexpr int N2 = line_split<N>();
        constexpr int RS = ZL ? LS : 1, ES = ZL ? 1 : LS;
        const int l = tid % TC, p = tid / TC;
        reg::split_a<V, N, N2, -1, RS, ES, true>(lds, tw, l, p);
        __syncthreads();
        reg::split_c<V, N, N2, NT, RS, ES, true, true, TC>(lds, tw, tid, [&](int, int k) { return fk[k]; });
        __syncthreads();
        reg::split_a<V, N, N2, +1, RS, ES, false>(lds, tw, l, p);
        __syncthreads();
    }
    if constexpr (DO_FWD_FIRST) {
        if constexpr (ZL) lds_fft<T, N, -1, TC, LS, 1, true, NT>(lds, tw, tid);
        else lds_fft<T, N, -1, TC, 1, LS, true, NT>(lds, tw, tid);
    }
    if constexpr (CONV && !SPLIT) {
        // x the line axis' factor of the separable kernel (op.sep holds the m1, m2, m3
        // factors) and op.c0
        const int off = AXIS == 0 ? 0 : AXIS == 1 ? g.m1 : g.m1 + g.m2;
        const double* f = job.op.sep + off;
        const double c0 = job.op.c0;
#pragma unroll
        for (int i = tid; i < N * TC; i += NT) {
            const int cc = ZL ? i / N : i % TC;
            const int t = ZL ? i % N : i / TC;
            V* p = &lds[ZL ? cc * LS + t : t * LS + cc];
            *p = cscale(*p, static_cast<T>(f[t] * c0));
        }
        __syncthreads();
    }
    if constexpr (MUL) {
        static_assert(!ZL, "the k-space multiply runs on x or y lines");
        static_assert(NT % TC == 0, "line threads must cover whole rows");
        const KOp op = job.op;
        const int cc = tid % TC;
        const LineKop lk = line_kop_setup(op, g, AXIS, col0 + cc);
        const size_t cb = line_base(AXIS, col0 + cc, g.m2, g.m3);
#pragma unroll
        for (int i = tid; i < N * TC; i += NT) {
            const int t = i / TC;
            V* p = &lds[t * LS + cc];
            *p = cscale(*p, static_cast<T>(line_kop_value(op, g, lk, t, cb + t * S)));
        }
        // k-space noise (semi-implicit steps), one element at a time: the Philox rounds
        // would otherwise multiply the registers of the unrolled loop above
        if constexpr (MODE == LINE_FWD_MUL_INV_NOISE) {
#pragma unroll 1
            for (int i = tid; i < N * TC; i += NT) {
                const int t = i / TC;
                const double2 nz = kop_noise<sizeof(T) == 4>(op, g.m1, g.m2, g.m3, cb + t * S);
                lds[t * LS + cc].x += static_cast<T>(nz.x);
                lds[t * LS + cc].y += static_cast<T>(nz.y);
            }
        }
        __syncthreads();
    }
    if constexpr (DO_INV || MUL || (CONV && !SPLIT)) {
        if constexpr (ZL) lds_fft<T, N, +1, TC, LS, 1, true, NT>(lds, tw, tid);
        else lds_fft<T, N, +1, TC, 1, LS, true, NT>(lds, tw, tid);
    }
    if constexpr (MODE == LINE_INV_E_FWD || MODE == LINE_INV_E || MODE == LINE_CONV_E) {
#pragma unroll
        for (int i = tid; i < N * TC; i += NT) {
            const int cc = ZL ? i / N : i % TC;
            const int t = ZL ? i % N : i / TC;
            const size_t gi = line_base(AXIS, col0 + cc, g.m2, g.m3) + t * S;
            V* p = &lds[ZL ? cc * LS + t : t * LS + cc];
            V v = cmul(*p, job.emw[gi]);
            if (job.norm_in) v = cscale(v, c);
            *p = v;
            job.qout[gi] = v;
            nrm = fmax(nrm, static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y);
        }
        __syncthreads();
    }
    if constexpr (PRODUCES_Q) {
        if (job.norm_out) {
            const double m = block_max<NT>(nrm, red);
            if (tid == 0) job.norm_out[blockIdx.x] = m;
        }
    }
    if constexpr (MODE == LINE_INV_E_FWD || MODE == LINE_E_FWD) {
        if constexpr (ZL) lds_fft<T, N, -1, TC, LS, 1, true, NT>(lds, tw, tid);
        else lds_fft<T, N, -1, TC, 1, LS, true, NT>(lds, tw, tid);
    }
    if constexpr (MODE != LINE_INV_E && MODE != LINE_CONV_E && MODE != LINE_E) {
#pragma unroll
        for (int i = tid; i < N * TC; i += NT) {
            const int cc = ZL ? i / N : i % TC;
            const int t = ZL ? i % N : i / TC;
            const size_t gi = line_base(AXIS, col0 + cc, g.m2, g.m3) + t * S;
            job.data[gi] = lds[ZL ? cc * LS + t : t * LS + cc];
        }
    }
}


// ---- register line pass for strided lines (x, y): see fft_reg.h --------------
template <int N> constexpr int rline_threads() { return LINE_TC * reg::LineSplit<N>::N2; }

template <typename T, int N, int AXIS, int MODE>
__global__ void __launch_bounds__(rline_threads<N>()) rline_kernel(LineJob<T> j0, LineJob<T> j1, const LineJob<T>* __restrict__ table, FftGeom g, const V2<T>* __restrict__ twg)
{
    using V = V2<T>;
    constexpr int N1 = reg::LineSplit<N>::N1, N2 = reg::LineSplit<N>::N2, TC = LINE_TC, NT = TC * N2;
    constexpr int KPT = N1 / N2;      // k1 values per thread in the cross-thread step
    constexpr int ROW = N2 * TC + TC / 2; // LDS elements per k1 (padded)
    __shared__ V xch[N1 * ROW];
    __shared__ V tw[N];
    const int tid = threadIdx.x;
    const int c = tid % TC, p = tid / TC;
    const LineJob<T> job = pick_job<T>(j0, j1, table);
    const size_t S = AXIS == 0 ? static_cast<size_t>(g.m2) * g.m3 : static_cast<size_t>(g.m3);
    const size_t base = line_base(AXIS, blockIdx.x * TC + c, g.m2, g.m3);
    for (int t = tid; t < N; t += NT) tw[t] = twg[t];
    V* data = job.data;
    V X[KPT][N2];
    if constexpr (MODE == LINE_FWD || MODE == LINE_FWD_MUL_INV) {
        V x[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) x[j] = data[base + static_cast<size_t>(N2 * j + p) * S];
        reg::RegDft<V, N1, -1>::run(x);
        __syncthreads(); // twiddle table
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1)
            if (p != 0) x[k1] = cmul(x[k1], tw[p * k1]);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) xch[k1 * ROW + p * TC + c] = x[k1];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < KPT; ++r) {
            const int k1 = p * KPT + r;
            V z[N2];
#pragma unroll
            for (int pp = 0; pp < N2; ++pp) z[pp] = xch[k1 * ROW + pp * TC + c];
            reg::RegDft<V, N2, -1>::run(z);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) X[r][k2] = z[k2];
        }
    } else {
        // inverse only: the thread's k-ordered values straight from memory
#pragma unroll
        for (int r = 0; r < KPT; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) X[r][k2] = data[base + static_cast<size_t>(p * KPT + r + N1 * k2) * S];
        __syncthreads(); // twiddle table
    }
    if constexpr (MODE == LINE_FWD) {
#pragma unroll
        for (int r = 0; r < KPT; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) data[base + static_cast<size_t>(p * KPT + r + N1 * k2) * S] = X[r][k2];
        return;
    } else {
        if constexpr (MODE == LINE_FWD_MUL_INV) {
            const KOp op = job.op;
            const LineKop lk = line_kop_setup(op, g, AXIS, blockIdx.x * TC + c);
#pragma unroll
            for (int r = 0; r < KPT; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    const int t = p * KPT + r + N1 * k2;
                    const size_t gi = base + static_cast<size_t>(t) * S;
                    X[r][k2] = cscale(X[r][k2], static_cast<T>(line_kop_value(op, g, lk, t, gi)));
                }
        }
        // inverse: DFT_N2 over k2, conjugate twiddle, exchange, DFT_N1 over k1
#pragma unroll
        for (int r = 0; r < KPT; ++r) {
            const int k1 = p * KPT + r;
            reg::RegDft<V, N2, +1>::run(X[r]);
#pragma unroll
            for (int pp = 0; pp < N2; ++pp) {
                V v = X[r][pp];
                if (pp != 0 && k1 != 0) {
                    V t = tw[pp * k1];
                    t.y = -t.y;
                    v = cmul(v, t);
                }
                xch[k1 * ROW + pp * TC + c] = v;
            }
        }
        __syncthreads();
        V x[N1];
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) x[k1] = xch[k1 * ROW + p * TC + c];
        reg::RegDft<V, N1, +1>::run(x);
#pragma unroll
        for (int j = 0; j < N1; ++j) data[base + static_cast<size_t>(N2 * j + p) * S] = x[j];
    }
}

template <typename T, int N, int AXIS, int MODE>
hipError_t rline_go(const FftGeom& g, const void* tw, const LineJob<T>* jobs, const LineJob<T>* table, int njobs, hipStream_t s)
{
    const size_t ncols = AXIS == 0 ? static_cast<size_t>(g.m2) * g.m3 : static_cast<size_t>(g.m1) * g.m3;
    dim3 grid(static_cast<unsigned>(ncols / LINE_TC), njobs);
    const LineJob<T> none{};
    hipLaunchKernelGGL((rline_kernel<T, N, AXIS, MODE>), grid, dim3(rline_threads<N>()), 0, s, jobs ? jobs[0] : none,
                       jobs ? jobs[njobs > 1 ? 1 : 0] : none, table, g, static_cast<const V2<T>*>(tw));
    return hipGetLastError();
}

template <typename T, int N, int AXIS, int MODE>
hipError_t line_go(const FftGeom& g, const void* tw, const LineJob<T>* jobs, const LineJob<T>* table, int njobs, hipStream_t s)
{
    const size_t ncols = AXIS == 0 ? static_cast<size_t>(g.m2) * g.m3 : AXIS == 1 ? static_cast<size_t>(g.m1) * g.m3 : static_cast<size_t>(g.m1) * g.m2;
    dim3 grid(static_cast<unsigned>(ncols / LINE_TC), njobs);
    const LineJob<T> none{};
    hipLaunchKernelGGL((line_kernel<T, N, AXIS, MODE>), grid, dim3(line_threads_for<N, MODE>()), 0, s, jobs ? jobs[0] : none,
                       jobs ? jobs[njobs > 1 ? 1 : 0] : none, table, g, static_cast<const V2<T>*>(tw));
    return hipGetLastError();
}

template <typename T, int N>
hipError_t line_dispatch(const FftGeom& g, int axis, int mode, const void* tw, const LineJob<T>* jobs, const LineJob<T>* table, int njobs,
                         hipStream_t s)
{
    if constexpr (reg::reg_line_ok<N>()) {
        if (axis == 0) {
            if (mode == LINE_FWD) return rline_go<T, N, 0, LINE_FWD>(g, tw, jobs, table, njobs, s);
            if (mode == LINE_INV) return rline_go<T, N, 0, LINE_INV>(g, tw, jobs, table, njobs, s);
            if (mode == LINE_FWD_MUL_INV) return rline_go<T, N, 0, LINE_FWD_MUL_INV>(g, tw, jobs, table, njobs, s);
        } else if (axis == 1) {
            if (mode == LINE_FWD) return rline_go<T, N, 1, LINE_FWD>(g, tw, jobs, table, njobs, s);
            if (mode == LINE_INV) return rline_go<T, N, 1, LINE_INV>(g, tw, jobs, table, njobs, s);
        }
    }
    if (mode == LINE_CONV) {
        if (axis == 0) return line_go<T, N, 0, LINE_CONV>(g, tw, jobs, table, njobs, s);
        if (axis == 1) return line_go<T, N, 1, LINE_CONV>(g, tw, jobs, table, njobs, s);
        if (axis == 2) return line_go<T, N, 2, LINE_CONV>(g, tw, jobs, table, njobs, s);
    }
    if (mode == LINE_CONV_E && axis == 1) return line_go<T, N, 1, LINE_CONV_E>(g, tw, jobs, table, njobs, s);
    if (mode == LINE_E && axis == 1) return line_go<T, N, 1, LINE_E>(g, tw, jobs, table, njobs, s);
    if (axis == 0) {
        if (mode == LINE_FWD) return line_go<T, N, 0, LINE_FWD>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_INV) return line_go<T, N, 0, LINE_INV>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_FWD_MUL_INV) return line_go<T, N, 0, LINE_FWD_MUL_INV>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_FWD_MUL_INV_NOISE) return line_go<T, N, 0, LINE_FWD_MUL_INV_NOISE>(g, tw, jobs, table, njobs, s);
    } else if (axis == 1) {
        if (mode == LINE_FWD) return line_go<T, N, 1, LINE_FWD>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_INV) return line_go<T, N, 1, LINE_INV>(g, tw, jobs, table, njobs, s);
    } else if (axis == 2) {
        if (mode == LINE_FWD) return line_go<T, N, 2, LINE_FWD>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_INV) return line_go<T, N, 2, LINE_INV>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_INV_E_FWD) return line_go<T, N, 2, LINE_INV_E_FWD>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_INV_E) return line_go<T, N, 2, LINE_INV_E>(g, tw, jobs, table, njobs, s);
        if (mode == LINE_E_FWD) return line_go<T, N, 2, LINE_E_FWD>(g, tw, jobs, table, njobs, s);
    }
    return hipErrorInvalidValue;
}

} // namespace line_detail

template <typename T>
hipError_t line_any(const FftGeom& g, int axis, int mode, const void* tw, const LineJob<T>* jobs, const LineJob<T>* table, int njobs,
                    hipStream_t s)
{
    using namespace line_detail;
    const int n = axis == 0 ? g.m1 : axis == 1 ? g.m2 : g.m3;
    if (mode == LINE_FWD_MUL_INV && jobs && (jobs[0].op.namp != 0.0 || (njobs > 1 && jobs[1].op.namp != 0.0))) mode = LINE_FWD_MUL_INV_NOISE;
    switch (n) {
    case 8: return line_dispatch<T, 8>(g, axis, mode, tw, jobs, table, njobs, s);
    case 12: return line_dispatch<T, 12>(g, axis, mode, tw, jobs, table, njobs, s);
    case 16: return line_dispatch<T, 16>(g, axis, mode, tw, jobs, table, njobs, s);
    case 24: return line_dispatch<T, 24>(g, axis, mode, tw, jobs, table, njobs, s);
    case 32: return line_dispatch<T, 32>(g, axis, mode, tw, jobs, table, njobs, s);
    case 48: return line_dispatch<T, 48>(g, axis, mode, tw, jobs, table, njobs, s);
    case 64: return line_dispatch<T, 64>(g, axis, mode, tw, jobs, table, njobs, s);
    case 96: return line_dispatch<T, 96>(g, axis, mode, tw, jobs, table, njobs, s);
    case 128: return line_dispatch<T, 128>(g, axis, mode, tw, jobs, table, njobs, s);
    case 192: return line_dispatch<T, 192>(g, axis, mode, tw, jobs, table, njobs, s);
    case 256: return line_dispatch<T, 256>(g, axis, mode, tw, jobs, table, njobs, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_line(const FftGeom& g, int axis, int mode, const void* tw, const LineJob<T>* jobs, int njobs, hipStream_t s)
{
    return line_any<T>(g, axis, mode, tw, jobs, nullptr, njobs, s);
}
template <typename T>
hipError_t launch_line_table(const FftGeom& g, int axis, int mode, const void* tw, const LineJob<T>* dev_jobs, int njobs, hipStream_t s)
{
    return line_any<T>(g, axis, mode, tw, nullptr, dev_jobs, njobs, s);
}

} // namespace fts
