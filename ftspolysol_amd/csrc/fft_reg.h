// fft_reg.h -- register-resident line FFTs.
//
// A length-N line is split N = N1 * N2 over N2 threads: thread p holds the N1
// points x[N2*j + p].  The forward transform is
//   (A) DFT_N1 over j in registers -> Y[p][k1]
//   (B) twiddle Y[p][k1] *= W_N^{p k1}
//   (C) one LDS exchange, then DFT_N2 over p in registers -> X[k1 + N1 k2]
// (Cooley-Tukey with N = N1 N2), and the inverse runs the same steps backwards,
// starting from the k-ordered values a thread holds after (C).  A forward,
// k-space multiply and inverse therefore cost two LDS exchanges, where the
// staged Stockham line FFT needs a read and a write of LDS per radix stage.
// The in-register DFTs are compile-time Cooley-Tukey trees over the radix-2/3/4/
// 6/8 butterflies with constant twiddles.
#pragma once

#include "fft.h"

namespace fts {
namespace reg {

// ---- compile-time sin / cos of 2 pi m / N --------------------------------------
// Quadrant reduction to |r| <= pi/4, then Taylor series (|error| ~ 1e-17).
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double r)
{
    double term = r, sum = r;
    for (int n = 1; n < 16; ++n) {
        term *= -r * r / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double r)
{
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 16; ++n) {
        term *= -r * r / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// returns cos (which = 0) or sin (which = 1) of 2 pi m / N
constexpr double turn(int m, int N, int which)
{
    m %= N;
    if (m < 0) m += N;
    // quadrant q = round(4 m / N), remainder angle r = 2 pi (m - q N / 4) / N
    const int q = (8 * m + N) / (2 * N);
    const double r = 2.0 * kPi * (static_cast<double>(m) - 0.25 * q * N) / N;
    const double c = taylor_cos(r), s = taylor_sin(r);
    const int qq = q % 4;
    const double cs = qq == 0 ? c : qq == 1 ? -s : qq == 2 ? -c : s;
    const double sn = qq == 0 ? s : qq == 1 ? c : qq == 2 ? -s : -c;
    return which == 0 ? cs : sn;
}

// W_N^{m} with the transform's sign: e^{SIGN 2 pi i m / N}
template <typename V, int N, int M, int SIGN> __device__ __forceinline__ V tw_const()
{
    using T = decltype(V::x);
    constexpr double c = turn(M, N, 0);
    constexpr double s = SIGN * turn(M, N, 1);
    return cmk<V>(static_cast<T>(c), static_cast<T>(s));
}

template <int N> struct Split;            // N = A * B with a direct radix A
template <> struct Split<12> { static constexpr int A = 4, B = 3; };
template <> struct Split<16> { static constexpr int A = 4, B = 4; };
template <> struct Split<24> { static constexpr int A = 8, B = 3; };
template <> struct Split<32> { static constexpr int A = 8, B = 4; };

template <int N> constexpr bool is_direct() { return N == 2 || N == 3 || N == 4 || N == 6 || N == 8; }

// In-register DFT of length N, natural order in and out.
template <typename V, int N, int SIGN> struct RegDft {
    __device__ __forceinline__ static void run(V* x)
    {
        if constexpr (is_direct<N>()) {
            Dft<V, N, SIGN>::run(x);
        } else {
            constexpr int A = Split<N>::A, B = Split<N>::B;
            // n = B n1 + n2, k = k1 + A k2
            V y[B][A];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
                for (int n1 = 0; n1 < A; ++n1) y[n2][n1] = x[B * n1 + n2];
                RegDft<V, A, SIGN>::run(y[n2]);
            }
            twiddle_all<A, B>(y, std::make_integer_sequence<int, A * B>{});
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                V z[B];
#pragma unroll
                for (int n2 = 0; n2 < B; ++n2) z[n2] = y[n2][k1];
                RegDft<V, B, SIGN>::run(z);
#pragma unroll
                for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = z[k2];
            }
        }
    }

    template <int A, int B, int... I>
    __device__ __forceinline__ static void twiddle_all(V (&y)[B][A], std::integer_sequence<int, I...>)
    {
        (twiddle_one<A, B, I>(y), ...);
    }
    template <int A, int B, int I> __device__ __forceinline__ static void twiddle_one(V (&y)[B][A])
    {
        constexpr int n2 = I / A, k1 = I % A;
        if constexpr (n2 * k1 != 0) y[n2][k1] = cmul(y[n2][k1], tw_const<V, N, n2 * k1, SIGN>());
    }
};

// register split of a line length over threads: N = N1 * N2
template <int N> struct LineSplit { static constexpr int N1 = 0, N2 = 0; };
template <> struct LineSplit<32> { static constexpr int N1 = 8, N2 = 4; };
template <> struct LineSplit<48> { static constexpr int N1 = 12, N2 = 4; };
template <> struct LineSplit<64> { static constexpr int N1 = 16, N2 = 4; };
template <> struct LineSplit<96> { static constexpr int N1 = 24, N2 = 4; };
template <> struct LineSplit<128> { static constexpr int N1 = 32, N2 = 4; };
template <> struct LineSplit<192> { static constexpr int N1 = 24, N2 = 8; };
template <> struct LineSplit<256> { static constexpr int N1 = 32, N2 = 8; };
template <int N> constexpr bool reg_line_ok() { return LineSplit<N>::N2 != 0; }

// ---- split transforms of the lines of a tile in LDS ------------------------------
// Line l, element t at lds[l * RS + t * ES]; N = N1 * N2.  Step A runs on thread
// (l, p), p < N2, over the thread's own slots N2 j + p: DFT_N1 over j in registers,
// then (forward) the twiddle W_N^{p k1}, written back to slot N2 k1 + p.  Step C
// runs on work items (l, k1), k1 < N1, over the contiguous slots N2 k1 + m: the
// forward DFT_N2 over m leaves X[k1 + N1 k2] in slot N2 k1 + k2 (digit-reversed
// order); an optional k-space multiply; the inverse DFT_N2 over k2 and the twiddle
// W_N^{-p k1} take the values back to slot N2 k1 + p, where the inverse step A (no
// twiddle) finishes the line in natural order.  Each step reads and writes only
// slots that no other thread touches in that step, so a barrier separates steps
// and none is needed inside one.
template <typename V, int N, int N2, int SIGN, int RS, int ES, bool TWID>
__device__ __forceinline__ void split_a(V* lds, const V* tw, int l, int p)
{
    constexpr int N1 = N / N2;
    V x[N1];
#pragma unroll
    for (int j = 0; j < N1; ++j) x[j] = lds[l * RS + (N2 * j + p) * ES];
    RegDft<V, N1, SIGN>::run(x);
    if constexpr (TWID) {
        if (p != 0) {
#pragma unroll
            for (int k1 = 1; k1 < N1; ++k1) {
                V t = tw[p * k1];
                if (SIGN > 0) t.y = -t.y;
                x[k1] = cmul(x[k1], t);
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[l * RS + (N2 * k1 + p) * ES] = x[k1];
}

// Step A on values already in registers: x[j] is point N2 j + p of line l (a pass
// that reads its lines straight from global memory loads them in this order, so
// the LDS staging round trip before the first step disappears).  Writes the step's
// results to LDS as split_a does.
template <typename V, int N, int N2, int SIGN, int RS, int ES, bool TWID>
__device__ __forceinline__ void split_a_regs(V (&x)[N / N2], V* lds, const V* tw, int l, int p)
{
    constexpr int N1 = N / N2;
    RegDft<V, N1, SIGN>::run(x);
    if constexpr (TWID) {
        if (p != 0) {
#pragma unroll
            for (int k1 = 1; k1 < N1; ++k1) {
                V t = tw[p * k1];
                if (SIGN > 0) t.y = -t.y;
                x[k1] = cmul(x[k1], t);
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[l * RS + (N2 * k1 + p) * ES] = x[k1];
}

// The last (inverse, untwiddled) step A of a pass, leaving its results in registers:
// on return x[k1] is point N2 k1 + p of line l in natural order, for the caller to
// store straight to global memory.
template <typename V, int N, int N2, int SIGN, int RS, int ES>
__device__ __forceinline__ void split_a_to_regs(const V* lds, V (&x)[N / N2], int l, int p)
{
    constexpr int N1 = N / N2;
#pragma unroll
    for (int j = 0; j < N1; ++j) x[j] = lds[l * RS + (N2 * j + p) * ES];
    RegDft<V, N1, SIGN>::run(x);
}

// NL lines (the tile's, N by default); mul(l, k): the k-space factor of line l at
// frequency k (FWD and INV together)
template <typename V, int N, int N2, int NT, int RS, int ES, bool FWD, bool INV, int NL = N, typename F>
__device__ __forceinline__ void split_c(V* lds, const V* tw, int tid, F mul)
{
    using T = decltype(V::x);
    constexpr int N1 = N / N2, ITEMS = NL * N1, IT = (ITEMS + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * NT;
        if (ITEMS % NT == 0 || w < ITEMS) {
            const int l = w % NL, k1 = w / NL;
            V z[N2];
#pragma unroll
            for (int m = 0; m < N2; ++m) z[m] = lds[l * RS + (N2 * k1 + m) * ES];
            if constexpr (FWD) RegDft<V, N2, -1>::run(z);
            if constexpr (FWD && INV) {
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) z[k2] = cscale(z[k2], static_cast<T>(mul(l, k1 + N1 * k2)));
            }
            if constexpr (INV) {
                RegDft<V, N2, +1>::run(z);
                if (k1 != 0) {
#pragma unroll
                    for (int pp = 1; pp < N2; ++pp) {
                        V t = tw[pp * k1];
                        t.y = -t.y;
                        z[pp] = cmul(z[pp], t);
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < N2; ++m) lds[l * RS + (N2 * k1 + m) * ES] = z[m];
        }
    }
}

struct NoMul {
    __device__ double operator()(int, int) const { return 1.0; }
};

// the plane pass's split N = (N / 8) * 8 (0: none): 8 N threads per plane
template <int N> constexpr int plane_split() { return (N % 8 == 0 && N >= 16 && 8 * N <= 1024) ? 8 : 0; }

} // namespace reg
} // namespace fts
