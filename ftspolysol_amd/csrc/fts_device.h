// fts_device.h -- device building blocks for the gfx950 engine: complex
// arithmetic over double2/float2, radix-2/3/4/6/8 butterflies, an LDS-resident
// mixed-radix Stockham line FFT, and Philox4x32-10 noise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fts {

template <typename T> struct V2t;
template <> struct V2t<double> { using type = double2; };
template <> struct V2t<float> { using type = float2; };
template <typename T> using V2 = typename V2t<T>::type;

template <typename V> __device__ __forceinline__ V cmk(decltype(V::x) x, decltype(V::x) y) { V r; r.x = x; r.y = y; return r; }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) { return cmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename V> __device__ __forceinline__ V cscale(V a, decltype(V::x) s) { return cmk<V>(s * a.x, s * a.y); }
// multiply by i*s : (x + iy) * i s = -s y + i s x
template <typename V, int S> __device__ __forceinline__ V cmul_is(V a) { return S > 0 ? cmk<V>(-a.y, a.x) : cmk<V>(a.y, -a.x); }

__device__ __forceinline__ double2 d2(double x, double y) { double2 r; r.x = x; r.y = y; return r; }

// Smith-scaled complex inverse (the formula of common.h:127-139)
__device__ __forceinline__ double2 cinv(double2 z)
{
    const double ax = fabs(z.x), ay = fabs(z.y);
    if (ax >= ay) {
        const double r = z.y / z.x;
        const double d = z.x * (1.0 + r * r);
        return d2(1.0 / d, -r / d);
    }
    const double r = z.x / z.y;
    const double d = z.y * (1.0 + r * r);
    return d2(r / d, -1.0 / d);
}

__device__ __forceinline__ double2 cexp(double2 z)
{
    const double r = exp(z.x);
    double s, c;
    sincos(z.y, &s, &c);
    return d2(r * c, r * s);
}

// ---------------------------------------------------------------------------
// Compile-time FFT plans: radix sequence and elements per thread E (a common
// multiple of the radices, so every stage gives a thread E/R whole butterflies).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int plan_e(int n)
{
    return n == 8 ? 8 : n == 12 ? 12 : n == 16 ? 4 : n == 24 ? 12 : n == 32 ? 8 : n == 48 ? 12 : n == 64 ? 8
         : n == 96 ? 12 : n == 128 ? 8 : n == 192 ? 12 : n == 256 ? 8 : 0;
}
__host__ __device__ constexpr int plan_nst(int n)
{
    return n == 8 ? 1 : n == 12 ? 2 : n == 16 ? 2 : n == 24 ? 2 : n == 32 ? 2 : n == 48 ? 3 : n == 64 ? 2
         : n == 96 ? 3 : n == 128 ? 3 : n == 192 ? 4 : n == 256 ? 3 : 0;
}
__host__ __device__ constexpr int plan_radix(int n, int s)
{
    return n == 8 ? 8
         : n == 12 ? (s == 0 ? 4 : 3)
         : n == 16 ? 4
         : n == 24 ? (s == 0 ? 4 : 6)
         : n == 32 ? (s == 0 ? 8 : 4)
         : n == 48 ? (s < 2 ? 4 : 3)
         : n == 64 ? 8
         : n == 96 ? (s < 2 ? 4 : 6)
         : n == 128 ? (s == 0 ? 8 : 4)
         : n == 192 ? (s < 3 ? 4 : 3)
         : n == 256 ? (s < 2 ? 8 : 4) : 0;
}
__host__ __device__ constexpr bool fft_length_supported(int n) { return plan_e(n) != 0; }

// ---------------------------------------------------------------------------
// Radix-R DFT in registers, X_m = sum_j u_j e^{SIGN 2 pi i j m / R}.
// ---------------------------------------------------------------------------
template <typename V, int R, int SIGN> struct Dft;

template <typename V, int SIGN> struct Dft<V, 2, SIGN> {
    __device__ __forceinline__ static void run(V* u)
    {
        const V a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    }
};

template <typename V, int SIGN> struct Dft<V, 3, SIGN> {
    __device__ __forceinline__ static void run(V* u)
    {
        using T = decltype(V::x);
        const T h = T(0.86602540378443864676372317075294);
        const V a = u[0];
        const V t1 = cadd(u[1], u[2]);
        const V t2 = csub(u[1], u[2]);
        const V m = csub(a, cscale(t1, T(0.5)));
        const V r = cscale(cmul_is<V, SIGN>(t2), h); // SIGN * i * (sqrt3/2) * t2
        u[0] = cadd(a, t1);
        u[1] = cadd(m, r);
        u[2] = csub(m, r);
    }
};

template <typename V, int SIGN> struct Dft<V, 4, SIGN> {
    __device__ __forceinline__ static void run(V* u)
    {
        const V a = cadd(u[0], u[2]);
        const V b = csub(u[0], u[2]);
        const V c = cadd(u[1], u[3]);
        const V d = cmul_is<V, SIGN>(csub(u[1], u[3]));
        u[0] = cadd(a, c);
        u[2] = csub(a, c);
        u[1] = cadd(b, d);
        u[3] = csub(b, d);
    }
};

template <typename V, int SIGN> struct Dft<V, 6, SIGN> {
    __device__ __forceinline__ static void run(V* u)
    {
        using T = decltype(V::x);
        V A[3] = {u[0], u[2], u[4]};
        V B[3] = {u[1], u[3], u[5]};
        Dft<V, 3, SIGN>::run(A);
        Dft<V, 3, SIGN>::run(B);
        const T h = T(0.86602540378443864676372317075294);
        // w6^1 = 1/2 + SIGN i sqrt3/2, w6^2 = -1/2 + SIGN i sqrt3/2
        const V w1 = cmk<V>(T(0.5), T(SIGN) * h);
        const V w2 = cmk<V>(T(-0.5), T(SIGN) * h);
        const V b1 = cmul(B[1], w1);
        const V b2 = cmul(B[2], w2);
        u[0] = cadd(A[0], B[0]);
        u[3] = csub(A[0], B[0]);
        u[1] = cadd(A[1], b1);
        u[4] = csub(A[1], b1);
        u[2] = cadd(A[2], b2);
        u[5] = csub(A[2], b2);
    }
};

template <typename V, int SIGN> struct Dft<V, 8, SIGN> {
    __device__ __forceinline__ static void run(V* u)
    {
        using T = decltype(V::x);
        V E[4] = {u[0], u[2], u[4], u[6]};
        V O[4] = {u[1], u[3], u[5], u[7]};
        Dft<V, 4, SIGN>::run(E);
        Dft<V, 4, SIGN>::run(O);
        const T s = T(0.70710678118654752440084436210485);
        // w = e^{SIGN 2 pi i / 8}
        const V o1 = cmul(O[1], cmk<V>(s, T(SIGN) * s));
        const V o2 = cmul_is<V, SIGN>(O[2]);
        const V o3 = cmul(O[3], cmk<V>(-s, T(SIGN) * s));
        u[0] = cadd(E[0], O[0]);
        u[4] = csub(E[0], O[0]);
        u[1] = cadd(E[1], o1);
        u[5] = csub(E[1], o1);
        u[2] = cadd(E[2], o2);
        u[6] = csub(E[2], o2);
        u[3] = cadd(E[3], o3);
        u[7] = csub(E[3], o3);
    }
};

// ---------------------------------------------------------------------------
// LDS Stockham line FFT.
//   NL lines of length N; element j of line l is lds[l*LS + j*ES].
//   Work item w in [0, NL*P), P = N/E threads per line; LINE_FAST maps
//   line = w % NL (use for strided lines, ES > 1) else q = w % P.
//   NT threads take part; every thread of the block must call this.
//   tw: e^{-2 pi i t / N}, t in [0, N) (forward); conjugated for SIGN = +1.
// One barrier separates the reads of a stage from its writes, so the transform
// runs in place in a single LDS buffer: each thread keeps its E values of the
// stage in registers across the barrier.
// ---------------------------------------------------------------------------
template <typename T, int N, int SIGN, int NL, int LS, int ES, bool LINE_FAST, int NT, int ST, int PC>
__device__ __forceinline__ void lds_fft_stage(V2<T>* lds, const V2<T>* tw, int tid)
{
    if constexpr (ST < plan_nst(N)) {
        using V = V2<T>;
        constexpr int R = plan_radix(N, ST);
        constexpr int E = plan_e(N);
        constexpr int P = N / E;
        constexpr int NW = NL * P;
        constexpr int ITEMS = (NW + NT - 1) / NT;
        constexpr int BPT = E / R;
        constexpr int NR = N / R;
        constexpr int TWS = N / (PC * R);
        V v[ITEMS][E];
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            const int w = tid + it * NT;
            if (w < NW) {
                const int line = LINE_FAST ? (w % NL) : (w / P);
                const int q = LINE_FAST ? (w / NL) : (w % P);
                const V* base = lds + line * LS;
#pragma unroll
                for (int m = 0; m < BPT; ++m) {
                    const int i = q + m * P;
#pragma unroll
                    for (int j = 0; j < R; ++j) v[it][m * R + j] = base[(i + j * NR) * ES];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            const int w = tid + it * NT;
            if (w < NW) {
                const int line = LINE_FAST ? (w % NL) : (w / P);
                const int q = LINE_FAST ? (w / NL) : (w % P);
                V* base = lds + line * LS;
#pragma unroll
                for (int m = 0; m < BPT; ++m) {
                    const int i = q + m * P;
                    const int k = i % PC;
                    V* u = &v[it][m * R];
                    if constexpr (PC > 1) {
#pragma unroll
                        for (int j = 1; j < R; ++j) {
                            V t = tw[j * k * TWS];
                            if (SIGN > 0) t.y = -t.y;
                            u[j] = cmul(u[j], t);
                        }
                    }
                    Dft<V, R, SIGN>::run(u);
                    const int ob = (i - k) * R + k;
#pragma unroll
                    for (int j = 0; j < R; ++j) base[(ob + j * PC) * ES] = u[j];
                }
            }
        }
        __syncthreads();
        lds_fft_stage<T, N, SIGN, NL, LS, ES, LINE_FAST, NT, ST + 1, PC * R>(lds, tw, tid);
    }
}

template <typename T, int N, int SIGN, int NL, int LS, int ES, bool LINE_FAST, int NT>
__device__ __forceinline__ void lds_fft(V2<T>* lds, const V2<T>* tw, int tid)
{
    lds_fft_stage<T, N, SIGN, NL, LS, ES, LINE_FAST, NT, 0, 1>(lds, tw, tid);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and the CL noise draw.
//   counter = {site, field, step_lo, step_hi}, key = {seed_lo, seed_hi}
//   u1, u2 = 53-bit uniforms in (0,1) from (r0,r1) and (r2,r3); n = sqrt(-2 ln u1) cos(2 pi u2)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double cl_normal(uint64_t seed, uint64_t step, uint32_t field, uint32_t site)
{
    uint32_t c[4] = {site, field, (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    const uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    const double u1 = ((double)a + 0.5) * 0x1.0p-53;
    const double u2 = ((double)b + 0.5) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// both Box-Muller normals of one draw: cl_normal's value (cos) and its sine partner;
// F32: the logarithm, root and sine/cosine in FP32 (for FP32 transforms)
template <bool F32 = false>
__device__ __forceinline__ double2 cl_normal_pair(uint64_t seed, uint64_t step, uint32_t field, uint32_t site)
{
    uint32_t c[4] = {site, field, (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    const uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    const double u1 = ((double)a + 0.5) * 0x1.0p-53;
    const double u2 = ((double)b + 0.5) * 0x1.0p-53;
    if constexpr (F32) {
        const float r = sqrtf(-2.0f * logf(static_cast<float>(u1)));
        float sn, cs;
        sincospif(2.0f * static_cast<float>(u2), &sn, &cs);
        return make_double2(r * cs, r * sn);
    }
    const double r = sqrt(-2.0 * log(u1));
    return make_double2(r * cos(6.283185307179586 * u2), r * sin(6.283185307179586 * u2));
}

} // namespace fts
