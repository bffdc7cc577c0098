// fft.hip -- plane and pencil FFT passes for gfx950 (see fft.h).
//
// Both passes hold their tile in LDS, so a pass costs one HBM read and one HBM
// write of the field.  The plane pass holds a whole m2 x m3 plane (96 x 97
// complex doubles = 149 KB of gfx950's 160 KB LDS at 96^3 FP64); the pencil pass
// holds m1 rows of 16 or 32 consecutive points of the flattened (y, z) plane, so
// each of its row reads is one contiguous segment.  The bead-step and field
// transforms run the register-split transforms of fft_reg.h, whose first and last
// steps read and write global memory directly; the other modes run the staged
// Stockham line FFT of fts_device.h.
#include "fft.h"
#include "fft_reg.h"

namespace fts {
namespace {

constexpr int PENCIL_ZC = 16;

template <int N> constexpr int plane_threads()
{
    return (N * N / plan_e(N)) > 1024 ? 1024 : (N * N / plan_e(N));
}
template <int N1, int ZC = PENCIL_ZC> constexpr int pencil_threads() { return ZC * (N1 / plan_e(N1)); }
// PLANE_CONV_E runs the register-split transforms where the plane length allows
// (8 threads per line), every other mode the staged Stockham transforms
template <int N, int MODE> constexpr bool plane_split_mode() { return MODE == PLANE_CONV_E && reg::plane_split<N>() != 0; }
// the plane passes that transform a field (forward or inverse, optionally from or to
// FP64) run the split transforms too, with the spectrum stored in natural order
template <int N, int MODE> constexpr bool plane_field_split()
{
    return (MODE == PLANE_FWD || MODE == PLANE_FWD_D2F || MODE == PLANE_INV || MODE == PLANE_INV_F2D) && reg::plane_split<N>() != 0;
}
template <int N, int MODE> constexpr int plane_threads_for()
{
    return (plane_split_mode<N, MODE>() || plane_field_split<N, MODE>()) ? reg::plane_split<N>() * N : plane_threads<N>();
}
// PENCIL_CONV likewise, with 8 threads per x line, up to N1 = 96
template <int N1, int MODE> constexpr bool pencil_split_mode()
{
    return (MODE == PENCIL_CONV || MODE == PENCIL_FWD_MUL_INV || MODE == PENCIL_FWD || MODE == PENCIL_INV || MODE == PENCIL_MUL_INV ||
            MODE == PENCIL_FWD_F2D || MODE == PENCIL_MUL_INV_D2F) &&
           N1 % 8 == 0 && N1 >= 16 && N1 <= 96;
}
template <int N1, int MODE, int ZC> constexpr int pencil_threads_for()
{
    return pencil_split_mode<N1, MODE>() ? ZC * 8 : pencil_threads<N1, ZC>();
}

// two consecutive complex FP32 values as one 16-byte access (16-byte aligned: every
// field starts at a multiple of an even point count)
template <typename V> __device__ __forceinline__ void ld_pair(const V* p, V& a, V& b)
{
    const float4 w = *reinterpret_cast<const float4*>(p);
    a = cmk<V>(w.x, w.y);
    b = cmk<V>(w.z, w.w);
}
// the same accesses marked non-temporal (streaming): the bead-step passes read the
// scratch field and q_s once and write q_{s+1} for a later pass, so these bypass the
// caches' retention and leave L2 and the Infinity Cache to the e^{-W} fields, which
// every bead step of a sweep reads again (measured: 175 -> 162 us per 16-job FP32
// launch group at 96^3)
typedef float fts_f4 __attribute__((ext_vector_type(4)));
template <typename V> __device__ __forceinline__ void ld_pair_stream(const V* p, V& a, V& b)
{
    const fts_f4 w = __builtin_nontemporal_load(reinterpret_cast<const fts_f4*>(p));
    a = cmk<V>(w.x, w.y);
    b = cmk<V>(w.z, w.w);
}
template <typename V> __device__ __forceinline__ void st_pair_stream(V* p, V a, V b)
{
    const fts_f4 w = {a.x, a.y, b.x, b.y};
    __builtin_nontemporal_store(w, reinterpret_cast<fts_f4*>(p));
}
// one complex FP64 value, the same streaming access (16 bytes)
typedef double fts_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2* p)
{
    const fts_d2 w = __builtin_nontemporal_load(reinterpret_cast<const fts_d2*>(p));
    return make_double2(w.x, w.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 v)
{
    const fts_d2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<fts_d2*>(p));
}
// one point of a bead-step input read by a split pass's first step: FP32 streams it
// (q_s and the scratch field are read once per bead step), FP64 reads it plainly
// (streaming measured slower there: FP64 pencil pass 13.4 -> 17.1 us for one system)
typedef float fts_f2 __attribute__((ext_vector_type(2)));
template <typename T> __device__ __forceinline__ V2<T> ld_point(const V2<T>* p)
{
    if constexpr (sizeof(T) == 4) {
        const fts_f2 w = __builtin_nontemporal_load(reinterpret_cast<const fts_f2*>(p));
        return make_float2(w.x, w.y);
    } else {
        return *p;
    }
}
// one point of q_{s+1} stored by the split plane pass's last step (streaming: the
// next bead step reads it once)
template <typename T> __device__ __forceinline__ void st_point(V2<T>* p, V2<T> v)
{
    if constexpr (sizeof(T) == 4) {
        const fts_f2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<fts_f2*>(p));
    } else {
        st_stream(p, v);
    }
}

// running max for the per-bead renormalisation: FP64 keeps max |q|^2; FP32 keeps
// max(|Re q|, |Im q|) in FP32 (one instruction per point) and squares it at the end.
// Either is a scale estimate within a factor 2 of max |q|^2; the scale itself is a
// power of two, which every later operation carries exactly.
template <typename T> struct NormAcc {
    double m = 0.0;
    __device__ __forceinline__ void add(V2<T> v) { m = fmax(m, static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y); }
    __device__ __forceinline__ double squared() const { return m; }
};
template <> struct NormAcc<float> {
    float m = 0.0f;
    __device__ __forceinline__ void add(float2 v) { m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y))); }
    __device__ __forceinline__ double squared() const { return static_cast<double>(m) * m; }
};

template <typename V> __device__ __forceinline__ void st_pair(V* p, V a, V b)
{
    float4 w;
    w.x = a.x;
    w.y = a.y;
    w.z = b.x;
    w.w = b.y;
    *reinterpret_cast<float4*>(p) = w;
}

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((plane_threads_for<N, MODE>())) plane_kernel(PlaneJob<T> j0, PlaneJob<T> j1, const PlaneJob<T>* __restrict__ table, const V2<T>* __restrict__ twg)
{
    using V = V2<T>;
    constexpr int NT = plane_threads_for<N, MODE>();
    constexpr bool SPLIT = plane_split_mode<N, MODE>();
    constexpr int LS = N + 1;
    __shared__ V lds[N * LS];
    __shared__ V tw[N];
    __shared__ T sfac[SPLIT ? 2 * N : 1]; // the split path's y and z factors of the bond kernel
    constexpr int NRED = NT % 64 == 0 ? NT / 64 : NT;
    __shared__ double red[NRED];
    __shared__ T scale_sh;
    const int tid = threadIdx.x;
    const PlaneJob<T> job = pick_job<T>(j0, j1, table);
    const size_t plane = static_cast<size_t>(blockIdx.x) * (N * N);
    for (int t = tid; t < N; t += NT) tw[t] = twg[t];
    if constexpr (plane_field_split<N, MODE>()) {
        // 2D transform of the plane, N = N1 N2, lanes as in the bead step: the outer
        // steps on (sr, sq) = row, point N2 j + sq of it; the middle steps on lines l.
        // After a forward step C, slot N2 k1 + k2 of a line holds frequency k1 + N1 k2,
        // so the spectrum leaves (and an inverse's enters) LDS through slot(k)
        constexpr int N2 = reg::plane_split<N>(), N1 = N / N2;
        const int sr = tid / N2, sq = tid % N2;
        const int l = tid % N, p = tid / N;
        auto slot = [](int k) { return N2 * (k % N1) + k / N1; };
        V x[N1];
        if constexpr (MODE == PLANE_FWD || MODE == PLANE_FWD_D2F) {
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const size_t gi = plane + sr * N + sq + N2 * j;
                if constexpr (MODE == PLANE_FWD_D2F) {
                    const double2 d = job.src64[gi];
                    x[j] = cmk<V>(static_cast<T>(d.x), static_cast<T>(d.y));
                } else {
                    x[j] = job.src[gi];
                }
            }
            __syncthreads(); // the twiddles
            reg::split_a_regs<V, N, N2, -1, LS, 1, true>(x, lds, tw, sr, sq); // z
            __syncthreads();
            reg::split_c<V, N, N2, NT, LS, 1, true, false>(lds, tw, tid, reg::NoMul{});
            __syncthreads();
            reg::split_a<V, N, N2, -1, 1, LS, true>(lds, tw, l, p); // y
            __syncthreads();
            reg::split_c<V, N, N2, NT, 1, LS, true, false>(lds, tw, tid, reg::NoMul{});
            __syncthreads();
#pragma unroll 4
            for (int i = tid; i < N * N; i += NT) job.dst[plane + i] = lds[slot(i / N) * LS + slot(i % N)];
        } else {
#pragma unroll 4
            for (int i = tid; i < N * N; i += NT) lds[slot(i / N) * LS + slot(i % N)] = job.src[plane + i];
            __syncthreads();
            reg::split_c<V, N, N2, NT, 1, LS, false, true>(lds, tw, tid, reg::NoMul{}); // y
            __syncthreads();
            reg::split_a<V, N, N2, +1, 1, LS, false>(lds, tw, l, p);
            __syncthreads();
            reg::split_c<V, N, N2, NT, LS, 1, false, true>(lds, tw, tid, reg::NoMul{}); // z
            __syncthreads();
            reg::split_a_to_regs<V, N, N2, +1, LS, 1>(lds, x, sr, sq);
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) {
                const size_t gi = plane + sr * N + N2 * k1 + sq;
                if constexpr (MODE == PLANE_INV_F2D) job.dst64[gi] = make_double2(static_cast<double>(x[k1].x), static_cast<double>(x[k1].y));
                else job.dst[gi] = x[k1];
            }
        }
        return;
    }
    if constexpr (SPLIT) {
        const double* sy = job.op.sep + gridDim.x; // sep holds the m1, m2, m3 factors; gridDim.x = m1
        for (int t = tid; t < 2 * N; t += NT) sfac[t] = static_cast<T>(sy[t]);
    }
    constexpr bool PRODUCES_Q = MODE == PLANE_INIT || MODE == PLANE_STEP || MODE == PLANE_LAST || MODE == PLANE_CONV_E;

    // renormalisation factor from the previous bead's per-plane norms: a max, so every
    // block gets the same value whatever the order; read by all threads at once (one
    // thread walking the m1 norms cost 11 us per 8-job launch at 96^3)
    T c = T(1);
    auto renorm = [&]() {
        if (job.norm_in) {
            double m = 0.0;
            for (int x = tid; x < gridDim.x; x += NT) m = fmax(m, job.norm_in[x]);
            if constexpr (NT % 64 == 0) {
                for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
                if ((tid & 63) == 0) red[tid >> 6] = m;
            } else {
                red[tid] = m;
            }
            __syncthreads();
            if (tid == 0) {
                double ms = 0.0;
                for (int w = 0; w < NRED; ++w) ms = fmax(ms, red[w]);
                // power-of-two scale: exact in binary floating point
                const double kmax = sizeof(T) == 4 ? 120.0 : 1000.0;
                double k = (ms > 0.0 && isfinite(ms)) ? rint(-0.5 * log2(ms)) : 0.0;
                k = fmin(fmax(k, -kmax), kmax);
                scale_sh = static_cast<T>(exp2(k));
                if (blockIdx.x == 0) *job.log_scale = k;
            }
            __syncthreads();
            c = scale_sh;
        } else if (job.log_scale && blockIdx.x == 0 && tid == 0) {
            *job.log_scale = 0.0;
        }
    };
    // INIT scales the plane as it loads it; the bead step needs the factor only after
    // the inverse transform, so it loads first and reads the norms while the loads fly
    if constexpr (MODE == PLANE_INIT) renorm();
    NormAcc<T> nrm;

    const V* src = (MODE == PLANE_INIT) ? job.emw : job.src;
    // FP32: pairs of points per 16-byte access.  Pair j = tid + k NT starts at plane
    // index 2 j; when 2 NT is a multiple of N, all pairs of a thread share one column
    // and lie 2 NT / N rows apart (COLFIX: the slot without a division per pair)
    constexpr bool PAIRS = sizeof(T) == 4 && N % 2 == 0;
    constexpr bool COLFIX = PAIRS && (2 * NT) % N == 0 && (N * N / 2) % NT == 0;
    const int slot0 = (2 * tid / N) * LS + (2 * tid) % N;
    auto pair_slot = [&](int k, int i) { return COLFIX ? slot0 + k * ((2 * NT / N) * LS) : (i / N) * LS + (i % N); };
    // the split bead step's first z step reads its rows from global memory itself: lane
    // (sr, sq) holds points z = SN2 j + sq of row sr (consecutive lanes, consecutive z)
    constexpr int SN2 = SPLIT ? reg::plane_split<N>() : 1, SN1 = N / SN2;
    const int sr = tid / SN2, sq = tid % SN2;
    V xs[SPLIT ? SN1 : 1];
    if constexpr (SPLIT) {
        const V* row = src + plane + sr * N + sq;
#pragma unroll
        for (int j = 0; j < SN1; ++j) xs[j] = ld_point<T>(row + SN2 * j);
    } else if constexpr (MODE == PLANE_FWD_D2F) {
#pragma unroll 4
        for (int i = tid; i < N * N; i += NT) {
            const double2 d = job.src64[plane + i];
            lds[(i / N) * LS + (i % N)] = cmk<V>(static_cast<T>(d.x), static_cast<T>(d.y));
        }
    } else if constexpr (PAIRS) {
#pragma unroll
        for (int k = 0, j = tid; j < N * N / 2; ++k, j += NT) {
            const int i = 2 * j;
            V v0, v1;
            if constexpr (MODE == PLANE_STEP || MODE == PLANE_LAST || MODE == PLANE_CONV_E) ld_pair_stream(src + plane + i, v0, v1);
            else ld_pair(src + plane + i, v0, v1);
            if constexpr (MODE == PLANE_INIT) {
                v0 = cscale(v0, c);
                v1 = cscale(v1, c);
                st_pair_stream(job.qout + plane + i, v0, v1);
                nrm.add(v0);
                nrm.add(v1);
            }
            V* l = &lds[pair_slot(k, i)];
            l[0] = v0;
            l[1] = v1;
        }
    } else {
#pragma unroll
        for (int i = tid; i < N * N; i += NT) {
            V v;
            if constexpr (sizeof(T) == 8 && (MODE == PLANE_STEP || MODE == PLANE_LAST || MODE == PLANE_CONV_E)) v = ld_stream(src + plane + i);
            else v = src[plane + i];
            if constexpr (MODE == PLANE_INIT) {
                v = cscale(v, c);
                job.qout[plane + i] = v;
                nrm.add(v);
            }
            lds[(i / N) * LS + (i % N)] = v;
        }
    }
    // bead step: e^{-W} read into registers now, so its latency hides behind the
    // transforms.  The FP64 plane fills the LDS (one workgroup per CU, 3 waves per SIMD),
    // which leaves the registers for it; the FP32 split pass stays at 79 VGPRs, within
    // its 6 waves per SIMD (measured 115.8 -> 108.8 us per 16-job group at 96^3)
    constexpr bool EPF = MODE == PLANE_CONV_E && (SPLIT || (sizeof(T) == 8 && (N * N) % NT == 0));
    constexpr int NPF = EPF ? (SPLIT ? SN1 : N * N / NT) : 1;
    V epf[NPF];
    if constexpr (EPF) {
#pragma unroll
        for (int k = 0; k < NPF; ++k) epf[k] = job.emw[plane + (SPLIT ? sr * N + SN2 * k + sq : tid + k * NT)];
    }
    if constexpr (MODE == PLANE_STEP || MODE == PLANE_LAST || MODE == PLANE_CONV_E) renorm();
    __syncthreads();

    if constexpr (SPLIT) {
        // the plane's share of the convolution (the pencil pass has applied phi_x / M)
        // with the split transforms, N = N1 N2: z forward on the rows leaves slot
        // N2 k1 + k2 holding kz = k1 + N1 k2, so column c holds kz = c / N2 + N1 (c % N2);
        // y forward, x phi_y phi_z and y inverse on the columns (the multiply between
        // the forward and inverse halves of step C); z inverse back to natural order
        // (the middle steps put lanes on consecutive lines, l = tid % N)
        constexpr int N2 = SN2, N1 = SN1;
        const int l = tid % N, p = tid / N;
        reg::split_a_regs<V, N, N2, -1, LS, 1, true>(xs, lds, tw, sr, sq);
        __syncthreads();
        reg::split_c<V, N, N2, NT, LS, 1, true, false>(lds, tw, tid, reg::NoMul{});
        __syncthreads();
        reg::split_a<V, N, N2, -1, 1, LS, true>(lds, tw, l, p);
        __syncthreads();
        const T* fy = sfac;
        const T* fz = sfac + N;
        reg::split_c<V, N, N2, NT, 1, LS, true, true>(lds, tw, tid, [&](int c, int ky) { return fy[ky] * fz[c / N2 + N1 * (c % N2)]; });
        __syncthreads();
        reg::split_a<V, N, N2, +1, 1, LS, false>(lds, tw, l, p);
        __syncthreads();
        reg::split_c<V, N, N2, NT, LS, 1, false, true>(lds, tw, tid, reg::NoMul{});
        __syncthreads();
        // the last z step ends in registers: x e^{-W}, x 2^k.  FP64 stores q_{s+1} from
        // there (eight lanes, one 128-byte row segment); FP32 would write 64-byte
        // segments, which the memory counts as 1.38x the bytes (PMC), so its results go
        // back to their own slots and leave as 16-byte pairs
        reg::split_a_to_regs<V, N, N2, +1, LS, 1>(lds, xs, sr, sq);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            const size_t gi = plane + sr * N + N2 * k1 + sq;
            V e;
            if constexpr (EPF) e = epf[k1];
            else e = job.emw[gi];
            V v = cmul(xs[k1], e);
            if (job.norm_in) v = cscale(v, c);
            if constexpr (PAIRS) lds[sr * LS + N2 * k1 + sq] = v;
            else st_point<T>(job.qout + gi, v);
            nrm.add(v);
        }
        if constexpr (PAIRS) {
            __syncthreads();
#pragma unroll 2
            for (int k = 0, j = tid; j < N * N / 2; ++k, j += NT) {
                const int i = 2 * j;
                const V* q = &lds[pair_slot(k, i)];
                st_pair_stream(job.qout + plane + i, q[0], q[1]);
            }
        }
    } else if constexpr (MODE == PLANE_CONV_E) {
        // the plane's share of the convolution: forward along z and y, x phi_y phi_z
        // (the pencil pass has applied phi_x / M), inverse below
        lds_fft<T, N, -1, N, LS, 1, true, NT>(lds, tw, tid);
        lds_fft<T, N, -1, N, 1, LS, true, NT>(lds, tw, tid);
        const double* sy = job.op.sep + gridDim.x; // sep holds the m1, m2, m3 factors; gridDim.x = m1
        const double* sz = sy + N;
        if constexpr (NT % N == 0) {
            // every element of this thread has z = tid % N
            const double szt = sz[tid % N];
#pragma unroll 4
            for (int i = tid; i < N * N; i += NT) {
                V* p = &lds[(i / N) * LS + (i % N)];
                *p = cscale(*p, static_cast<T>(sy[i / N] * szt));
            }
        } else {
#pragma unroll 4
            for (int i = tid; i < N * N; i += NT) {
                V* p = &lds[(i / N) * LS + (i % N)];
                *p = cscale(*p, static_cast<T>(sy[i / N] * sz[i % N]));
            }
        }
        __syncthreads();
    }
    if constexpr (MODE == PLANE_INV || MODE == PLANE_INV_F2D || MODE == PLANE_STEP || MODE == PLANE_LAST || (MODE == PLANE_CONV_E && !SPLIT)) {
        lds_fft<T, N, +1, N, LS, 1, true, NT>(lds, tw, tid); // along z
        lds_fft<T, N, +1, N, 1, LS, true, NT>(lds, tw, tid);  // along y
    }
    if constexpr (SPLIT) {
        // stored above
    } else if constexpr (PAIRS && (MODE == PLANE_STEP || MODE == PLANE_LAST || MODE == PLANE_CONV_E)) {
#pragma unroll 2
        for (int k = 0, j = tid; j < N * N / 2; ++k, j += NT) {
            const int i = 2 * j;
            V* p = &lds[pair_slot(k, i)];
            V e0, e1;
            ld_pair(job.emw + plane + i, e0, e1);
            V v0 = cmul(p[0], e0), v1 = cmul(p[1], e1);
            if (job.norm_in) {
                v0 = cscale(v0, c);
                v1 = cscale(v1, c);
            }
            if constexpr (MODE == PLANE_STEP) {
                // the forward transform below reads q_{s+1} back from LDS
                p[0] = v0;
                p[1] = v1;
            }
            st_pair_stream(job.qout + plane + i, v0, v1);
            nrm.add(v0);
            nrm.add(v1);
        }
        if constexpr (MODE == PLANE_STEP) __syncthreads();
    } else if constexpr (EPF) {
#pragma unroll
        for (int k = 0; k < NPF; ++k) {
            const int i = tid + k * NT;
            V v = cmul(lds[(i / N) * LS + (i % N)], epf[k]);
            if (job.norm_in) v = cscale(v, c);
            st_stream(job.qout + plane + i, v);
            nrm.add(v);
        }
    } else if constexpr (MODE == PLANE_STEP || MODE == PLANE_LAST || MODE == PLANE_CONV_E) {
#pragma unroll 4
        for (int i = tid; i < N * N; i += NT) {
            V* p = &lds[(i / N) * LS + (i % N)];
            V v = cmul(*p, job.emw[plane + i]);
            if (job.norm_in) v = cscale(v, c);
            *p = v;
            if constexpr (sizeof(T) == 8) st_stream(job.qout + plane + i, v);
            else job.qout[plane + i] = v;
            nrm.add(v);
        }
        __syncthreads();
    }
    if constexpr (PRODUCES_Q) {
        if (job.norm_out) {
            double n2 = nrm.squared();
            if constexpr (NT % 64 == 0) {
                for (int off = 32; off > 0; off >>= 1) n2 = fmax(n2, __shfl_down(n2, off, 64));
                if ((tid & 63) == 0) red[tid >> 6] = n2;
            } else {
                red[tid] = n2;
            }
            __syncthreads();
            if (tid == 0) {
                double s2 = 0.0;
                for (int w = 0; w < NRED; ++w) s2 = fmax(s2, red[w]);
                job.norm_out[blockIdx.x] = s2;
            }
        }
    }
    if constexpr (MODE == PLANE_FWD || MODE == PLANE_FWD_D2F || MODE == PLANE_INIT || MODE == PLANE_STEP) {
        lds_fft<T, N, -1, N, LS, 1, true, NT>(lds, tw, tid);
        lds_fft<T, N, -1, N, 1, LS, true, NT>(lds, tw, tid);
    }
    if constexpr (MODE == PLANE_INV_F2D) {
#pragma unroll 4
        for (int i = tid; i < N * N; i += NT) {
            const V v = lds[(i / N) * LS + (i % N)];
            job.dst64[plane + i] = make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
        }
    } else if constexpr (MODE != PLANE_LAST && MODE != PLANE_CONV_E) {
        if constexpr (PAIRS) {
#pragma unroll 2
            for (int j = tid; j < N * N / 2; j += NT) {
                const int i = 2 * j;
                const V* l = &lds[(i / N) * LS + (i % N)];
                st_pair(job.dst + plane + i, l[0], l[1]);
            }
        } else {
#pragma unroll 4
            for (int i = tid; i < N * N; i += NT) job.dst[plane + i] = lds[(i / N) * LS + (i % N)];
        }
    }
}

__device__ __forceinline__ double kop_value(const KOp& op, const double* kvec, int m1, int m2, int x, int y, int z, size_t idx)
{
    double v = 1.0;
    if (op.sep) v = op.sep[x] * op.sep[m1 + y] * op.sep[m1 + m2 + z];
    if (op.table) v *= kop_table(op, op.table[idx]);
    if (op.c1 != 0.0 || op.c2 != 0.0) {
        const double k[3] = {kvec[x], kvec[m1 + y], kvec[m1 + m2 + z]};
        double poly = op.c0;
        if (op.c1 != 0.0) poly += op.c1 * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        if (op.c2 != 0.0) poly += op.c2 * (k[op.ka] * k[op.kb]);
        v *= poly;
    } else {
        v *= op.c0;
    }
    return v;
}

// ZCW: (y, z) columns per workgroup; the bead-step multiplies of batches take 32
// (256-byte row segments), everything else PENCIL_ZC
template <typename T, int N1, int MODE, int ZCW>
__global__ void __launch_bounds__((pencil_threads_for<N1, MODE, ZCW>())) pencil_kernel(PencilJob<T> j0, PencilJob<T> j1, const PencilJob<T>* __restrict__ table, FftGeom g)
{
    using V = V2<T>;
    constexpr int NT = pencil_threads_for<N1, MODE, ZCW>();
    constexpr bool SPLIT = pencil_split_mode<N1, MODE>();
    constexpr int ZC = ZCW;
    constexpr int LS = ZC + 1;
    __shared__ V lds[N1 * LS];
    __shared__ V tw[N1];
    __shared__ T fx[SPLIT ? N1 : 1]; // the split path's phi_x(k) c0
    const int tid = threadIdx.x;
    const PencilJob<T> job = pick_job<T>(j0, j1, table);
    const size_t sx = static_cast<size_t>(g.m2) * g.m3;
    const size_t jbase = static_cast<size_t>(blockIdx.x) * ZC;
    const V* twg = static_cast<const V*>(g.tw_x);
    for (int t = tid; t < N1; t += NT) {
        tw[t] = twg[t];
        if constexpr (MODE == PENCIL_CONV && SPLIT) fx[t] = static_cast<T>(job.op.sep[t] * job.op.c0);
    }
    V* data = job.data;
    const V* in = job.src ? job.src : data;
    if constexpr (SPLIT) {
        // Transforms along x with the split transforms (N1 = NJ x 8), whose outer steps
        // read and write global memory themselves: lane (l, p) holds points x = 8 j + p
        // of column l (consecutive lanes, consecutive columns: each row segment of the
        // tile is one contiguous access); in LDS column l is the line at lds[l + x LS].
        // A spectrum that enters or leaves in natural order goes through slot(k): after
        // a forward step C, slot 8 k1 + k2 holds frequency k1 + NJ k2.
        constexpr int NJ = N1 / 8;
        constexpr bool FWD = MODE == PENCIL_CONV || MODE == PENCIL_FWD_MUL_INV || MODE == PENCIL_FWD || MODE == PENCIL_FWD_F2D;
        constexpr bool INV = MODE != PENCIL_FWD && MODE != PENCIL_FWD_F2D;
        const int l = tid % ZC, p = tid / ZC;
        auto slot = [](int k) { return 8 * (k % NJ) + k / NJ; };
        // the k-space kernel at frequency x of tile column lc (as the staged path)
        const KOp op = job.op;
        auto kv = [&](int lc, int x) -> double {
            const int j = static_cast<int>(jbase) + lc;
            const int y = j / g.m3, z = j % g.m3;
            double v = op.sep ? op.sep[x] * (op.sep[g.m1 + y] * op.sep[g.m1 + g.m2 + z]) : 1.0;
            if (op.table) v *= kop_table(op, op.table[static_cast<size_t>(x) * sx + j]);
            if (op.c1 != 0.0 || op.c2 != 0.0) {
                const double k[3] = {g.kvec[x], g.kvec[g.m1 + y], g.kvec[g.m1 + g.m2 + z]};
                double poly = op.c0;
                if (op.c1 != 0.0) poly += op.c1 * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
                if (op.c2 != 0.0) poly += op.c2 * (k[op.ka] * k[op.kb]);
                v *= poly;
            } else {
                v *= op.c0;
            }
            return v;
        };
        V x[NJ];
        if constexpr (FWD) {
            const V* col = in + jbase + l;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                if constexpr (MODE == PENCIL_CONV) x[j] = ld_point<T>(col + static_cast<size_t>(8 * j + p) * sx);
                else x[j] = col[static_cast<size_t>(8 * j + p) * sx];
            }
            __syncthreads(); // the twiddles and factors
            reg::split_a_regs<V, N1, 8, -1, 1, LS, true>(x, lds, tw, l, p);
            __syncthreads();
            if constexpr (MODE == PENCIL_CONV) reg::split_c<V, N1, 8, NT, 1, LS, true, true, ZC>(lds, tw, tid, [&](int, int k) { return fx[k]; });
            else if constexpr (INV) reg::split_c<V, N1, 8, NT, 1, LS, true, true, ZC>(lds, tw, tid, kv);
            else reg::split_c<V, N1, 8, NT, 1, LS, true, false, ZC>(lds, tw, tid, reg::NoMul{});
            __syncthreads();
        } else {
            // a spectrum in natural order (times the kernel for the MUL modes) into the slots
#pragma unroll 4
            for (int i = tid; i < N1 * ZC; i += NT) {
                const int xx = i / ZC, c = i % ZC;
                const size_t gi = static_cast<size_t>(xx) * sx + jbase + c;
                V v;
                if constexpr (MODE == PENCIL_MUL_INV_D2F) v = cmk<V>(static_cast<T>(job.src64[gi].x), static_cast<T>(job.src64[gi].y));
                else v = in[gi];
                if constexpr (MODE != PENCIL_INV) v = cscale(v, static_cast<T>(kv(c, xx)));
                lds[c + slot(xx) * LS] = v;
            }
            __syncthreads();
            reg::split_c<V, N1, 8, NT, 1, LS, false, true, ZC>(lds, tw, tid, reg::NoMul{});
            __syncthreads();
        }
        if constexpr (INV) {
            reg::split_a_to_regs<V, N1, 8, +1, 1, LS>(lds, x, l, p);
            V* out = data + jbase + l;
#pragma unroll
            for (int k1 = 0; k1 < NJ; ++k1) out[static_cast<size_t>(8 * k1 + p) * sx] = x[k1];
        } else {
#pragma unroll 4
            for (int i = tid; i < N1 * ZC; i += NT) {
                const int xx = i / ZC, c = i % ZC;
                const size_t gi = static_cast<size_t>(xx) * sx + jbase + c;
                const V v = lds[c + slot(xx) * LS];
                if constexpr (MODE == PENCIL_FWD_F2D) job.dst64[gi] = make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
                else data[gi] = v;
            }
        }
        return;
    }
    // FP32: pairs of columns per 16-byte access (ZC and sx are even)
    constexpr bool PAIRS = sizeof(T) == 4 && ZC % 2 == 0;
    if constexpr (MODE == PENCIL_MUL_INV_D2F) {
#pragma unroll 4
        for (int i = tid; i < N1 * ZC; i += NT) {
            const int x = i / ZC, c = i % ZC;
            const double2 d = job.src64[static_cast<size_t>(x) * sx + jbase + c];
            lds[x * LS + c] = cmk<V>(static_cast<T>(d.x), static_cast<T>(d.y));
        }
    } else if constexpr (PAIRS) {
#pragma unroll 4
        for (int j = tid; j < N1 * ZC / 2; j += NT) {
            const int i = 2 * j;
            const int x = i / ZC, c = i % ZC;
            if constexpr (MODE == PENCIL_CONV) ld_pair_stream(in + static_cast<size_t>(x) * sx + jbase + c, lds[x * LS + c], lds[x * LS + c + 1]);
            else ld_pair(in + static_cast<size_t>(x) * sx + jbase + c, lds[x * LS + c], lds[x * LS + c + 1]);
        }
    } else {
#pragma unroll 4
        for (int i = tid; i < N1 * ZC; i += NT) {
            const int x = i / ZC, c = i % ZC;
            lds[x * LS + c] = in[static_cast<size_t>(x) * sx + jbase + c];
        }
    }
    __syncthreads();
    if constexpr (MODE == PENCIL_FWD || MODE == PENCIL_FWD_F2D || MODE == PENCIL_FWD_MUL_INV || MODE == PENCIL_FWD_MUL_INV_NOISE || (MODE == PENCIL_CONV && !SPLIT))
        lds_fft<T, N1, -1, ZC, 1, LS, true, NT>(lds, tw, tid);
    if constexpr (MODE == PENCIL_CONV && !SPLIT) {
        const double* sxf = job.op.sep;
        const double c0 = job.op.c0;
#pragma unroll 4
        for (int i = tid; i < N1 * ZC; i += NT) {
            const int x = i / ZC;
            lds[x * LS + i % ZC] = cscale(lds[x * LS + i % ZC], static_cast<T>(sxf[x] * c0));
        }
        __syncthreads();
    }
    if constexpr (MODE == PENCIL_FWD_MUL_INV || MODE == PENCIL_FWD_MUL_INV_NOISE || MODE == PENCIL_MUL || MODE == PENCIL_MUL_INV || MODE == PENCIL_MUL_INV_D2F) {
        // every element a thread touches shares its column c (NT is a multiple of ZC):
        // the (y, z) part of the kernel is one per-thread constant
        const KOp op = job.op;
        static_assert(NT % ZC == 0, "pencil threads must cover whole rows");
        const int c = tid % ZC;
        const int j = static_cast<int>(jbase) + c;
        const int y = j / g.m3, z = j % g.m3;
        const double syz = op.sep ? op.sep[g.m1 + y] * op.sep[g.m1 + g.m2 + z] : 1.0;
        const double ky = g.kvec[g.m1 + y], kz = g.kvec[g.m1 + g.m2 + z];
#pragma unroll 4
        for (int i = tid; i < N1 * ZC; i += NT) {
            const int x = i / ZC;
            double kv = op.sep ? op.sep[x] * syz : 1.0;
            if (op.table) kv *= kop_table(op, op.table[static_cast<size_t>(x) * sx + j]);
            if (op.c1 != 0.0 || op.c2 != 0.0) {
                const double k[3] = {g.kvec[x], ky, kz};
                double poly = op.c0;
                if (op.c1 != 0.0) poly += op.c1 * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
                if (op.c2 != 0.0) poly += op.c2 * (k[op.ka] * k[op.kb]);
                kv *= poly;
            } else {
                kv *= op.c0;
            }
            lds[x * LS + c] = cscale(lds[x * LS + c], static_cast<T>(kv));
        }
        // k-space noise (semi-implicit steps), one element at a time: the Philox rounds
        // would otherwise multiply the registers of the unrolled loop above
        if constexpr (MODE == PENCIL_FWD_MUL_INV_NOISE) {
#pragma unroll 1
            for (int i = tid; i < N1 * ZC; i += NT) {
                const int x = i / ZC;
                const double2 nz = kop_noise<sizeof(T) == 4>(op, g.m1, g.m2, g.m3, static_cast<size_t>(x) * sx + j);
                lds[x * LS + c].x += static_cast<T>(nz.x);
                lds[x * LS + c].y += static_cast<T>(nz.y);
            }
        }
        __syncthreads();
    }
    if constexpr (MODE == PENCIL_INV || MODE == PENCIL_FWD_MUL_INV || MODE == PENCIL_FWD_MUL_INV_NOISE || MODE == PENCIL_MUL_INV || MODE == PENCIL_MUL_INV_D2F ||
                  (MODE == PENCIL_CONV && !SPLIT))
        lds_fft<T, N1, +1, ZC, 1, LS, true, NT>(lds, tw, tid);
    if constexpr (MODE == PENCIL_FWD_F2D) {
#pragma unroll 4
        for (int i = tid; i < N1 * ZC; i += NT) {
            const int x = i / ZC, c = i % ZC;
            const V v = lds[x * LS + c];
            job.dst64[static_cast<size_t>(x) * sx + jbase + c] = make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
        }
    } else if constexpr (PAIRS) {
#pragma unroll 4
        for (int j = tid; j < N1 * ZC / 2; j += NT) {
            const int i = 2 * j;
            const int x = i / ZC, c = i % ZC;
            st_pair(data + static_cast<size_t>(x) * sx + jbase + c, lds[x * LS + c], lds[x * LS + c + 1]);
        }
    } else {
#pragma unroll 4
        for (int i = tid; i < N1 * ZC; i += NT) {
            const int x = i / ZC, c = i % ZC;
            data[static_cast<size_t>(x) * sx + jbase + c] = lds[x * LS + c];
        }
    }
}

// ---------------------------------------------------------------------------
// The x pass of the conv schedule as DFT-as-GEMM on the matrix cores (FP32, the
// opt-in FTS_SCHED_CONV_MFMA): a length N1 = 16 R line is split n = R n1 + n2,
// k = k1 + 16 k2.  The DFT-16 over n1 runs on v_mfma_f32_16x16x4f32 (a 16 x 16
// complex matrix times 16 columns of the tile, as four real products per complex
// one), the twiddle e^{-2 pi i n2 k1 / N1} and the DFT-R over n2 in registers; then
// x phi_x(k) c0 and the same steps inverse.  One wave holds 16 columns in registers
// from load to store (no LDS).  Operand layout of the instruction: A[i][kk] from lane
// i + 16 kk, B[kk][j] from lane j + 16 kk, D[4 q + r][j] in register r of lane
// j + 16 q; chunk s of the sum over n1 (forward) takes n1 = 4 s + kk, and chunk s of
// the sum over k1 (inverse) takes k1 = 4 kk + s, the register s a lane holds.
// ---------------------------------------------------------------------------
constexpr int MFMA_WAVES = 4; // waves (16 columns each) per workgroup

__device__ __forceinline__ fts_f4 mfma4(float a, float b, fts_f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <int N1>
// (3 waves per SIMD: 138 VGPRs; the compiler's own choice was 2, and 4 spills)
__global__ void __launch_bounds__(64 * MFMA_WAVES) __attribute__((amdgpu_waves_per_eu(3))) pencil_mfma_kernel(PencilJob<float> j0, PencilJob<float> j1, const PencilJob<float>* __restrict__ table, FftGeom g)
{
    constexpr int R = N1 / 16;
    const PencilJob<float> job = pick_job<float>(j0, j1, table);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, c = lane & 15;
    const size_t sx = static_cast<size_t>(g.m2) * g.m3;
    const size_t col = static_cast<size_t>(blockIdx.x) * (16 * MFMA_WAVES) + wave * 16 + c;
    const float2* in = job.src ? job.src : job.data;
    const float2* twg = static_cast<const float2*>(g.tw_x); // e^{-2 pi i t / N1}, t < N1
    // DFT-16 entries this lane feeds: forward F[c][4 s + q], inverse conj F[c][4 q + s],
    // F[k][n] = e^{-2 pi i k n / 16} = twg[((k n) mod 16) R]
    float far[4], fai[4], iar[4], iai[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const float2 wf = twg[((c * (4 * s + q)) & 15) * R];
        const float2 wi = twg[((c * (4 * q + s)) & 15) * R];
        far[s] = wf.x;
        fai[s] = wf.y;
        iar[s] = wi.x;
        iai[s] = -wi.y;
    }
    // forward DFT-16 over n1 for each n2: D = sum_s A_s B_s, complex as four real products
    // (the four points of one n2 are loaded just before its products: fewer live registers)
    fts_f4 dr[R], di[R];
#pragma unroll
    for (int n2 = 0; n2 < R; ++n2) {
        float2 xv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) xv[s] = ld_point<float>(in + static_cast<size_t>(R * (4 * s + q) + n2) * sx + col);
        fts_f4 a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            a = mfma4(far[s], xv[s].x, a);
            a = mfma4(-fai[s], xv[s].y, a);
            b = mfma4(far[s], xv[s].y, b);
            b = mfma4(fai[s], xv[s].x, b);
        }
        dr[n2] = a;
        di[n2] = b;
    }
    // register r holds k1 = 4 q + r: twiddle, DFT-R over n2, x phi_x c0, inverse DFT-R, twiddle
    const double* sep = job.op.sep;
    const double c0 = job.op.c0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k1 = 4 * q + r;
        float2 y[R];
#pragma unroll
        for (int n2 = 0; n2 < R; ++n2) {
            y[n2] = make_float2(dr[n2][r], di[n2][r]);
            if (n2 != 0 && k1 != 0) y[n2] = cmul(y[n2], twg[n2 * k1]);
        }
        Dft<float2, R, -1>::run(y);
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) y[k2] = cscale(y[k2], static_cast<float>(sep[k1 + 16 * k2] * c0));
        Dft<float2, R, +1>::run(y);
#pragma unroll
        for (int n2 = 0; n2 < R; ++n2) {
            if (n2 != 0 && k1 != 0) {
                float2 t = twg[n2 * k1];
                t.y = -t.y;
                y[n2] = cmul(y[n2], t);
            }
            dr[n2][r] = y[n2].x;
            di[n2][r] = y[n2].y;
        }
    }
    // inverse DFT-16 over k1 for each n2; lane holds x[R (4 q + r) + n2] of its column
    float2* out = job.data;
#pragma unroll
    for (int n2 = 0; n2 < R; ++n2) {
        fts_f4 a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            a = mfma4(iar[s], dr[n2][s], a);
            a = mfma4(-iai[s], di[n2][s], a);
            b = mfma4(iar[s], di[n2][s], b);
            b = mfma4(iai[s], dr[n2][s], b);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) out[static_cast<size_t>(R * (4 * q + r) + n2) * sx + col] = make_float2(a[r], b[r]);
    }
}

template <typename T, int N1>
hipError_t pencil_mfma_go(const FftGeom& g, const PencilJob<T>* jobs, const PencilJob<T>* table, int njobs, hipStream_t s)
{
    if constexpr (sizeof(T) == 4 && (N1 == 32 || N1 == 48 || N1 == 64 || N1 == 96)) {
        const size_t cols = static_cast<size_t>(g.m2) * g.m3;
        if (cols % (16 * MFMA_WAVES) != 0) return hipErrorInvalidValue;
        dim3 grid(static_cast<unsigned>(cols / (16 * MFMA_WAVES)), njobs);
        const PencilJob<T> none{};
        hipLaunchKernelGGL((pencil_mfma_kernel<N1>), grid, dim3(64 * MFMA_WAVES), 0, s, jobs ? jobs[0] : none, jobs ? jobs[njobs > 1 ? 1 : 0] : none,
                           table, g);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

template <typename T, int N, int MODE>
hipError_t plane_go(const FftGeom& g, const PlaneJob<T>* jobs, const PlaneJob<T>* table, int njobs, hipStream_t s)
{
    dim3 grid(g.m1, njobs);
    const PlaneJob<T> none{};
    hipLaunchKernelGGL((plane_kernel<T, N, MODE>), grid, dim3(plane_threads_for<N, MODE>()), 0, s, jobs ? jobs[0] : none,
                       jobs ? jobs[njobs > 1 ? 1 : 0] : none, table, static_cast<const V2<T>*>(g.tw_yz));
    return hipGetLastError();
}

template <typename T, int N>
hipError_t plane_mode(const FftGeom& g, int mode, const PlaneJob<T>* jobs, const PlaneJob<T>* table, int njobs, hipStream_t s)
{
    switch (mode) {
    case PLANE_FWD: return plane_go<T, N, PLANE_FWD>(g, jobs, table, njobs, s);
    case PLANE_INV: return plane_go<T, N, PLANE_INV>(g, jobs, table, njobs, s);
    case PLANE_FWD_D2F: return plane_go<T, N, PLANE_FWD_D2F>(g, jobs, table, njobs, s);
    case PLANE_INV_F2D: return plane_go<T, N, PLANE_INV_F2D>(g, jobs, table, njobs, s);
    case PLANE_INIT: return plane_go<T, N, PLANE_INIT>(g, jobs, table, njobs, s);
    case PLANE_STEP: return plane_go<T, N, PLANE_STEP>(g, jobs, table, njobs, s);
    case PLANE_LAST: return plane_go<T, N, PLANE_LAST>(g, jobs, table, njobs, s);
    case PLANE_CONV_E: return plane_go<T, N, PLANE_CONV_E>(g, jobs, table, njobs, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int N1, int MODE, int ZC>
hipError_t pencil_launch(const FftGeom& g, const PencilJob<T>* jobs, const PencilJob<T>* table, int njobs, hipStream_t s)
{
    const size_t tiles = static_cast<size_t>(g.m2) * g.m3 / ZC;
    dim3 grid(static_cast<unsigned>(tiles), njobs);
    const PencilJob<T> none{};
    hipLaunchKernelGGL((pencil_kernel<T, N1, MODE, ZC>), grid, dim3(pencil_threads_for<N1, MODE, ZC>()), 0, s, jobs ? jobs[0] : none,
                       jobs ? jobs[njobs > 1 ? 1 : 0] : none, table, g);
    return hipGetLastError();
}

// Launches of four or more jobs (replica batches) run the bead-step multiplies on
// 32-column tiles when the (y, z) plane divides: measured 1.7 % (FP32) and 4 % (FP64)
// faster per 8-job group, 2 % slower for a single system.  Same arithmetic per column.
template <typename T, int N1, int MODE>
hipError_t pencil_go(const FftGeom& g, const PencilJob<T>* jobs, const PencilJob<T>* table, int njobs, hipStream_t s)
{
    if constexpr (MODE == PENCIL_CONV || MODE == PENCIL_FWD_MUL_INV) {
        if (njobs >= 4 && (static_cast<size_t>(g.m2) * g.m3) % 32 == 0) return pencil_launch<T, N1, MODE, 32>(g, jobs, table, njobs, s);
    }
    return pencil_launch<T, N1, MODE, PENCIL_ZC>(g, jobs, table, njobs, s);
}

template <typename T, int N1>
hipError_t pencil_mode(const FftGeom& g, int mode, const PencilJob<T>* jobs, const PencilJob<T>* table, int njobs, hipStream_t s)
{
    switch (mode) {
    case PENCIL_FWD_MUL_INV: return pencil_go<T, N1, PENCIL_FWD_MUL_INV>(g, jobs, table, njobs, s);
    case PENCIL_FWD: return pencil_go<T, N1, PENCIL_FWD>(g, jobs, table, njobs, s);
    case PENCIL_INV: return pencil_go<T, N1, PENCIL_INV>(g, jobs, table, njobs, s);
    case PENCIL_MUL: return pencil_go<T, N1, PENCIL_MUL>(g, jobs, table, njobs, s);
    case PENCIL_MUL_INV: return pencil_go<T, N1, PENCIL_MUL_INV>(g, jobs, table, njobs, s);
    case PENCIL_CONV: return pencil_go<T, N1, PENCIL_CONV>(g, jobs, table, njobs, s);
    case PENCIL_FWD_F2D: return pencil_go<T, N1, PENCIL_FWD_F2D>(g, jobs, table, njobs, s);
    case PENCIL_MUL_INV_D2F: return pencil_go<T, N1, PENCIL_MUL_INV_D2F>(g, jobs, table, njobs, s);
    case PENCIL_FWD_MUL_INV_NOISE: return pencil_go<T, N1, PENCIL_FWD_MUL_INV_NOISE>(g, jobs, table, njobs, s);
    case PENCIL_CONV_MFMA: return pencil_mfma_go<T, N1>(g, jobs, table, njobs, s);
    default: return hipErrorInvalidValue;
    }
}

constexpr int kSizes[] = {8, 12, 16, 24, 32, 48, 64, 96, 128};

bool in_sizes(int n)
{
    for (int s : kSizes)
        if (s == n) return true;
    return false;
}

} // namespace

bool line_supported(int m1, int m2, int m3)
{
    auto ok = [](int n) { return n == 192 || n == 256 || in_sizes(n); };
    if (!ok(m1) || !ok(m2) || !ok(m3)) return false;
    return (m2 * m3) % LINE_TC == 0 && (m1 * m3) % LINE_TC == 0 && (m1 * m2) % LINE_TC == 0;
}

bool conv_mfma_supported(int m1, int m2, int m3, bool fp32)
{
    return fp32 && fft_supported(m1, m2, m3, fp32) && (m1 == 32 || m1 == 48 || m1 == 64 || m1 == 96) &&
           (static_cast<size_t>(m2) * m3) % (16 * MFMA_WAVES) == 0;
}

bool fft_supported(int m1, int m2, int m3, bool fp32)
{
    if (!in_sizes(m1) || !in_sizes(m2) || m2 != m3) return false;
    const size_t elem = fp32 ? 8 : 16;
    const size_t lds = static_cast<size_t>(m2) * (m2 + 1) * elem + m2 * elem;
    return lds <= 160 * 1024;
}

template <typename T>
hipError_t plane_any(const FftGeom& g, int mode, const PlaneJob<T>* jobs, const PlaneJob<T>* table, int njobs, hipStream_t s)
{
    switch (g.m2) {
    case 8: return plane_mode<T, 8>(g, mode, jobs, table, njobs, s);
    case 12: return plane_mode<T, 12>(g, mode, jobs, table, njobs, s);
    case 16: return plane_mode<T, 16>(g, mode, jobs, table, njobs, s);
    case 24: return plane_mode<T, 24>(g, mode, jobs, table, njobs, s);
    case 32: return plane_mode<T, 32>(g, mode, jobs, table, njobs, s);
    case 48: return plane_mode<T, 48>(g, mode, jobs, table, njobs, s);
    case 64: return plane_mode<T, 64>(g, mode, jobs, table, njobs, s);
    case 96: return plane_mode<T, 96>(g, mode, jobs, table, njobs, s);
    case 128:
        if constexpr (sizeof(T) == 4) return plane_mode<T, 128>(g, mode, jobs, table, njobs, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t pencil_any(const FftGeom& g, int mode, const PencilJob<T>* jobs, const PencilJob<T>* table, int njobs, hipStream_t s)
{
    if (mode == PENCIL_FWD_MUL_INV && jobs && (jobs[0].op.namp != 0.0 || (njobs > 1 && jobs[1].op.namp != 0.0))) mode = PENCIL_FWD_MUL_INV_NOISE;
    switch (g.m1) {
    case 8: return pencil_mode<T, 8>(g, mode, jobs, table, njobs, s);
    case 12: return pencil_mode<T, 12>(g, mode, jobs, table, njobs, s);
    case 16: return pencil_mode<T, 16>(g, mode, jobs, table, njobs, s);
    case 24: return pencil_mode<T, 24>(g, mode, jobs, table, njobs, s);
    case 32: return pencil_mode<T, 32>(g, mode, jobs, table, njobs, s);
    case 48: return pencil_mode<T, 48>(g, mode, jobs, table, njobs, s);
    case 64: return pencil_mode<T, 64>(g, mode, jobs, table, njobs, s);
    case 96: return pencil_mode<T, 96>(g, mode, jobs, table, njobs, s);
    case 128: return pencil_mode<T, 128>(g, mode, jobs, table, njobs, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_plane(const FftGeom& g, int mode, const PlaneJob<T>* jobs, int njobs, hipStream_t s)
{
    return plane_any<T>(g, mode, jobs, nullptr, njobs, s);
}
template <typename T>
hipError_t launch_plane_table(const FftGeom& g, int mode, const PlaneJob<T>* dev_jobs, int njobs, hipStream_t s)
{
    return plane_any<T>(g, mode, nullptr, dev_jobs, njobs, s);
}
template <typename T>
hipError_t launch_pencil(const FftGeom& g, int mode, const PencilJob<T>* jobs, int njobs, hipStream_t s)
{
    return pencil_any<T>(g, mode, jobs, nullptr, njobs, s);
}
template <typename T>
hipError_t launch_pencil_table(const FftGeom& g, int mode, const PencilJob<T>* dev_jobs, int njobs, hipStream_t s)
{
    return pencil_any<T>(g, mode, nullptr, dev_jobs, njobs, s);
}

template hipError_t launch_plane<double>(const FftGeom&, int, const PlaneJob<double>*, int, hipStream_t);
template hipError_t launch_plane<float>(const FftGeom&, int, const PlaneJob<float>*, int, hipStream_t);
template hipError_t launch_pencil<double>(const FftGeom&, int, const PencilJob<double>*, int, hipStream_t);
template hipError_t launch_pencil<float>(const FftGeom&, int, const PencilJob<float>*, int, hipStream_t);
template hipError_t launch_plane_table<double>(const FftGeom&, int, const PlaneJob<double>*, int, hipStream_t);
template hipError_t launch_plane_table<float>(const FftGeom&, int, const PlaneJob<float>*, int, hipStream_t);
template hipError_t launch_pencil_table<double>(const FftGeom&, int, const PencilJob<double>*, int, hipStream_t);
template hipError_t launch_pencil_table<float>(const FftGeom&, int, const PencilJob<float>*, int, hipStream_t);

} // namespace fts
