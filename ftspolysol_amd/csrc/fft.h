// fft.h -- host interface of the engine's in-house 3D FFT passes (fft.hip).
//
// A 3D transform of an m1 x m2 x m3 grid (x slowest) is split into two kinds of
// pass, each one HBM read and one HBM write of the field:
//   plane pass   one workgroup per x-plane: the whole m2 x m3 plane is staged in
//                LDS and transformed along z and y (2D FFT);
//   pencil pass  one workgroup per (y, z-chunk) tile of m1 x ZC points:
//                transform along x, optionally multiply by a k-space kernel and
//                transform back.
// The propagator bead step q' = e^{-W} IFFT[phi FFT[q] / M] is then
//   pencil(FWD, x phi/M, INV)   then   plane(INV, x e^{-W}, store q', FWD)
// with the forward plane pass of the next bead folded into the current one.
// Each pass can carry two independent jobs (blockIdx.y), so the forward and
// backward propagator sweeps share every launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "fts_device.h"

namespace fts {

// k-space kernel applied in the pencil pass:
//   K(k) = [sep ? fx[kx] fy[ky] fz[kz] : 1] * [table ? table[idx] : 1] * (c0 + c1 k^2 + c2 k_a k_b)
// sep holds m1 + m2 + m3 doubles (x, y, z factors); table holds M doubles.
struct KOp {
    const double* sep;
    const double* table;
    double c0, c1, c2;
    int ka, kb;
    int table_form; // the factor taken from the table value t (kop_table)
    // k-space noise of a semi-implicit step (namp != 0): the multiply pass adds
    // namp f(t) g(k), g the spectrum of real white noise of unit variance per site
    // divided by sqrt(M) (kop_noise), f(t) = sqrt(t (1 + t) / 2), or 2 t / (1 + t) when
    // ncn; nzero drops k = 0.  nkey: the device noise key {seed, step} of the step (read
    // when the pass runs, so a replayed graph draws the noise of the step it replays)
    const unsigned long long* nkey;
    unsigned nfield;
    double namp;
    int nzero;
    int ncn;
};

// Factors of a semi-implicit table t = P = 1 / (1 + a), a = dt lam A(k):
//   KOP_T   t                        the semi-implicit step's filter
//   KOP_CN  2 t / (1 + t) = 1/(1 + a/2)   the predictor-corrector's corrector (Crank-Nicolson)
//   KOP_CNR (1 - t) / (1 + t) = 1 - KOP_CN  its share of the predictor's increment
enum KopForm { KOP_T = 0, KOP_CN = 2, KOP_CNR = 3 };
__host__ __device__ __forceinline__ double kop_table(const KOp& op, double t)
{
    return op.table_form == KOP_CN ? 2.0 * t / (1.0 + t) : op.table_form == KOP_CNR ? (1.0 - t) / (1.0 + t) : t;
}

// PLANE_FWD_D2F / PLANE_INV_F2D: PLANE_FWD reading src64, PLANE_INV writing dst64 (complex
// FP64 fields transformed in the pass's precision)
enum PlaneMode { PLANE_FWD = 0, PLANE_INV = 1, PLANE_INIT = 2, PLANE_STEP = 3, PLANE_LAST = 4, PLANE_CONV_E = 5, PLANE_FWD_D2F = 6, PLANE_INV_F2D = 7 };
// PENCIL_CONV: forward along x, x op.sep[x] op.c0 (the x factor of a separable
// kernel), inverse along x; the plane pass PLANE_CONV_E applies the (y, z) factors
// PENCIL_FWD_F2D: PENCIL_FWD writing dst64; PENCIL_MUL_INV_D2F: PENCIL_MUL_INV reading src64
// PENCIL_FWD_MUL_INV_NOISE / LINE_FWD_MUL_INV_NOISE: the multiply adds the k-space
// noise of its KOp (the launch functions pick them when a job's op has namp != 0; the
// Philox rounds stay out of the registers of the plain multiply)
// PENCIL_CONV_MFMA: PENCIL_CONV as DFT-as-GEMM on the matrix cores (FP32, m1 in {32, 48, 64, 96},
// m2 m3 a multiple of 64; the opt-in FTS_SCHED_CONV_MFMA)
enum PencilMode { PENCIL_FWD_MUL_INV = 0, PENCIL_FWD = 1, PENCIL_INV = 2, PENCIL_MUL = 3, PENCIL_MUL_INV = 4, PENCIL_CONV = 5, PENCIL_FWD_F2D = 6,
                  PENCIL_MUL_INV_D2F = 7, PENCIL_FWD_MUL_INV_NOISE = 8, PENCIL_CONV_MFMA = 9 };

template <typename T> struct PlaneJob {
    const V2<T>* src;
    V2<T>* dst;
    V2<T>* qout;
    const V2<T>* emw;
    // Per-bead renormalisation (INIT / STEP / LAST).  When norm_out is set, the
    // pass writes max |q_out|^2 of each x-plane to norm_out[x]; when norm_in is
    // set, q_out is scaled by c = 2^k, k = round(-log2(max |q_prev|)), and block 0
    // stores k in *log_scale.  Powers of two scale exactly, so the recursion's
    // values change only by the exponent, and |q| stays below ~1.4 times the
    // growth of one bead step.
    const double* norm_in;
    double* norm_out;
    double* log_scale;
    KOp op; // PLANE_CONV_E: a separable k-space kernel; its (y, z) factors are applied here
    const double2* src64; // PLANE_FWD_D2F
    double2* dst64;       // PLANE_INV_F2D
};

template <typename T> struct PencilJob {
    V2<T>* data;
    KOp op;
    const V2<T>* src; // read from here instead of data when set (the result still goes to data)
    const double2* src64; // PENCIL_MUL_INV_D2F
    double2* dst64;       // PENCIL_FWD_F2D
};

// Line pass: a tile of LINE_TC lines of one axis, transformed along that axis in LDS.
//   axis 0 (x): column j in [0, m2 m3),  point t at  t*m2*m3 + j
//   axis 1 (y): column j in [0, m1 m3),  point t at  (j/m3)*m2*m3 + t*m3 + j%m3
//   axis 2 (z): column j in [0, m1 m2),  point t at  j*m3 + t
// Modes: FWD, INV, FWD_MUL_INV (k-space kernel, as the pencil pass), and on z only:
// INV_E_FWD (inverse, x e^{-W} (x 2^k), store q, forward), INV_E (same without the
// forward transform) and E_FWD (q = e^{-W} (x 2^k), store q, forward).  The bead step
// with line passes is x(FWD_MUL_INV) -> y(INV) -> z(INV_E_FWD) -> y(FWD); it moves
// more bytes than the plane/pencil pair but every pass has m^2/16 workgroups.
// The separable schedule convolves one axis per pass with that axis' factor of the bond
// kernel: x(CONV, q_s from src) -> z(CONV) -> y(CONV_E: x e^{-W} (x 2^k), store q); E
// stores q = e^{-W} (x 2^k) alone (the first bead of a sweep).
constexpr int LINE_TC = 16;
enum LineMode { LINE_FWD = 0, LINE_INV = 1, LINE_FWD_MUL_INV = 2, LINE_INV_E_FWD = 3, LINE_INV_E = 4, LINE_E_FWD = 5, LINE_FWD_MUL_INV_NOISE = 6,
                LINE_CONV = 7, LINE_CONV_E = 8, LINE_E = 9 };

template <typename T> struct LineJob {
    V2<T>* data;          // transformed in place (E_FWD writes its result here)
    V2<T>* qout;          // z modes that produce a propagator
    const V2<T>* emw;     // e^{-W} of the bead (z modes)
    KOp op;               // FWD_MUL_INV
    const double* norm_in;
    double* norm_out;     // per-block max |q|^2 (gridDim.x entries)
    double* log_scale;
    const V2<T>* src;     // CONV: read from here instead of data when set (the result still goes to data)
};

// The job of a workgroup: row blockIdx.y of a device job table, or (no table) j0
// for blockIdx.y == 0 and j1 otherwise.  Built member by member: a select between
// the two by-value kernel arguments as whole structs is lowered through private
// per-lane copies of both (120 bytes of scratch written by every lane).
__device__ __forceinline__ KOp pick_kop(KOp a, KOp b, bool f)
{
    KOp k;
    k.sep = f ? a.sep : b.sep;
    k.table = f ? a.table : b.table;
    k.c0 = f ? a.c0 : b.c0;
    k.c1 = f ? a.c1 : b.c1;
    k.c2 = f ? a.c2 : b.c2;
    k.ka = f ? a.ka : b.ka;
    k.kb = f ? a.kb : b.kb;
    k.table_form = f ? a.table_form : b.table_form;
    k.nkey = f ? a.nkey : b.nkey;
    k.nfield = f ? a.nfield : b.nfield;
    k.namp = f ? a.namp : b.namp;
    k.nzero = f ? a.nzero : b.nzero;
    k.ncn = f ? a.ncn : b.ncn;
    return k;
}

// g(k) at grid index gi = (x m2 + y) m3 + z: Hermitian, so that its inverse transform is
// real.  The pair {k, -k} takes the two Box-Muller normals a, b of one Philox draw
// (seed and step from nkey, nfield, site = the pair's smaller index): g = (a +- i b) / sqrt 2, + at
// the smaller index; a self-conjugate k (every component 0 or m/2) takes a alone: g = a.
// This is the distribution of FFT[eta] / sqrt(M) for eta real white noise of unit variance.
// F32 (FP32 transforms): the normals' transcendentals in FP32.
template <bool F32 = false>
__device__ __forceinline__ double2 kop_noise(const KOp& op, int m1, int m2, int m3, size_t gi)
{
    double2 r = make_double2(0.0, 0.0);
    if (op.namp == 0.0 || !op.table) return r;
    const size_t s23 = static_cast<size_t>(m2) * m3;
    const int x = static_cast<int>(gi / s23), y = static_cast<int>((gi / m3) % m2), z = static_cast<int>(gi % m3);
    const size_t pi = (static_cast<size_t>((m1 - x) % m1) * m2 + (m2 - y) % m2) * m3 + (m3 - z) % m3;
    if (op.nzero && gi == 0) return r;
    const double t = op.table[gi];
    const double amp = op.namp * (op.ncn ? 2.0 * t / (1.0 + t) : sqrt(0.5 * t * (1.0 + t)));
    const double2 ab = cl_normal_pair<F32>(op.nkey[0], op.nkey[1], op.nfield, static_cast<uint32_t>(gi < pi ? gi : pi));
    if (pi == gi) {
        r.x = amp * ab.x;
        return r;
    }
    const double h = amp * 0.70710678118654752440;
    r.x = h * ab.x;
    r.y = gi < pi ? h * ab.y : -h * ab.y;
    return r;
}
template <typename T> __device__ __forceinline__ PlaneJob<T> pick_job(PlaneJob<T> a, PlaneJob<T> b, const PlaneJob<T>* table)
{
    PlaneJob<T> j;
    const PlaneJob<T>* t = table + blockIdx.y;
    const bool tb = table != nullptr, f = blockIdx.y == 0;
    j.src = tb ? t->src : f ? a.src : b.src;
    j.dst = tb ? t->dst : f ? a.dst : b.dst;
    j.qout = tb ? t->qout : f ? a.qout : b.qout;
    j.emw = tb ? t->emw : f ? a.emw : b.emw;
    j.norm_in = tb ? t->norm_in : f ? a.norm_in : b.norm_in;
    j.norm_out = tb ? t->norm_out : f ? a.norm_out : b.norm_out;
    j.log_scale = tb ? t->log_scale : f ? a.log_scale : b.log_scale;
    j.op = tb ? pick_kop(t->op, t->op, true) : pick_kop(a.op, b.op, f);
    j.src64 = tb ? t->src64 : f ? a.src64 : b.src64;
    j.dst64 = tb ? t->dst64 : f ? a.dst64 : b.dst64;
    return j;
}
template <typename T> __device__ __forceinline__ PencilJob<T> pick_job(PencilJob<T> a, PencilJob<T> b, const PencilJob<T>* table)
{
    PencilJob<T> j;
    const bool f = blockIdx.y == 0;
    if (table) {
        const PencilJob<T>* t = table + blockIdx.y;
        j.data = t->data;
        j.op = pick_kop(t->op, t->op, true);
        j.src = t->src;
        j.src64 = t->src64;
        j.dst64 = t->dst64;
    } else {
        j.data = f ? a.data : b.data;
        j.op = pick_kop(a.op, b.op, f);
        j.src = f ? a.src : b.src;
        j.src64 = f ? a.src64 : b.src64;
        j.dst64 = f ? a.dst64 : b.dst64;
    }
    return j;
}
template <typename T> __device__ __forceinline__ LineJob<T> pick_job(LineJob<T> a, LineJob<T> b, const LineJob<T>* table)
{
    LineJob<T> j;
    const bool f = blockIdx.y == 0;
    if (table) {
        const LineJob<T>* t = table + blockIdx.y;
        j.data = t->data;
        j.qout = t->qout;
        j.emw = t->emw;
        j.op = pick_kop(t->op, t->op, true);
        j.norm_in = t->norm_in;
        j.norm_out = t->norm_out;
        j.log_scale = t->log_scale;
        j.src = t->src;
    } else {
        j.data = f ? a.data : b.data;
        j.qout = f ? a.qout : b.qout;
        j.emw = f ? a.emw : b.emw;
        j.op = pick_kop(a.op, b.op, f);
        j.norm_in = f ? a.norm_in : b.norm_in;
        j.norm_out = f ? a.norm_out : b.norm_out;
        j.log_scale = f ? a.log_scale : b.log_scale;
        j.src = f ? a.src : b.src;
    }
    return j;
}

struct FftGeom {
    int m1, m2, m3;
    const double* kvec;   // device: kx[m1], ky[m2], kz[m3]
    const void* tw_x;     // device twiddles e^{-2 pi i t / m1} (T complex)
    const void* tw_yz;    // device twiddles e^{-2 pi i t / m2}, m2 == m3
};

// True when the plane and pencil kernels exist for this grid and precision.
bool fft_supported(int m1, int m2, int m3, bool fp32);
// True when the x pass of FTS_SCHED_CONV_MFMA exists for this grid (FP32 only)
bool conv_mfma_supported(int m1, int m2, int m3, bool fp32);
// the conv schedule and its matrix-core form share every pass but the x pass
inline bool sched_is_conv(int sched) { return sched == 3 || sched == 5; }
inline int conv_pencil_mode(int sched) { return sched == 5 ? PENCIL_CONV_MFMA : PENCIL_CONV; }

// Launch one plane pass over all x-planes for njobs (1 or 2) jobs.
template <typename T>
hipError_t launch_plane(const FftGeom& g, int mode, const PlaneJob<T>* jobs, int njobs, hipStream_t s);

// True when line kernels exist for this grid (any m1, m2, m3 in the size set).
bool line_supported(int m1, int m2, int m3);

// Launch one line pass along `axis` for njobs (1 or 2) jobs.  tw: twiddles of that axis.
template <typename T>
hipError_t launch_line(const FftGeom& g, int axis, int mode, const void* tw, const LineJob<T>* jobs, int njobs, hipStream_t s);

// Launch one pencil pass for njobs (1 or 2) jobs.
template <typename T>
hipError_t launch_pencil(const FftGeom& g, int mode, const PencilJob<T>* jobs, int njobs, hipStream_t s);

// The same passes over njobs jobs read from a device-resident job table (blockIdx.y
// indexes the table): the propagator sweeps of many replicas share each launch.
template <typename T>
hipError_t launch_plane_table(const FftGeom& g, int mode, const PlaneJob<T>* dev_jobs, int njobs, hipStream_t s);
template <typename T>
hipError_t launch_pencil_table(const FftGeom& g, int mode, const PencilJob<T>* dev_jobs, int njobs, hipStream_t s);
template <typename T>
hipError_t launch_line_table(const FftGeom& g, int axis, int mode, const void* tw, const LineJob<T>* dev_jobs, int njobs, hipStream_t s);

} // namespace fts
