// IMRPhenomXAS on MI355X (gfx950): the per-binary setup kernel, the
// per-(binary, frequency) evaluation kernels on a shared grid and on ragged
// per-binary grids, and the C ABI the Python layer binds with ctypes
// (declared in include/waveformx_amd.h).
//
// Launch shapes:
//  * setup: one thread per binary (per binary and direction for the
//    derivatives), 64-thread blocks;
//  * shared grid: 256-thread blocks; a block owns 256 consecutive
//    frequencies and a group of binaries, computes each frequency's powers
//    once and loops over the group (the record is wave-uniform: scalar
//    loads), so the per-frequency work is amortised over the group; blocks
//    are remapped so that one group's frequency tiles share an XCD's L2;
//  * ragged: a persistent grid walks (binary, tile) work items from a
//    prefix sum of tiles per binary; one thread per bin.
#include "phenomd_point.h"
#include "phenomd_adjoint.h"
#include "xas_adjoint.h"
#include "xhm_adjoint.h"
#include "xhm_point.h"

#include <atomic>
#include <climits>
#include <cstdio>
#include <limits>
#include <mutex>
#include <string>

using namespace wfx;

namespace {

thread_local std::string g_error;

int fail(const char* what) {
    g_error = what;
    return -1;
}

int check_launch(const char* what) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        g_error = std::string(what) + ": " + hipGetErrorString(err);
        return -1;
    }
    return 0;
}

constexpr int kPrecF64 = 0;
constexpr int kPrecF32 = 1;
constexpr int kPrecSplit = 2;
constexpr int kPrecSplitDF = 3;

int cu_count() {
    static std::atomic<int> cached{0};
    int n = cached.load();
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cached.store(n);
    }
    return n;
}

// blocks are dealt round-robin to the 8 XCDs; give XCD x a contiguous range
// of work ids so that neighbouring work (one group's tiles) shares its L2
__device__ inline int64_t xcd_work_id(int64_t block, int64_t nblocks) {
    const int64_t full = nblocks & ~int64_t(7);
    if (block >= full) return block;
    return (block & 7) * (full >> 3) + (block >> 3);
}

// the same for output-bound launches: the XCDs take turns at consecutive
// chunks of `chunk` work ids (XCD x the x-th chunk of each round of 8), so
// that the blocks running at one time on all XCDs work on one window of the
// output, while each chunk's neighbouring work stays on one XCD; the last,
// partial round splits evenly. A bijection of [0, nblocks).
__device__ inline int64_t xcd_chunk_work_id(int64_t block, int64_t nblocks, int64_t chunk) {
    const int64_t full = nblocks & ~int64_t(7);
    if (block >= full) return block;
    const int64_t xcd = block & 7, k = block >> 3, per = full >> 3;
    const int64_t round = k / chunk, within = k % chunk;
    const int64_t rounds = per / chunk;
    if (round < rounds) return round * 8 * chunk + xcd * chunk + within;
    return rounds * 8 * chunk + xcd * (per - rounds * chunk) + within;
}

// ------------------------------------------------------------------ setup

template <class R>
__device__ inline void with_reference_mf(R* p, double reference_mf) {
    // xas/model.py:35-45: f_ref <- reference_mf / M where f_ref was zero
    if (reference_mf != reference_mf) return;
    R scaled = R(reference_mf) / (R(kMtsunSI) * (p[kMass1] + p[kMass2]));
    p[kFRef] = value_of(p[kFRef]) == 0.0 ? scaled : R(__builtin_nan(""));
}

// FP64 record
__global__ __launch_bounds__(64) void setup_f64_kernel(const double* __restrict__ params, int64_t nb,
                                                       const double* __restrict__ f_min, int64_t f_min_stride,
                                                       double amp_scale, double reference_mf,
                                                       double* __restrict__ records, uint32_t* __restrict__ flags) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double p[kParams];
    for (int k = 0; k < kParams; ++k) p[k] = params[b * kParams + k];
    with_reference_mf(p, reference_mf);
    double rec[kSlots];
    const unsigned fl = xas_record<double>(p, f_min[b * f_min_stride], amp_scale, rec);
    rec[kT0Lo] = 0.0;
    rec[kAnchorLo] = 0.0;
    double* out = records + b * kSlots;
    for (int k = 0; k < kSlots; ++k) out[k] = rec[k];
    if (flags) flags[b] = fl;
}

// FP32 record from FP64 slots: every slot rounded to FP32 but the
// post-inspiral constants (slots const_a, const_b: IMRPhenomXAS's by
// default), which take -A modulo 2 pi first (fold_anchor); T0 and A also
// as hi + lo and as doubles in slots 60..63 (low word first)
__device__ inline void write_f32_record(const double* rec, float* __restrict__ out, int const_a = kIntConst,
                                        int const_b = kRdConst) {
    for (int k = 0; k < 60; ++k) out[k] = (float)rec[k];
    const double t0 = rec[kT0], anchor = rec[kAnchor];
    out[const_a] = (float)fold_anchor(rec[const_a], anchor);
    out[const_b] = (float)fold_anchor(rec[const_b], anchor);
    out[kT0] = (float)t0;
    out[kT0Lo] = (float)(t0 - (double)out[kT0]);
    out[kAnchor] = (float)anchor;
    out[kAnchorLo] = (float)(anchor - (double)out[kAnchor]);
    out[kT0F64] = __int_as_float(__double2loint(t0));
    out[kT0F64 + 1] = __int_as_float(__double2hiint(t0));
    out[kAnchorF64] = __int_as_float(__double2loint(anchor));
    out[kAnchorF64 + 1] = __int_as_float(__double2hiint(anchor));
}

// FP32 record of the FP64 setup (split precision)
__global__ __launch_bounds__(64) void setup_split_kernel(const double* __restrict__ params, int64_t nb,
                                                         const double* __restrict__ f_min, int64_t f_min_stride,
                                                         double amp_scale, double reference_mf,
                                                         float* __restrict__ records, uint32_t* __restrict__ flags) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double p[kParams];
    for (int k = 0; k < kParams; ++k) p[k] = params[b * kParams + k];
    with_reference_mf(p, reference_mf);
    double rec[kSlots];
    const unsigned fl = xas_record<double>(p, f_min[b * f_min_stride], amp_scale, rec);
    write_f32_record(rec, records + b * kSlots);
    if (flags) flags[b] = fl;
}

// plain FP32 record: the whole setup in FP32 (then written as above, T0 and
// A exact in FP32, so their low parts are zero)
__global__ __launch_bounds__(64) void setup_f32_kernel(const double* __restrict__ params, int64_t nb,
                                                       const double* __restrict__ f_min, int64_t f_min_stride,
                                                       double amp_scale, double reference_mf,
                                                       float* __restrict__ records, uint32_t* __restrict__ flags) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    float p[kParams];
    for (int k = 0; k < kParams; ++k) p[k] = (float)params[b * kParams + k];
    with_reference_mf(p, reference_mf);
    float rec[kSlots];
    const unsigned fl = xas_record<float>(p, (float)f_min[b * f_min_stride], amp_scale, rec);
    double wide[kSlots];
    for (int k = 0; k < kSlots; ++k) wide[k] = (double)rec[k];
    write_f32_record(wide, records + b * kSlots);
    if (flags) flags[b] = fl;
}

// The models' per-binary setups, for the derivative setup: the record of a
// binary in any scalar (with reference_mf applied as the model does), each
// slot handed to out(k, value) once, and the slots of the FP32 record that
// take the anchor (write_f32_record).
struct XasRecord {
    static constexpr int kFoldA = kIntConst, kFoldB = kRdConst;
    template <class R, class Out>
    __device__ __forceinline__ static unsigned record_to(R* p, const R& f_min, double amp_scale,
                                                         double reference_mf, Out& out,
                                                         XasCoeffs22<R>* keep = nullptr) {
        with_reference_mf(p, reference_mf);
        return xas_record_to<R>(p, f_min, amp_scale, out, keep);
    }
    // The record's derivative with respect to parameter dir (d_L,
    // inclination, phi_ref or f_ref) from its FP64 values v, into col[k *
    // stride] (the slots it leaves alone are zero): d_L divides the
    // amplitude's scale, the inclination sets the two polarization factors,
    // phi_ref and f_ref move the phase constants by 2 phi_ref - phase(f_ref)
    // and f_ref the anchor T0 f_ref^(-5/3) (xas/core.py:257-379)
    __device__ static void closed_column(int dir, const double* p, double reference_mf, unsigned flags,
                                         const double* v, double* col, int stride) {
        using D = Dual<1>;
        if (dir == kDistance) {
            const double inv = -1.0 / p[kDistance];
            for (int k = 0; k < 10; ++k) col[(kAmpIns + k) * stride] = v[kAmpIns + k] * inv;
            col[kAmpIntN * stride] = v[kAmpIntN] * inv;
            col[kAmpRdN * stride] = v[kAmpRdN] * inv;
        } else if (dir == kInclination) {
            D iota(p[kInclination]);
            iota.d[0] = 1.0;
            const D ci = m_cos(iota);
            const D pfac = D(0.5) * (D(1.0) + ci * ci);
            col[kKp * stride] = flags ? 0.0 : (D(-kYlm) * pfac).d[0];
            col[kKc * stride] = flags ? 0.0 : (D(kYlm) * ci).d[0];
        } else if (dir == kPhiRef) {
            col[kInsConst * stride] = 2.0;
            col[kIntConst * stride] = 2.0;
            col[kRdConst * stride] = 2.0;
        } else if (reference_mf != reference_mf && p[kFRef] != 0.0) {
            // (with reference_mf, or f_ref = 0 meaning f_min, there is no
            // f_ref to differentiate)
            D f(p[kFRef]);
            f.d[0] = 1.0;
            bool past;
            const double rest = phase_rest(v, freq_powers(f), past).d[0];
            col[kInsConst * stride] = -rest;
            col[kIntConst * stride] = -rest;
            col[kRdConst * stride] = -rest;
            col[kAnchor * stride] = past ? 0.0 : (D(v[kT0]) * m_exp(D(-5.0 / 3.0) * m_log(f))).d[0];
        }
    }
};

struct PhenomDRecord {
    static constexpr int kFoldA = kPdIntConst, kFoldB = kPdMrdConst;
    template <class R, class Out>
    __device__ __forceinline__ static unsigned record_to(R* p, const R& f_min, double amp_scale,
                                                         double reference_mf, Out& out) {
        R rec[kSlots];
        const unsigned fl = pd_record<R>(p, f_min, amp_scale, reference_mf, rec);
#pragma unroll
        for (int k = 0; k < kSlots; ++k) out(k, rec[k]);
        return fl;
    }
    // as XasRecord::closed_column, for IMRPhenomD's record
    // (phenomd/core.py:274-377)
    __device__ static void closed_column(int dir, const double* p, double reference_mf, unsigned flags,
                                         const double* v, double* col, int stride) {
        using D = Dual<1>;
        const bool with_mf = reference_mf == reference_mf;
        if (dir == kDistance) {
            const double inv = -1.0 / p[kDistance];
            for (int k = 0; k < 10; ++k) col[(kPdAmpIns + k) * stride] = v[kPdAmpIns + k] * inv;
            for (int j = 0; j < 5; ++j) col[(kPdAmpInt + j) * stride] = v[kPdAmpInt + j] * inv;
            col[kPdAmpRdN * stride] = v[kPdAmpRdN] * inv;
        } else if (dir == kInclination) {
            const bool error = flags != 0 || (with_mf && p[kFRef] != 0.0);
            D iota(p[kInclination]);
            iota.d[0] = 1.0;
            const D cos_i = m_cos(iota);
            col[kPdKp * stride] = error ? 0.0 : (D(0.5) * (D(1.0) + cos_i * cos_i)).d[0];
            col[kPdKc * stride] = error ? 0.0 : (-cos_i).d[0];
        } else if (dir == kPhiRef) {
            col[kPdInsConst * stride] = 2.0;
            col[kPdIntConst * stride] = 2.0;
            col[kPdMrdConst * stride] = 2.0;
        } else if (!with_mf && p[kFRef] != 0.0) {
            D f(p[kFRef]);
            f.d[0] = 1.0;
            bool inspiral;
            const double rest = pd_phase_rest(v, pd_powers(f), inspiral).d[0];
            col[kPdInsConst * stride] = -rest;
            col[kPdIntConst * stride] = -rest;
            col[kPdMrdConst * stride] = -rest;
            col[kAnchor * stride] = inspiral ? (D(v[kT0]) * m_exp(D(-5.0 / 3.0) * m_log(f))).d[0] : 0.0;
        }
    }
};

// the record and its derivatives with respect to the 8 parameters (m1, m2,
// chi1, chi2, distance, inclination, phi_ref, f_ref), jac (B, 64, 8): one
// thread per (binary, direction) for the masses and spins, through the
// whole setup in dual numbers; the thread of direction j also writes the
// column of parameter j + 4, which enters the record only through a few
// slots (S::closed_column). The direction-0 thread writes the record, FP64
// (records) and/or FP32 split (split_records), either NULL. With Keep
// (IMRPhenomXAS), the thread of (b, dir) also writes its coefficient structs
// in dual numbers to keep[4 b + dir], for IMRPhenomXHM's higher modes.
// Each slot leaves the set-up as soon as it is final: its derivative and its
// value (direction 0's) to the wave's LDS, so that the set-up's registers
// never hold the whole record in dual numbers. After the set-up the wave
// stores its binaries' Jacobian rows whole (derivatives and the closed-form
// half's zeros, contiguous 64-byte rows), then the closed-form entries over
// them: rows stored as 32-byte halves (each thread its column) moved 58 MB
// for 36 MB at 8192 binaries, whole rows 41 MB.
struct JvpSink {
    double* col;   // the thread's derivatives (LDS): slot k at col[4 k]
    double* vals;  // the binary's values, slot k at vals[k]
    bool lead;     // direction 0: writes the values
    __device__ void operator()(int k, const Dual<1>& x) const {
        col[k * 4] = k < 60 ? x.d[0] : 0.0;
        if (lead) vals[k] = x.v;
    }
};

template <class S, bool Keep = false>
__global__ __launch_bounds__(64) void setup_jvp_kernel(const double* __restrict__ params, int64_t nb,
                                                       const double* __restrict__ f_min, int64_t f_min_stride,
                                                       double amp_scale, double reference_mf,
                                                       double* __restrict__ records, float* __restrict__ split_records,
                                                       double* __restrict__ jac, uint32_t* __restrict__ flags,
                                                       XasCoeffs22<Dual<1>>* __restrict__ keep = nullptr) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = gid >> 2;
    const int dir = (int)(gid & 3);
    const bool live = b < nb;
    // the 16 binaries' values of the block (one wave), and their Jacobian
    // rows' derivative halves (after those are stored, their FP32 records)
    __shared__ double s_vals[16][kSlots];
    __shared__ double s_dual[16][kSlots][4];
    double* vals = s_vals[threadIdx.x >> 2];
    using D = Dual<1>;
    unsigned fl = 0;
    if (live) {
        D p[kParams];
        for (int k = 0; k < kParams; ++k) {
            p[k] = D(params[b * kParams + k]);
            p[k].d[0] = k == dir ? 1.0 : 0.0;
        }
        JvpSink out{&s_dual[threadIdx.x >> 2][0][dir], vals, dir == 0};
        if constexpr (Keep)
            fl = S::template record_to<D>(p, D(f_min[b * f_min_stride]), amp_scale, reference_mf, out, keep + gid);
        else
            fl = S::template record_to<D>(p, D(f_min[b * f_min_stride]), amp_scale, reference_mf, out);
    }
    __syncthreads();
    // the block's binaries' rows (contiguous), element e of row e / 8:
    // the derivatives in columns 0..3, zeros in the closed-form columns
    const int64_t first = (int64_t)blockIdx.x * (blockDim.x >> 2);
    const int64_t count = (nb - first < 16 ? nb - first : 16) * kSlots;
    double* rows = jac + first * kSlots * 8;
    for (int64_t e = threadIdx.x; e < count * 8; e += blockDim.x) {
        const int64_t row = e >> 3;
        const int c = (int)(e & 7);
        rows[e] = c < 4 ? s_dual[row / kSlots][row % kSlots][c] : 0.0;
    }
    // the barrier orders those zeros before the closed-form entries that
    // other threads of the block store over them
    __syncthreads();
    float(*s_split)[kSlots] = reinterpret_cast<float(*)[kSlots]>(&s_dual[0][0][0]);
    if (live) {
        double pv[kParams];
        for (int k = 0; k < kParams; ++k) pv[k] = params[b * kParams + k];
        S::closed_column(dir + 4, pv, reference_mf, fl, vals, jac + b * kSlots * 8 + dir + 4, 8);
        if (dir == 0) {
            double v[kSlots];
            for (int k = 0; k < kSlots; ++k) v[k] = vals[k];
            v[kT0Lo] = 0.0;
            v[kAnchorLo] = 0.0;
            if (split_records) write_f32_record(v, s_split[threadIdx.x >> 2], S::kFoldA, S::kFoldB);
            if (flags) flags[b] = fl;
        }
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < count; e += blockDim.x) {
        const int slot = (int)(e % kSlots);
        if (records)
            records[first * kSlots + e] = (slot == kT0Lo || slot == kAnchorLo) ? 0.0 : s_vals[e / kSlots][slot];
        if (split_records) split_records[first * kSlots + e] = s_split[e / kSlots][slot];
    }
}

__global__ __launch_bounds__(64) void reference_phase_kernel(const double* __restrict__ params, int64_t nb,
                                                             const double* __restrict__ mf,
                                                             double* __restrict__ phase) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double p[kParams];
    for (int k = 0; k < kParams; ++k) p[k] = params[b * kParams + k];
    phase[b] = lal_phase_22<double>(p, mf[b]);
}

// IMRPhenomD: the FP64 setup (phenomd_setup.h), written as the FP64 record
// or as the FP32 split record (the anchor folded into the intermediate and
// merger-ringdown constants); the table is FP64 whatever the evaluation's
// arithmetic, as the reference's (phenomd/model.py:33-50)
__global__ __launch_bounds__(64) void setup_pd_kernel(const double* __restrict__ params, int64_t nb,
                                                      const double* __restrict__ f_min, int64_t f_min_stride,
                                                      double amp_scale, double reference_mf,
                                                      double* __restrict__ records64, float* __restrict__ records32,
                                                      uint32_t* __restrict__ flags) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double p[kParams];
    for (int k = 0; k < kParams; ++k) p[k] = params[b * kParams + k];
    double rec[kSlots];
    const unsigned fl = pd_record<double>(p, f_min[b * f_min_stride], amp_scale, reference_mf, rec);
    rec[kT0Lo] = 0.0;
    rec[kAnchorLo] = 0.0;
    if (records64)
        for (int k = 0; k < kSlots; ++k) records64[b * kSlots + k] = rec[k];
    if (records32) write_f32_record(rec, records32 + b * kSlots, kPdIntConst, kPdMrdConst);
    if (flags) flags[b] = fl;
}

__global__ __launch_bounds__(64) void reference_phase_pd_kernel(const double* __restrict__ params, int64_t nb,
                                                                const double* __restrict__ mf,
                                                                double* __restrict__ phase) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double p[kParams];
    for (int k = 0; k < kParams; ++k) p[k] = params[b * kParams + k];
    phase[b] = pd_reference_phase<double>(p, mf[b]);
}

// the reference phase and its derivatives grad[B][5] with respect to (m1,
// m2, chi1, chi2, mf): one thread per (binary, direction) in dual numbers;
// Model 0 is IMRPhenomXAS's lal_phase_22, 1 IMRPhenomD's pd_reference_phase
template <int Model>
__global__ __launch_bounds__(64) void reference_phase_grad_kernel(const double* __restrict__ params, int64_t nb,
                                                                  const double* __restrict__ mf,
                                                                  double* __restrict__ phase,
                                                                  double* __restrict__ grad) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = gid / 5;
    const int dir = (int)(gid % 5);
    if (b >= nb) return;
    using D = Dual<1>;
    D p[kParams];
    for (int k = 0; k < kParams; ++k) {
        p[k] = D(params[b * kParams + k]);
        p[k].d[0] = k == dir ? 1.0 : 0.0;
    }
    D x(mf[b]);
    x.d[0] = dir == 4 ? 1.0 : 0.0;
    const D v = Model == 0 ? lal_phase_22<D>(p, x) : pd_reference_phase<D>(p, x);
    grad[b * 5 + dir] = v.d[0];
    if (dir == 0 && phase) phase[b] = v.v;
}

// ------------------------------------------------------------ evaluation

constexpr int kBlock = 256;

template <class F>
__device__ inline double load_freq(const F* freqs, int64_t i) {
    return (double)freqs[i];
}

// The models that share the evaluation kernels: a frequency's powers and
// the per-bin formula on a record, in FP64 and in FP32 (Lead: how the
// inspiral's leading term is carried).
struct XasModel {
    using Pow64 = FreqPowers<double>;
    using Pow32 = PowF32;
    __device__ static Pow64 powers64(double f) { return freq_powers<double>(f); }
    __device__ static void point64(const double* c, const Pow64& p, double& a, double& b, double& x, double& y) {
        point_f64(c, p, a, b, x, y);
    }
    __device__ static Pow32 powers32(float f) { return pow_f32(f); }
    template <int L>
    __device__ static void point32(const float* c, const Pow32& p, float2& hp, float2& hc) {
        point_f32<L>(c, p, hp, hc);
    }
};

struct PhenomDModel {
    using Pow64 = PdPowers<double>;
    using Pow32 = PdPowF32;
    __device__ static Pow64 powers64(double f) { return pd_powers<double>(f); }
    __device__ static void point64(const double* c, const Pow64& p, double& a, double& b, double& x, double& y) {
        point_pd_f64(c, p, a, b, x, y);
    }
    __device__ static Pow32 powers32(float f) { return pd_pow_f32(f); }
    template <int L>
    __device__ static void point32(const float* c, const Pow32& p, float2& hp, float2& hc) {
        point_pd_f32<L>(c, p, hp, hc);
    }
};

// shared grid, FP64
template <class M, class F>
__global__ __launch_bounds__(kBlock) void eval_f64_shared(const double* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, int64_t n, int64_t ntiles,
                                                          int group, double2* __restrict__ hp,
                                                          double2* __restrict__ hc) {
    const int64_t w = xcd_work_id(blockIdx.x, gridDim.x);
    const int64_t tile = w % ntiles;
    const int64_t b0 = (w / ntiles) * group;
    const int64_t i = tile * kBlock + threadIdx.x;
    const bool valid = i < n;
    const typename M::Pow64 fp = M::powers64(load_freq(freqs, valid ? i : n - 1));
    const int64_t b1 = b0 + group < nb ? b0 + group : nb;
    for (int64_t b = b0; b < b1; ++b) {
        double a, bb, c, d;
        M::point64(records + b * kSlots, fp, a, bb, c, d);
        if (valid) {
            hp[b * n + i] = make_double2(a, bb);
            hc[b * n + i] = make_double2(c, d);
        }
    }
}

// work ids per XCD chunk of the FP32 shared-grid evaluation (measured at
// 65,536 x 4017, 4 binaries x 4 frequencies per thread: 0.741-0.747 ms,
// chunks of 32 and 64 0.746-0.750 ms, one contiguous range per XCD
// 0.768-0.774 ms)
constexpr int64_t kWorkChunk = 128;

// shared grid, FP32 paths: a block owns kBlock * FPT frequencies (thread t
// takes t, t + kBlock, ...) and a group of binaries; each binary's record
// is read once per iteration (scalar loads) and serves FPT bins per thread.
//
// The shape trades three costs, measured on MI355X at 65,536 x 4017: the
// powers (amortised over the group), the record's scalar loads and their
// latency (amortised over FPT bins per thread), and the write pattern: the
// output streams best when the blocks running at one time on an XCD write a
// narrow, contiguous window of it. Group g therefore takes the binaries g,
// g + ngroups, g + 2 ngroups, ... (ngroups = ceil(B / group)): the XCD's
// blocks, which hold consecutive groups, write consecutive rows at each
// step of their loops, where groups of consecutive binaries would write
// rows `group` apart; and the XCDs take turns at chunks of work
// (xcd_chunk_work_id), so that all of them write near one another. 4
// binaries x 4 frequencies per thread balances the three costs.
template <class M, int L, int FPT, class F>
__global__ __launch_bounds__(kBlock) void eval_f32_shared(const float* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, int64_t n, int64_t ntiles,
                                                          int group, float2* __restrict__ hp,
                                                          float2* __restrict__ hc) {
    const int64_t w = xcd_chunk_work_id(blockIdx.x, gridDim.x, kWorkChunk);
    const int64_t tile = w % ntiles;
    const int64_t g = w / ntiles;
    const int64_t i0 = tile * (kBlock * FPT) + threadIdx.x;
    typename M::Pow32 p[FPT];
#pragma unroll
    for (int j = 0; j < FPT; ++j) {
        const int64_t i = i0 + j * kBlock;
        p[j] = M::powers32((float)freqs[i < n ? i : n - 1]);
    }
    // (a trip bound, not an exit inside the loop: the latter made the
    // compiler load the record with vector loads)
    const int64_t ngroups = (nb + group - 1) / group;
    const int64_t b1 = g + (int64_t)group * ngroups < nb ? g + (int64_t)group * ngroups : nb;
    for (int64_t b = g; b < b1; b += ngroups) {
        float c[kSlots];
        const float* rec = records + b * kSlots;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) c[k] = rec[k];
#pragma unroll
        for (int j = 0; j < FPT; ++j) {
            const int64_t i = i0 + j * kBlock;
            float2 a, x;
            M::template point32<L>(c, p[j], a, x);
            if (i < n) {
                hp[b * n + i] = a;
                hc[b * n + i] = x;
            }
        }
    }
}

// ragged grids: work item w -> (binary, tile) through the prefix sum of
// tiles per binary (tiles[b] .. tiles[b + 1]): the largest b with tiles[b]
// <= w, by a 64-ary search of the whole wave (the wave must be converged):
// each round lane l tests the candidate lo + l step, the lanes that pass
// form a prefix (tiles is non-decreasing), and the range shrinks to the last
// passing candidate's step. Three rounds of one vector load for 2^18
// binaries, where a binary search takes a chain of 18 dependent loads.
__device__ inline int64_t owner(const int64_t* __restrict__ tiles, int64_t nb, int64_t w) {
    int64_t lo = 0, hi = nb;  // tiles[lo] <= w < tiles[hi]
    const int lane = threadIdx.x & 63;
    while (hi - lo > 1) {
        const int64_t step = (hi - lo + 63) >> 6;
        const int64_t cand = lo + (int64_t)lane * step;
        const bool pass = cand < hi && tiles[cand] <= w;  // lane 0 (cand = lo) passes
        const int last = 63 - __clzll(__ballot(pass));
        lo += (int64_t)last * step;
        hi = lo + step < hi ? lo + step : hi;
    }
    return lo;
}

template <class F>
__device__ inline double ragged_freq(const F* freqs, const double* f0, const double* df, int64_t b, int64_t i,
                                     int64_t k) {
    return freqs ? (double)freqs[i] : f0[b] + (double)k * df[b];
}

// a work item is a tile of kBlock * BPT consecutive bins of one binary;
// thread t takes bins t, t + kBlock, ... of the tile
template <class M, int BPT, class F>
__global__ __launch_bounds__(kBlock) void eval_f64_ragged(const double* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, const double* __restrict__ f0,
                                                          const double* __restrict__ df,
                                                          const int64_t* __restrict__ offsets,
                                                          const int64_t* __restrict__ tiles,
                                                          double2* __restrict__ hp, double2* __restrict__ hc) {
    const int64_t total = tiles[nb];
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int64_t b = owner(tiles, nb, w);
        const int64_t start = offsets[b], stop = offsets[b + 1];
#pragma unroll
        for (int j = 0; j < BPT; ++j) {
            const int64_t k = (w - tiles[b]) * (kBlock * BPT) + j * kBlock + threadIdx.x;
            const int64_t i = start + k;
            if (i >= stop) break;
            const typename M::Pow64 fp = M::powers64(ragged_freq(freqs, f0, df, b, i, k));
            double a, bb, c, d;
            M::point64(records + b * kSlots, fp, a, bb, c, d);
            hp[i] = make_double2(a, bb);
            hc[i] = make_double2(c, d);
        }
    }
}

template <class M, int L, int BPT, class F>
__global__ __launch_bounds__(kBlock) void eval_f32_ragged(const float* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, const double* __restrict__ f0,
                                                          const double* __restrict__ df,
                                                          const int64_t* __restrict__ offsets,
                                                          const int64_t* __restrict__ tiles,
                                                          float2* __restrict__ hp, float2* __restrict__ hc) {
    const int64_t total = tiles[nb];
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int64_t b = owner(tiles, nb, w);
        const int64_t start = offsets[b], stop = offsets[b + 1];
        float rc[kSlots];
        const float* rec = records + b * kSlots;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) rc[k] = rec[k];
        // the tile's first bin and the number of its bins the binary has:
        // the per-bin index arithmetic is then 32-bit, the outputs'
        // addresses a wave-uniform base plus a 32-bit offset
        const int64_t first = (w - tiles[b]) * (kBlock * BPT);
        const int64_t left = stop - start - first;
        const int count = left < kBlock * BPT ? (int)left : kBlock * BPT;
        const double first_d = (double)first;
        float2* __restrict__ out_p = hp + start + first;
        float2* __restrict__ out_c = hc + start + first;
        const F* __restrict__ tile_f = freqs ? freqs + start + first : nullptr;
        const double g0 = freqs ? 0.0 : f0[b], gd = freqs ? 0.0 : df[b];
#pragma unroll
        for (int j = 0; j < BPT; ++j) {
            const int k = j * kBlock + (int)threadIdx.x;
            if (k >= count) break;
            // f0 + k df as ragged_freq computes it (the bin's index first +
            // k, exact in FP64)
            const double f = tile_f ? (double)tile_f[k] : g0 + (first_d + (double)k) * gd;
            const typename M::Pow32 p = M::powers32((float)f);
            float2 a, c;
            M::template point32<L>(rc, p, a, c);
            out_p[k] = a;
            out_c[k] = c;
        }
    }
}

// tiles[b] = sum_{j < b} ceil(len_j / tile): one block, a running prefix
// over chunks of 1024 binaries
__global__ __launch_bounds__(1024) void tile_prefix_kernel(const int64_t* __restrict__ offsets, int64_t nb,
                                                           int64_t tile, int64_t* __restrict__ tiles) {
    __shared__ int64_t scan[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t b = base + threadIdx.x;
        int64_t len = b < nb ? offsets[b + 1] - offsets[b] : 0;
        if (len < 0) len = 0;
        scan[threadIdx.x] = (len + tile - 1) / tile;
        __syncthreads();
        for (int s = 1; s < 1024; s <<= 1) {
            const int64_t v = threadIdx.x >= (unsigned)s ? scan[threadIdx.x - s] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (b < nb) tiles[b + 1] = carry + scan[threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 1023) carry += scan[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) tiles[0] = 0;
}

// fast_atan elementwise (its accuracy test's entry point)
__global__ __launch_bounds__(kBlock) void fast_atan_kernel(const float* __restrict__ x, float* __restrict__ y,
                                                           int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = fast_atan(x[i]);
}

// ---------------------------------------------------------------- adjoint

// dL/dtheta of each binary: one block per binary walks its frequencies,
// each thread accumulating dL/d(record slot) of its bins (FP32 or FP64 as
// the evaluation); the block then reduces the slots in FP64 and contracts
// them with the record's derivatives jac (B, 64, 8) from the FP64 setup.
// kLoss: the cotangents come from L = sum_f w(f) (|h+ - d+|^2 + |hx - dx|^2)
// (data d, weights w), whose per-binary value is written to loss.
constexpr int kFromCotangents = 0;
constexpr int kLoss = 1;

// Cross-lane moves of one wave for the reductions of the adjoint kernel.
// swap32(a, b): v_permlane32_swap, lanes 32-63 of a trade places with
// lanes 0-31 of b; swap16(a, b): v_permlane16_swap, the odd 16-lane rows of
// a with the even rows of b. dpp<Ctrl>(x): x of the lane that the DPP
// control names (row_ror:8 = 0x128, row_half_mirror = 0x141, quad_perm
// [2,3,0,1] = 0x4e, [1,0,3,2] = 0xb1). Doubles move as two 32-bit halves.
__device__ inline void swap32(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ inline void swap16(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ inline void swap32(double& a, double& b) {
    const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ inline void swap16(double& a, double& b) {
    const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}
template <int Ctrl>
__device__ inline float dpp(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), Ctrl, 0xF, 0xF, false));
}
template <int Ctrl>
__device__ inline double dpp(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, Ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), Ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// One level of a reduce-scatter over the wave: of each pair (v[k], v[k+H]),
// k < H, the lanes with bit Bit of their index clear keep the first and the
// others the second, each adding its partner's copy (the partner differs in
// that bit and agrees in the higher ones).
template <int Bit, int H, class T>
__device__ inline void scatter_level(T* v, int lane) {
    if constexpr (Bit == 32) {
#pragma unroll
        for (int k = 0; k < H; ++k) {
            swap32(v[k], v[k + H]);
            v[k] += v[k + H];
        }
    } else if constexpr (Bit == 16) {
#pragma unroll
        for (int k = 0; k < H; ++k) {
            swap16(v[k], v[k + H]);
            v[k] += v[k + H];
        }
    } else {
        constexpr int ctrl = Bit == 8 ? 0x128 : Bit == 4 ? 0x141 : Bit == 2 ? 0x4e : 0xb1;
        const bool high = lane & Bit;
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const T keep = high ? v[k + H] : v[k];
            const T send = high ? v[k] : v[k + H];
            v[k] = keep + dpp<ctrl>(send);
        }
    }
}

// v[0..63] per lane in, the sum over the wave's 64 lanes of v[lane] out
// (in v[0]: each level halves the live values and lane L keeps index L).
// Every lane of the wave must be active.
template <class T>
__device__ inline T reduce_scatter64(T* v, int lane) {
    scatter_level<32, 32>(v, lane);
    scatter_level<16, 16>(v, lane);
    scatter_level<8, 8>(v, lane);
    scatter_level<4, 4>(v, lane);
    scatter_level<2, 2>(v, lane);
    scatter_level<1, 1>(v, lane);
    return v[0];
}

// sum over the wave of s[lane] * jac_row[lane][d] for the 8 parameters d:
// the value of parameter lane / 8, in every lane of its group of 8
__device__ inline double contract8(double s, const double* __restrict__ jac_row, int lane) {
    double p[8];
    // (an exactly zero derivative contributes nothing, whatever s: as the
    // reference's derivatives through jnp.where, for a binary whose loss is
    // NaN)
#pragma unroll
    for (int d = 0; d < 8; ++d) p[d] = jac_row[d] == 0.0 ? 0.0 : s * jac_row[d];
    scatter_level<32, 4>(p, lane);
    scatter_level<16, 2>(p, lane);
    scatter_level<8, 1>(p, lane);
    double g = p[0];
    g += dpp<0x141>(g);
    g += dpp<0x4e>(g);
    g += dpp<0xb1>(g);
    return g;
}

template <class T>
__device__ inline BinPowers<T> adjoint_powers(double f);
template <>
__device__ inline BinPowers<float> adjoint_powers<float>(double f) {
    return bin_powers(pow_f32((float)f));
}
template <>
__device__ inline BinPowers<double> adjoint_powers<double>(double f) {
    return bin_powers(freq_powers<double>(f));
}

// the powers of frequency i: from the grid's table if there is one
template <class T, class F>
__device__ inline BinPowers<T> powers_at(const F* __restrict__ freqs, const PowerRow<T>* __restrict__ table,
                                         int64_t i) {
    if (table) {
        const PowerRow<T> r = table[i];
        return BinPowers<T>{r.f, r.t, r.m1, r.lf, r.m76, r.m53};
    }
    return adjoint_powers<T>((double)freqs[i]);
}

template <class T, class F>
__global__ __launch_bounds__(kBlock) void power_table_kernel(const F* __restrict__ freqs, int64_t n,
                                                             PowerRow<T>* __restrict__ table) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BinPowers<T> p = adjoint_powers<T>((double)freqs[i]);
    table[i] = PowerRow<T>{p.f, p.t, p.m1, p.lf, p.m76, p.m53};
}

// One wave per binary: lane l takes bins l, l + 64, ...; each lane sums its
// bins' dL/d(record slot) in registers, one reduce-scatter over the wave
// leaves slot l's total in lane l, and the wave contracts the totals with
// the record's derivatives. The loss rides in the slot of T0's low part,
// which carries no derivative of its own (T0 enters as hi + lo, accounted
// on kT0).
constexpr int kAdjointBlock = 64;
constexpr int kLossSlot = kT0Lo;

// The models' per-bin formula and its adjoint, for adjoint_kernel: forward
// gives phi and A of one bin, adjoint adds the bin's dL/d(record slot).
struct XasAdjoint {
    static constexpr int kKpSlot = kKp, kKcSlot = kKc;
    template <class T>
    __device__ __forceinline__ static T forward(const T* c, const BinPowers<T>& p, T& amp) {
        bool past, ringdown;
        T z;
        return phase_and_amp(c, p, past, ringdown, z, amp);
    }
    template <class T>
    __device__ __forceinline__ static void adjoint(const T* c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re,
                                                   T gx_im, T* acc) {
        adjoint_point(c, p, gp_re, gp_im, gx_re, gx_im, acc);
    }
    template <class T>
    __device__ __forceinline__ static T grid(const T* c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re, T gx_im) {
        return grid_adjoint_point(c, p, gp_re, gp_im, gx_re, gx_im);
    }
};

struct PhenomDAdjoint {
    static constexpr int kKpSlot = kPdKp, kKcSlot = kPdKc;
    template <class T>
    __device__ __forceinline__ static T forward(const T* c, const BinPowers<T>& p, T& amp) {
        bool inspiral, mrd;
        T z;
        return pd_phase_and_amp(c, p, pd_f34(p.f), inspiral, mrd, z, amp);
    }
    template <class T>
    __device__ __forceinline__ static void adjoint(const T* c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re,
                                                   T gx_im, T* acc) {
        pd_adjoint_point(c, p, gp_re, gp_im, gx_re, gx_im, acc);
    }
    template <class T>
    __device__ __forceinline__ static T grid(const T* c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re, T gx_im) {
        return pd_grid_adjoint_point(c, p, gp_re, gp_im, gx_re, gx_im);
    }
};

template <class A, class T, int Mode, class F>
__global__ __launch_bounds__(kAdjointBlock) void adjoint_kernel(const T* __restrict__ records,
                                                                const double* __restrict__ jac, int64_t nb,
                                                                const F* __restrict__ freqs, int64_t n,
                                                                const T* __restrict__ in_plus,
                                                                const T* __restrict__ in_cross,
                                                                const T* __restrict__ weights,
                                                                double* __restrict__ grads, double* __restrict__ loss,
                                                                const PowerRow<T>* __restrict__ table) {
    const int64_t b = blockIdx.x;
    if (b >= nb) return;
    const int lane = threadIdx.x;
    T c[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) c[k] = records[b * kSlots + k];
    T acc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) acc[k] = T(0);
    // IMRPhenomXAS's fused loss in FP32: each iteration loads the next one's
    // powers, inputs and weight before its own arithmetic, so that their
    // latency overlaps it (0.326 -> 0.307 ms for the loss and gradient at
    // 8192 x 4017, the same bits; the other variants lose a wave per SIMD to
    // the registers it takes: IMRPhenomD's loss 2% slower, FP64 has none to
    // spare)
    constexpr bool kAhead =
        std::is_same<T, float>::value && std::is_same<A, XasAdjoint>::value && Mode == kLoss;
    PowerRow<T> next_row{};
    T next_in[4] = {T(0), T(0), T(0), T(0)};
    T next_w = T(0);
    auto fetch = [&](int64_t j) __attribute__((always_inline)) {
        if (table) next_row = table[j];
        const int64_t o = 2 * (b * n + j);
        next_in[0] = in_plus[o];
        next_in[1] = in_plus[o + 1];
        next_in[2] = in_cross[o];
        next_in[3] = in_cross[o + 1];
        if (Mode == kLoss) next_w = weights[j];
    };
    if (kAhead && lane < n) fetch(lane);
    for (int64_t i = lane; i < n; i += kAdjointBlock) {
        BinPowers<T> p;
        T gp_re, gp_im, gx_re, gx_im, w_i = T(0);
        if constexpr (kAhead) {
            p = table ? BinPowers<T>{next_row.f, next_row.t, next_row.m1, next_row.lf, next_row.m76, next_row.m53}
                      : adjoint_powers<T>((double)freqs[i]);
            gp_re = next_in[0];
            gp_im = next_in[1];
            gx_re = next_in[2];
            gx_im = next_in[3];
            w_i = next_w;
            if (i + kAdjointBlock < n) fetch(i + kAdjointBlock);
        } else {
            p = powers_at<T>(freqs, table, i);
            const int64_t o = 2 * (b * n + i);
            gp_re = in_plus[o];
            gp_im = in_plus[o + 1];
            gx_re = in_cross[o];
            gx_im = in_cross[o + 1];
            if (Mode == kLoss) w_i = weights[i];
        }
        if (Mode == kLoss) {
            T amp;
            const T phi = A::forward(c, p, amp);
            T s, co;
            m_sincos(phi, &s, &co);
            const T w = w_i;
            const T ap = c[A::kKpSlot] * amp, ac = c[A::kKcSlot] * amp;
            const T rp_re = ap * co - gp_re, rp_im = ap * s - gp_im;
            const T rx_re = -ac * s - gx_re, rx_im = ac * co - gx_im;
            acc[kLossSlot] += w * (rp_re * rp_re + rp_im * rp_im + rx_re * rx_re + rx_im * rx_im);
            const T two_w = T(2) * w;
            gp_re = two_w * rp_re;
            gp_im = two_w * rp_im;
            gx_re = two_w * rx_re;
            gx_im = two_w * rx_im;
        }
        A::adjoint(c, p, gp_re, gp_im, gx_re, gx_im, acc);
    }
    double s = (double)reduce_scatter64(acc, lane);
    if (Mode == kLoss && lane == kLossSlot) loss[b] = s;
    // slots without a derivative: T0's and A's low parts (the FP32 record's
    // T0 and A enter in FP64 through hi + lo) and the FP64 copies 60..63
    if (lane == kT0Lo || lane == kAnchorLo || lane >= 60) s = 0.0;
    const double g = contract8(s, jac + (b * kSlots + lane) * 8, lane);
    if ((lane & 7) == 0) grads[b * 8 + (lane >> 3)] = g;
}

// dL/df of each frequency of a shared grid: the evaluation's launch shape
// (a thread per frequency, a loop over a group of kGridGroup binaries), the
// bins' contributions summed in FP64 in a register in the binaries' order,
// then partial[g][i] for group g; group_sum_kernel adds the groups in
// order (no atomics: the same bits on every run)
constexpr int kGridGroup = 64;
template <class A, class T, class F>
__global__ __launch_bounds__(kBlock) void grid_adjoint_kernel(const T* __restrict__ records, int64_t nb,
                                                              const F* __restrict__ freqs, int64_t n, int64_t ntiles,
                                                              const T* __restrict__ cot_plus,
                                                              const T* __restrict__ cot_cross,
                                                              double* __restrict__ partial) {
    const int64_t w = blockIdx.x;
    const int64_t tile = w % ntiles;
    const int64_t g = w / ntiles;
    const int64_t b0 = g * kGridGroup;
    const int64_t i = tile * kBlock + threadIdx.x;
    if (i >= n) return;
    const BinPowers<T> p = adjoint_powers<T>((double)freqs[i]);
    const int64_t b1 = b0 + kGridGroup < nb ? b0 + kGridGroup : nb;
    double acc = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
        const int64_t o = 2 * (b * n + i);
        acc += (double)A::grid(records + b * kSlots, p, cot_plus[o], cot_plus[o + 1], cot_cross[o],
                               cot_cross[o + 1]);
    }
    partial[g * n + i] = acc;
}

// out[i] = the sum over g < groups of partial[g * n + i], in order
__global__ __launch_bounds__(kBlock) void group_sum_kernel(const double* __restrict__ partial, int64_t groups,
                                                           int64_t n, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t g = 0; g < groups; ++g) s += partial[g * n + i];
    out[i] = s;
}

// the grid adjoint's workspace: one double per (group of binaries, frequency)
inline int64_t grid_adjoint_bytes(int64_t nb, int64_t n) {
    const int64_t groups = (nb + kGridGroup - 1) / kGridGroup;
    if (n > 0 && groups > INT64_MAX / 8 / n) return -1;
    return groups * n * 8;
}

struct AdjointArgs {
    const void* records;
    const double* jac;
    int64_t nb;
    const void* freqs;
    int64_t n;
    const void* in_plus;
    const void* in_cross;
    const void* weights;
    double* grads;
    double* loss;
    const void* powers;
    int blocks;
    hipStream_t stream;
};

template <class A, class T, int Mode, class F>
void launch_adjoint_t(const AdjointArgs& a) {
    adjoint_kernel<A, T, Mode, F><<<a.blocks, kAdjointBlock, 0, a.stream>>>(
        (const T*)a.records, a.jac, a.nb, (const F*)a.freqs, a.n, (const T*)a.in_plus, (const T*)a.in_cross,
        (const T*)a.weights, a.grads, a.loss, (const PowerRow<T>*)a.powers);
}

template <class A, int Mode>
bool launch_adjoint(const AdjointArgs& a, int precision, int freqs_f64) {
    if (precision == kPrecF64) {
        freqs_f64 ? launch_adjoint_t<A, double, Mode, double>(a) : launch_adjoint_t<A, double, Mode, float>(a);
        return true;
    }
    if (precision == kPrecF32 || precision == kPrecSplit || precision == kPrecSplitDF) {
        freqs_f64 ? launch_adjoint_t<A, float, Mode, double>(a) : launch_adjoint_t<A, float, Mode, float>(a);
        return true;
    }
    return false;
}

// the FP32 shared-grid evaluation's default shape (see eval_f32_shared)
constexpr int kDefaultGroup = 4;
// the FP64 evaluation's (its frequency powers are FP64 transcendentals:
// shared by more binaries)
constexpr int kDefaultGroupF64 = 8;
constexpr int kDefaultFreqsPerThread = 4;

struct SharedArgs {
    const void* records;
    int64_t nb;
    const void* freqs;
    int64_t n;
    int64_t ntiles;
    int group;
    void* hp;
    void* hc;
    int blocks;
    hipStream_t stream;
};

template <class M, int L, int FPT, class F>
void launch_f32_shared_t(const SharedArgs& a) {
    eval_f32_shared<M, L, FPT, F><<<a.blocks, kBlock, 0, a.stream>>>(
        (const float*)a.records, a.nb, (const F*)a.freqs, a.n, a.ntiles, a.group, (float2*)a.hp, (float2*)a.hc);
}

template <class M, int L>
void launch_f32_shared(const SharedArgs& a, int freqs_f64, int fpt) {
    if (freqs_f64) {
        if (fpt == 1) launch_f32_shared_t<M, L, 1, double>(a);
        else if (fpt == 2) launch_f32_shared_t<M, L, 2, double>(a);
        else if (fpt == 4) launch_f32_shared_t<M, L, 4, double>(a);
        else launch_f32_shared_t<M, L, 8, double>(a);
    } else {
        if (fpt == 1) launch_f32_shared_t<M, L, 1, float>(a);
        else if (fpt == 2) launch_f32_shared_t<M, L, 2, float>(a);
        else if (fpt == 4) launch_f32_shared_t<M, L, 4, float>(a);
        else launch_f32_shared_t<M, L, 8, float>(a);
    }
}

struct RaggedArgs {
    const void* records;
    int64_t nb;
    const void* freqs;
    const double* f0;
    const double* df;
    const int64_t* offsets;
    const int64_t* tiles;
    void* hp;
    void* hc;
    int blocks;
    hipStream_t stream;
};

template <class M, int BPT, class F>
void launch_ragged(const RaggedArgs& a, int precision) {
    const F* freqs = (const F*)a.freqs;
    if (precision == kPrecF64) {
        eval_f64_ragged<M, BPT, F><<<a.blocks, kBlock, 0, a.stream>>>(
            (const double*)a.records, a.nb, freqs, a.f0, a.df, a.offsets, a.tiles, (double2*)a.hp, (double2*)a.hc);
        return;
    }
    const float* rec = (const float*)a.records;
    float2* hp = (float2*)a.hp;
    float2* hc = (float2*)a.hc;
    if (precision == kPrecSplit)
        eval_f32_ragged<M, kLeadF64, BPT, F><<<a.blocks, kBlock, 0, a.stream>>>(rec, a.nb, freqs, a.f0, a.df,
                                                                               a.offsets, a.tiles, hp, hc);
    else if (precision == kPrecSplitDF)
        eval_f32_ragged<M, kLeadDF, BPT, F><<<a.blocks, kBlock, 0, a.stream>>>(rec, a.nb, freqs, a.f0, a.df,
                                                                              a.offsets, a.tiles, hp, hc);
    else
        eval_f32_ragged<M, kLeadF32, BPT, F><<<a.blocks, kBlock, 0, a.stream>>>(rec, a.nb, freqs, a.f0, a.df,
                                                                               a.offsets, a.tiles, hp, hc);
}

int grid_blocks(int64_t nblocks, const char* what) {
    if (nblocks <= 0) return 0;
    if (nblocks > 0x7fffffff) {
        fail(what);
        return -1;
    }
    return (int)nblocks;
}

// the shared-grid evaluation of model M behind its C entry point `name`
template <class M>
int eval_shared(const char* name, const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                int64_t n_freqs, int precision, int group, int freqs_per_thread, void* hplus, void* hcross,
                void* stream) {
    const std::string who(name);
    if (n_binaries < 0 || n_freqs < 0) return fail((who + ": negative size").c_str());
    if (n_binaries == 0 || n_freqs == 0) return 0;
    if (!records || !freqs || !hplus || !hcross) return fail((who + ": null pointer").c_str());
    if (precision < kPrecF64 || precision > kPrecSplitDF) return fail((who + ": unknown precision").c_str());
    if (freqs_per_thread != 0 && freqs_per_thread != 1 && freqs_per_thread != 2 && freqs_per_thread != 4 &&
        freqs_per_thread != 8)
        return fail((who + ": freqs_per_thread must be 1, 2, 4 or 8 (0: default)").c_str());
    if (precision == kPrecF64) freqs_per_thread = 1;
    if (freqs_per_thread == 0) freqs_per_thread = kDefaultFreqsPerThread;
    if (group <= 0) group = precision == kPrecF64 ? kDefaultGroupF64 : kDefaultGroup;
    const int64_t ntiles = (n_freqs + kBlock * freqs_per_thread - 1) / (kBlock * freqs_per_thread);
    const int64_t ngroups = (n_binaries + group - 1) / group;
    const int blocks = grid_blocks(ntiles * ngroups, (who + ": grid too large, raise the group size").c_str());
    if (blocks < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    const SharedArgs a{records, n_binaries, freqs, n_freqs, ntiles, group, hplus, hcross, blocks, s};
    switch (precision) {
        case kPrecF64:
            if (freqs_f64)
                eval_f64_shared<M, double><<<blocks, kBlock, 0, s>>>((const double*)records, n_binaries,
                                                                     (const double*)freqs, n_freqs, ntiles, group,
                                                                     (double2*)hplus, (double2*)hcross);
            else
                eval_f64_shared<M, float><<<blocks, kBlock, 0, s>>>((const double*)records, n_binaries,
                                                                    (const float*)freqs, n_freqs, ntiles, group,
                                                                    (double2*)hplus, (double2*)hcross);
            break;
        case kPrecSplit:
            launch_f32_shared<M, kLeadF64>(a, freqs_f64, freqs_per_thread);
            break;
        case kPrecSplitDF:
            launch_f32_shared<M, kLeadDF>(a, freqs_f64, freqs_per_thread);
            break;
        default:
            launch_f32_shared<M, kLeadF32>(a, freqs_f64, freqs_per_thread);
            break;
    }
    return check_launch(name);
}

// the ragged-grid evaluation of model M behind its C entry point `name`
template <class M>
int eval_ragged(const char* name, const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                const double* grid_f0, const double* grid_df, const int64_t* offsets, const int64_t* tiles,
                int precision, int bins_per_thread, int blocks, void* hplus, void* hcross, void* stream) {
    const std::string who(name);
    if (n_binaries < 0) return fail((who + ": negative size").c_str());
    if (n_binaries == 0) return 0;
    if (!records || !offsets || !tiles || !hplus || !hcross) return fail((who + ": null pointer").c_str());
    if (!freqs && (!grid_f0 || !grid_df)) return fail((who + ": give freqs or (grid_f0, grid_df)").c_str());
    if (bins_per_thread != 1 && bins_per_thread != 2 && bins_per_thread != 4 && bins_per_thread != 8 &&
        bins_per_thread != 16)
        return fail((who + ": bins_per_thread must be 1, 2, 4, 8 or 16").c_str());
    if (precision < kPrecF64 || precision > kPrecSplitDF) return fail((who + ": unknown precision").c_str());
    if (blocks <= 0) blocks = 8 * cu_count();
    const RaggedArgs a{records, n_binaries, freqs, grid_f0, grid_df, offsets, tiles, hplus, hcross, blocks,
                       (hipStream_t)stream};
    const bool f64 = freqs_f64 || !freqs;
    switch (bins_per_thread) {
        case 1:
            f64 ? launch_ragged<M, 1, double>(a, precision) : launch_ragged<M, 1, float>(a, precision);
            break;
        case 2:
            f64 ? launch_ragged<M, 2, double>(a, precision) : launch_ragged<M, 2, float>(a, precision);
            break;
        case 4:
            f64 ? launch_ragged<M, 4, double>(a, precision) : launch_ragged<M, 4, float>(a, precision);
            break;
        case 8:
            f64 ? launch_ragged<M, 8, double>(a, precision) : launch_ragged<M, 8, float>(a, precision);
            break;
        default:
            f64 ? launch_ragged<M, 16, double>(a, precision) : launch_ragged<M, 16, float>(a, precision);
            break;
    }
    return check_launch(name);
}

// the derivative setup of model S behind its C entry point `name`
template <class S>
int setup_jvp(const char* name, const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
              double amp_scale, double reference_mf, double* records, float* split_records, double* jacobian,
              uint32_t* flags, void* stream) {
    const std::string who(name);
    if (n_binaries < 0) return fail((who + ": negative number of binaries").c_str());
    if (n_binaries == 0) return 0;
    if (!params || !f_min || !jacobian) return fail((who + ": null pointer").c_str());
    if (!records && !split_records) return fail((who + ": no record output").c_str());
    if (f_min_stride != 0 && f_min_stride != 1) return fail((who + ": f_min_stride must be 0 or 1").c_str());
    const int blocks = grid_blocks((4 * n_binaries + 63) / 64, (who + ": too many binaries").c_str());
    if (blocks < 0) return -1;
    setup_jvp_kernel<S><<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, f_min, f_min_stride, amp_scale,
                                                                reference_mf, records, split_records, jacobian, flags);
    return check_launch(name);
}

// the per-binary adjoint of model A behind its C entry point `name`: from
// the cotangents (Mode kFromCotangents; in_plus/in_cross are the
// cotangents) or from the noise-weighted loss against data (Mode kLoss;
// in_plus/in_cross are the data, loss receives the loss)
template <class A, int Mode>
int adjoint(const char* name, const void* records, const double* jacobian, int64_t n_binaries, const void* freqs,
            int freqs_f64, int64_t n_freqs, int precision, const void* in_plus, const void* in_cross,
            const void* weights, const void* powers, double* loss, double* grads, void* stream) {
    const std::string who(name);
    if (n_binaries < 0 || n_freqs < 0) return fail((who + ": negative size").c_str());
    if (n_binaries == 0) return 0;
    const bool loss_mode = Mode == kLoss;
    if (!records || !jacobian || !grads || (loss_mode && !loss) ||
        (n_freqs && (!freqs || !in_plus || !in_cross || (loss_mode && !weights))))
        return fail((who + ": null pointer").c_str());
    const int blocks = grid_blocks(n_binaries, (who + ": too many binaries").c_str());
    if (blocks < 0) return -1;
    const AdjointArgs a{records, jacobian, n_binaries, freqs,  n_freqs, in_plus, in_cross, weights,
                        grads,   loss,     powers,     blocks, (hipStream_t)stream};
    if (!launch_adjoint<A, Mode>(a, precision, freqs_f64)) return fail((who + ": unknown precision").c_str());
    return check_launch(name);
}

// dL/df of a shared grid for model A behind its C entry point `name`
template <class A>
int grid_adjoint(const char* name, const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                 int64_t n_freqs, int precision, const void* cot_plus, const void* cot_cross, double* dldf,
                 void* workspace, int64_t workspace_bytes, void* stream) {
    const std::string who(name);
    if (n_binaries < 0 || n_freqs < 0) return fail((who + ": negative size").c_str());
    if (n_freqs == 0) return 0;
    if (!dldf) return fail((who + ": null pointer").c_str());
    hipStream_t s = (hipStream_t)stream;
    const int64_t ntiles = (n_freqs + kBlock - 1) / kBlock;
    const int out_blocks = grid_blocks(ntiles, (who + ": too large").c_str());
    if (out_blocks < 0) return -1;
    if (n_binaries == 0) {
        // no binary: a zero gradient (a kernel, not a memset node)
        group_sum_kernel<<<out_blocks, kBlock, 0, s>>>(nullptr, 0, n_freqs, dldf);
        return check_launch(name);
    }
    if (!records || !freqs || !cot_plus || !cot_cross) return fail((who + ": null pointer").c_str());
    const int64_t need = grid_adjoint_bytes(n_binaries, n_freqs);
    if (need < 0) return fail((who + ": too large").c_str());
    if (!workspace || workspace_bytes < need || (uintptr_t)workspace % 256)
        return fail((who + ": the workspace must be 256-aligned and hold wfx_xas_grid_adjoint_workspace() bytes")
                        .c_str());
    const int64_t groups = (n_binaries + kGridGroup - 1) / kGridGroup;
    const int blocks = grid_blocks(ntiles * groups, (who + ": too large").c_str());
    if (blocks < 0) return -1;
    double* partial = (double*)workspace;
    if (precision == kPrecF64) {
        if (freqs_f64)
            grid_adjoint_kernel<A, double, double><<<blocks, kBlock, 0, s>>>(
                (const double*)records, n_binaries, (const double*)freqs, n_freqs, ntiles, (const double*)cot_plus,
                (const double*)cot_cross, partial);
        else
            grid_adjoint_kernel<A, double, float><<<blocks, kBlock, 0, s>>>(
                (const double*)records, n_binaries, (const float*)freqs, n_freqs, ntiles, (const double*)cot_plus,
                (const double*)cot_cross, partial);
    } else if (precision == kPrecF32 || precision == kPrecSplit || precision == kPrecSplitDF) {
        if (freqs_f64)
            grid_adjoint_kernel<A, float, double><<<blocks, kBlock, 0, s>>>(
                (const float*)records, n_binaries, (const double*)freqs, n_freqs, ntiles, (const float*)cot_plus,
                (const float*)cot_cross, partial);
        else
            grid_adjoint_kernel<A, float, float><<<blocks, kBlock, 0, s>>>(
                (const float*)records, n_binaries, (const float*)freqs, n_freqs, ntiles, (const float*)cot_plus,
                (const float*)cot_cross, partial);
    } else {
        return fail((who + ": unknown precision").c_str());
    }
    group_sum_kernel<<<out_blocks, kBlock, 0, s>>>(partial, groups, n_freqs, dldf);
    return check_launch(name);
}

// ------------------------------------------------------------ IMRPhenomXHM

// f(std::integral_constant<int, k>) for the higher mode k (0..3): the mode's
// set-up inlined into f then sees k, and every branch and array index that
// follows from it, as a constant
template <class F>
__device__ __forceinline__ void with_mode(int k, F f) {
    switch (k) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        default: f(std::integral_constant<int, 3>{}); break;
    }
}

// one FP64 thread per (binary, group), five groups: the four higher modes'
// records (group k < 4) and the (2,2)'s IMRPhenomXAS record with the flags
// and the binary's head (the factors of the five modes in the
// polarizations, NaN where LAL fails). The group is uniform over a block
// (blocks [k P, (k + 1) P), P = ceil(nb / 64)), so a wave runs one mode's
// branch of the setup rather than each in turn.
__global__ __launch_bounds__(64) void xhm_setup_kernel(const double* __restrict__ params, int64_t nb,
                                                       const double* __restrict__ f_min, int64_t f_min_stride,
                                                       double amp_scale, double reference_mf, int split,
                                                       void* __restrict__ records22, uint32_t* __restrict__ flags,
                                                       void* __restrict__ head, void* __restrict__ modes) {
    const int64_t per_group = (nb + 63) / 64;
    const int k = (int)(blockIdx.x / per_group);
    const int64_t b = (int64_t)(blockIdx.x % per_group) * 64 + threadIdx.x;
    if (k > kHmModes || b >= nb) return;
    double p[kParams];
    for (int i = 0; i < kParams; ++i) p[i] = params[b * kParams + i];
    with_reference_mf(p, reference_mf);
    if (k == kHmModes) {
        double rec[kSlots];
        const unsigned fl = xas_record<double>(p, f_min[b * f_min_stride], amp_scale, rec);
        if (split) {
            write_f32_record(rec, (float*)records22 + b * kSlots);
        } else {
            rec[kT0Lo] = 0.0;
            rec[kAnchorLo] = 0.0;
            for (int s = 0; s < kSlots; ++s) ((double*)records22)[b * kSlots + s] = rec[s];
        }
        if (flags) flags[b] = fl;
        double h[kHeadSlots];
        for (int s = 0; s < kHeadSlots; ++s) h[s] = 0.0;
        const double inclination = p[kInclination];
        mode_factors(inclination, 2, 2, h + kHeadFactors);
        for (int m = 0; m < kHmModes; ++m) mode_factors(inclination, kHmEll[m], kHmEmm[m], h + kHeadFactors + 4 * (m + 1));
        if (fl)
            for (int s = 0; s < 20; ++s) h[kHeadFactors + s] = __builtin_nan("");
        for (int s = 0; s < kHeadSlots; ++s) {
            if (split)
                ((float*)head)[b * kHeadSlots + s] = (float)h[s];
            else
                ((double*)head)[b * kHeadSlots + s] = h[s];
        }
        return;
    }
    const WaveformVars<double> wf = waveform_vars(p, f_min[b * f_min_stride], amp_scale);
    const AmpCoeffs<double> amp22 = amp_coeffs(wf);
    const PhaseCoeffs<double> ph22 = phase_coeffs(wf);
    HmCoeffs<double> c;
    hm_setup(wf, amp22, ph22, k, c);
    const int64_t at = (b * kHmModes + k) * kHmSlots;
    if (split)
        write_mode_record_split(c, wf, (float*)modes + at);
    else
        write_mode_record(c, wf, (double*)modes + at);
}

// The higher modes' records with their derivatives in the masses and
// spins, dual numbers through the whole set-up: one thread per (binary,
// mode, direction), the (mode, direction) uniform over a block (blocks
// [g P, (g + 1) P), g = 4 k + direction, P = ceil(nb / 64)); the 17th group
// writes the head and its derivative in the inclination (one thread per
// binary). The records in the split layout, the gradient's: float for the
// split arithmetic, double for FP64, with the tangents of their slots;
// dmodes (B, 4 modes, kHmSlots, 4 directions), double or float as the
// records. The (2,2) and its derivatives come from the
// IMRPhenomXAS derivative setup (the same record).
__global__ __launch_bounds__(64) void xhm_setup_jvp_kernel(const double* __restrict__ params, int64_t nb,
                                                           const double* __restrict__ f_min, int64_t f_min_stride,
                                                           double amp_scale, double reference_mf, int split,
                                                           const XasCoeffs22<Dual<1>>* __restrict__ kept,
                                                           void* __restrict__ head, void* __restrict__ dhead,
                                                           void* __restrict__ modes, void* __restrict__ dmodes) {
    const int64_t per_group = (nb + 63) / 64;
    const int g = (int)(blockIdx.x / per_group);
    const int64_t b = (int64_t)(blockIdx.x % per_group) * 64 + threadIdx.x;
    if (g > 4 * kHmModes || b >= nb) return;
    if (g == 4 * kHmModes) {
        double p[kParams];
        for (int i = 0; i < kParams; ++i) p[i] = params[b * kParams + i];
        with_reference_mf(p, reference_mf);
        const unsigned fl = waveform_vars(p, f_min[b * f_min_stride], amp_scale).flags;
        using D = Dual<1>;
        D inclination(p[kInclination]);
        inclination.d[0] = 1.0;
        D h[kHeadSlots];
        for (int s = 0; s < kHeadSlots; ++s) h[s] = D(0.0);
        mode_factors(inclination, 2, 2, h + kHeadFactors);
        for (int m = 0; m < kHmModes; ++m)
            mode_factors(inclination, kHmEll[m], kHmEmm[m], h + kHeadFactors + 4 * (m + 1));
        for (int s = 0; s < kHeadSlots; ++s) {
            // a binary that fails LAL's checks: NaN factors and derivatives
            const double v = (fl && s < 20) ? __builtin_nan("") : h[s].v;
            const double dv = (fl && s < 20) ? __builtin_nan("") : h[s].d[0];
            if (split) {
                ((float*)head)[b * kHeadSlots + s] = (float)v;
                ((float*)dhead)[b * kHeadSlots + s] = (float)dv;
            } else {
                ((double*)head)[b * kHeadSlots + s] = v;
                ((double*)dhead)[b * kHeadSlots + s] = dv;
            }
        }
        return;
    }
    const int k = g >> 2, dir = g & 3;
    using D = Dual<1>;
    // the (2,2)'s coefficients in this direction, from its derivative
    // set-up (the same arithmetic), read where the set-up uses them
    const XasCoeffs22<D>& c22 = kept[4 * b + dir];
    const WaveformVars<D>& wf = c22.wf;
    const int64_t at = (b * kHmModes + k) * kHmSlots;
    with_mode(k, [&](auto mode) __attribute__((always_inline)) {
        HmCoeffs<D> c{};
        hm_setup(wf, c22.amp, c22.ph, decltype(mode)::value, c);
        if (split) {
            float values[kHmSlots], tangents[kHmSlots];
            write_mode_record_split(c, wf, values, tangents);
#pragma unroll
            for (int s = 0; s < kHmSlots; ++s) {
                if (dir == 0) ((float*)modes)[at + s] = values[s];
                ((float*)dmodes)[4 * (at + s) + dir] = tangents[s];
            }
        } else {
            double values[kHmSlots], tangents[kHmSlots];
            write_mode_record_split(c, wf, values, tangents);
#pragma unroll
            for (int s = 0; s < kHmSlots; ++s) {
                if (dir == 0) ((double*)modes)[at + s] = values[s];
                ((double*)dmodes)[4 * (at + s) + dir] = tangents[s];
            }
        }
    });
}

// one thread per (binary, frequency), 256-thread blocks, a block on 256
// consecutive frequencies of one binary (its records wave-uniform). Three
// waves per SIMD: the FP64 arithmetic is latency-bound at the two that 184
// registers allow (168 registers and 20 B of spills: 2.44 ms against 3.48
// ms at 4096 x 8193; four waves spill 140 B and take 2.83 ms)
constexpr int kXhmBlock = 256;
template <class F>
__global__ __launch_bounds__(kXhmBlock) __attribute__((amdgpu_waves_per_eu(3))) void xhm_eval_kernel(const double* __restrict__ rec22,
                                                             const double* __restrict__ head,
                                                             const double* __restrict__ modes, int64_t nb,
                                                             const F* __restrict__ freqs, int64_t n, int64_t chunks,
                                                             double2* __restrict__ hp, double2* __restrict__ hc,
                                                             double2* __restrict__ mode_out) {
    const int64_t b = blockIdx.x / chunks;
    const int64_t i = (blockIdx.x % chunks) * kXhmBlock + threadIdx.x;
    if (b >= nb || i >= n) return;
    double h[5][2];
    xhm_modes_f64(rec22 + b * kSlots, modes + b * kHmModes * kHmSlots, (double)freqs[i], h);
    double p_re, p_im, c_re, c_im;
    xhm_sum(head + b * kHeadSlots, h, p_re, p_im, c_re, c_im);
    hp[b * n + i] = make_double2(p_re, p_im);
    hc[b * n + i] = make_double2(c_re, c_im);
    if (mode_out) {
        const double nan = __builtin_nan("");
        const bool bad = head[b * kHeadSlots] != head[b * kHeadSlots];
        for (int k = 0; k < 5; ++k)
            mode_out[(k * nb + b) * n + i] = bad ? make_double2(nan, nan) : make_double2(h[k][0], h[k][1]);
    }
}

// Binaries on their own grids: the concatenated grid (binary b owns
// [offsets[b], offsets[b + 1]) of freqs) or the uniform grid f0[b] + k
// df[b] of that length, in tiles of kXhmBlock consecutive bins of one
// binary (tiles[b] from tile_prefix_kernel at kXhmBlock); a persistent
// grid walks the tiles, each wave finding its tile's binary by the 64-ary
// search; per bin the shared-grid kernels' arithmetic (the same bits for
// the same frequency). mode_out, if not null, holds the five modes (5,
// total).
template <class F>
__device__ inline double xhm_ragged_freq(const F* __restrict__ freqs, const double* __restrict__ f0,
                                         const double* __restrict__ df, int64_t b, int64_t i, int64_t k) {
    return freqs ? (double)freqs[i] : f0[b] + (double)k * df[b];
}

template <class F>
__global__ __launch_bounds__(kXhmBlock) __attribute__((amdgpu_waves_per_eu(3))) void xhm_eval_ragged_kernel(
    const double* __restrict__ rec22, const double* __restrict__ head, const double* __restrict__ modes, int64_t nb,
    const F* __restrict__ freqs, const double* __restrict__ f0, const double* __restrict__ df,
    const int64_t* __restrict__ offsets, const int64_t* __restrict__ tiles, double2* __restrict__ hp,
    double2* __restrict__ hc, double2* __restrict__ mode_out) {
    const int64_t total_tiles = tiles[nb];
    const int64_t total = offsets[nb];
    for (int64_t w = blockIdx.x; w < total_tiles; w += gridDim.x) {
        const int64_t b = owner(tiles, nb, w);
        const int64_t k = (w - tiles[b]) * kXhmBlock + threadIdx.x;
        const int64_t i = offsets[b] + k;
        if (i >= offsets[b + 1]) continue;
        double h[5][2];
        xhm_modes_f64(rec22 + b * kSlots, modes + b * kHmModes * kHmSlots, xhm_ragged_freq(freqs, f0, df, b, i, k),
                      h);
        double p_re, p_im, c_re, c_im;
        xhm_sum(head + b * kHeadSlots, h, p_re, p_im, c_re, c_im);
        hp[i] = make_double2(p_re, p_im);
        hc[i] = make_double2(c_re, c_im);
        if (mode_out) {
            const double nan = __builtin_nan("");
            const bool bad = head[b * kHeadSlots] != head[b * kHeadSlots];
            for (int m = 0; m < 5; ++m)
                mode_out[m * total + i] = bad ? make_double2(nan, nan) : make_double2(h[m][0], h[m][1]);
        }
    }
}

template <class F>
__global__ __launch_bounds__(kXhmBlock) void xhm_eval_split_ragged_kernel(
    const float* __restrict__ rec22, const float* __restrict__ head, const float* __restrict__ modes, int64_t nb,
    const F* __restrict__ freqs, const double* __restrict__ f0, const double* __restrict__ df,
    const int64_t* __restrict__ offsets, const int64_t* __restrict__ tiles, float2* __restrict__ hp,
    float2* __restrict__ hc, float2* __restrict__ mode_out) {
    const int64_t total_tiles = tiles[nb];
    const int64_t total = offsets[nb];
    for (int64_t w = blockIdx.x; w < total_tiles; w += gridDim.x) {
        const int64_t b = owner(tiles, nb, w);
        const int64_t k = (w - tiles[b]) * kXhmBlock + threadIdx.x;
        const int64_t i = offsets[b] + k;
        if (i >= offsets[b + 1]) continue;
        float h[5][2];
        xhm_modes_split(rec22 + b * kSlots, modes + b * kHmModes * kHmSlots,
                        (float)xhm_ragged_freq(freqs, f0, df, b, i, k), h);
        float2 p, c;
        xhm_sum_f32(head + b * kHeadSlots, h, p, c);
        hp[i] = p;
        hc[i] = c;
        if (mode_out) {
            const float nan = __builtin_nanf("");
            const bool bad = head[b * kHeadSlots] != head[b * kHeadSlots];
            for (int m = 0; m < 5; ++m)
                mode_out[m * total + i] = bad ? make_float2(nan, nan) : make_float2(h[m][0], h[m][1]);
        }
    }
}

// the FP64 polarizations contracted with cotangents, nothing per bin
// written (the gradient's differences and closed forms): the evaluation's
// shape; for each of n_heads heads (head + j * nb * kHeadSlots: the modes'
// factors in the polarizations), partial[(j * nb + b) * chunks + c] = the
// sum over block c's frequencies of Re(conj(cot+) h+ + conj(cotx) hx), with
// the cotangent row b % n_cot; the modes are computed once for all heads.
// The block's sum is in a fixed order (64-lane butterflies, then the waves
// in turn), so the result is reproducible.
constexpr int kMaxProjectHeads = 4;
__global__ __launch_bounds__(kXhmBlock) __attribute__((amdgpu_waves_per_eu(3))) void xhm_project_kernel(
    const double* __restrict__ rec22, const double* __restrict__ head, int n_heads, const double* __restrict__ modes,
    int64_t nb, const double* __restrict__ freqs, int64_t n, int64_t chunks, const double2* __restrict__ cot_plus,
    const double2* __restrict__ cot_cross, int64_t n_cot, double* __restrict__ partial) {
    __shared__ double wave_sums[kMaxProjectHeads][kXhmBlock / 64];
    const int64_t b = blockIdx.x / chunks;
    const int64_t c = blockIdx.x % chunks;
    const int64_t i = c * kXhmBlock + threadIdx.x;
    double v[kMaxProjectHeads] = {0.0, 0.0, 0.0, 0.0};
    if (b < nb && i < n) {
        double h[5][2];
        xhm_modes_f64(rec22 + b * kSlots, modes + b * kHmModes * kHmSlots, freqs[i], h);
        const int64_t row = (b % n_cot) * n + i;
        const double2 cp = cot_plus[row], cc = cot_cross[row];
        for (int j = 0; j < kMaxProjectHeads; ++j) {
            if (j >= n_heads) break;
            double p_re, p_im, c_re, c_im;
            xhm_sum(head + (j * nb + b) * kHeadSlots, h, p_re, p_im, c_re, c_im);
            v[j] = (cp.x * p_re + cp.y * p_im) + (cc.x * c_re + cc.y * c_im);
        }
    }
    for (int j = 0; j < kMaxProjectHeads; ++j) {
        for (int offset = 32; offset > 0; offset >>= 1) v[j] += __shfl_xor(v[j], offset, 64);
        if ((threadIdx.x & 63) == 0) wave_sums[j][threadIdx.x >> 6] = v[j];
    }
    __syncthreads();
    if (threadIdx.x < n_heads && b < nb) {
        const int j = threadIdx.x;
        double total = 0.0;
        for (int w = 0; w < kXhmBlock / 64; ++w) total += wave_sums[j][w];
        partial[(j * nb + b) * chunks + c] = total;
    }
}

// the split evaluation: the same shape, FP32 with the inspiral's leading
// phase terms in FP64
__global__ __launch_bounds__(kXhmBlock) void xhm_eval_split_kernel(const float* __restrict__ rec22,
                                                                   const float* __restrict__ head,
                                                                   const float* __restrict__ modes, int64_t nb,
                                                                   const float* __restrict__ freqs, int64_t n,
                                                                   int64_t chunks, float2* __restrict__ hp,
                                                                   float2* __restrict__ hc,
                                                                   float2* __restrict__ mode_out) {
    const int64_t b = blockIdx.x / chunks;
    const int64_t i = (blockIdx.x % chunks) * kXhmBlock + threadIdx.x;
    if (b >= nb || i >= n) return;
    float h[5][2];
    xhm_modes_split(rec22 + b * kSlots, modes + b * kHmModes * kHmSlots, freqs[i], h);
    float2 p, c;
    xhm_sum_f32(head + b * kHeadSlots, h, p, c);
    hp[b * n + i] = p;
    hc[b * n + i] = c;
    if (mode_out) {
        const float nan = __builtin_nanf("");
        const bool bad = head[b * kHeadSlots] != head[b * kHeadSlots];
        for (int k = 0; k < 5; ++k)
            mode_out[(k * nb + b) * n + i] = bad ? make_float2(nan, nan) : make_float2(h[k][0], h[k][1]);
    }
}

// The gradient of IMRPhenomXHM's polarizations, reverse mode, in the split
// layout's formula: FP32 (the split gradient) or FP64. Each binary's grid
// is dealt to `chunks` waves per mode in blocks of 64 frequencies, block j
// to wave j mod chunks (so that the waves of one binary and mode share its
// regions evenly, whatever part of the grid each region holds), lane l
// taking frequency l of each block: the mode's forward values recomputed, each
// bin's dL/d(record slot) summed per lane in registers (the (2,2): the
// IMRPhenomXAS adjoint's 64 slots; a higher mode: its 86 slots with a
// derivative), one reduce-scatter over the wave, and the slots' totals
// contracted once with the record's derivatives. The (3,2)'s ringdown
// recycles the (2,2): its wave writes the cotangent it passes to the (2,2)
// per bin (gh22), which the (2,2)'s wave, launched after, adds to its own.
// A higher mode's bins in the ringdown of both its amplitude and its phase
// (most of them on a linear grid) go to waves of their own with compact
// accumulators (xhm_hm_rd_adjoint_*: 16 slots for a plain mode, 26 for the
// mixing (3,2)), so that more waves fit a SIMD; the general waves skip them.
// So do a mode's bins in the intermediate region of both
// (xhm_hm_inter_adjoint_*: 24 slots) and in the inspiral of both
// (xhm_hm_ins_adjoint_*: 43 slots); the general waves keep the bins whose
// amplitude and phase regions differ.
// Per wave, partial[((b * 17 + row) * chunks + c) * 8 + j] (row 0 the (2,2),
// 1..4 the higher modes' general waves, 5..8 their ringdown waves, 9..12
// their intermediate waves, 13..16 their inspiral waves):
// Re conj(cot) dh/d(mass_1, mass_2, chi_1z, chi_2z)
// (0..3); Re conj(G) h (4); Re conj(cot) dF/d(iota) h (5); Re m conj(G) i h
// (6), G = conj(F+) cot+ + conj(Fx) cotx the mode's cotangent; 0 (7).
// With the grid's gradient, each higher mode's wave writes its bins' dL/df
// (gfm[k][b][i]), the (2,2)'s wave adds them to its own (into gfm[0][b][i]),
// and two passes sum them over the binaries in a fixed order.
constexpr int kXhmGradSlots = 8;
constexpr int kXhmGradRows = 17;
// the ints per binary and mode of xhm_cuts_kernel
constexpr int kCutsPer = 4;
// at least this many higher-mode waves per launch (the chip holds ~3,000
// at the kernels' occupancy), and at most this many blocks of 64
// frequencies per wave: a wave's bins may all fall in one region, and the
// longest waves, started last, end the launch
constexpr int64_t kXhmGradWaves = 12288;
constexpr int64_t kXhmGradBlocksPerWave = 64;
// the grid gradient's first pass: each thread sums this many binaries
constexpr int64_t kXhmGridGroup = 64;

// the waves per binary and higher mode on its general bins (its
// inspiral, the short part of its grid before the intermediate bins) and
// on its intermediate bins, at most the chunks (1, 2 or 3 of each within
// 3% of one another at 4096 x 8193, two the fastest)
constexpr int64_t kXhmGenChunks = 2;
constexpr int64_t kXhmInterChunks = 2;

struct XhmGradPlan {
    int64_t chunks, gen_chunks, inter_chunks, loss_chunks, grid_groups;
    int64_t table, cuts, gh22, partial, cot_plus, cot_cross, loss_partial, gfm, grid_partial, bytes;
};

inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// the chunks and the workspace's layout (each part 256-aligned) of B
// binaries on N frequencies in T (4- or 8-byte reals); bytes < 0 where the
// sizes overflow
inline XhmGradPlan xhm_grad_plan(int64_t nb, int64_t n, bool loss, bool grid, int64_t real) {
    XhmGradPlan g{};
    const int64_t blocks = (n + 63) / 64;
    int64_t chunks = (kXhmGradWaves + 4 * nb - 1) / (4 * nb);
    const int64_t least = (blocks + kXhmGradBlocksPerWave - 1) / kXhmGradBlocksPerWave;
    chunks = chunks < least ? least : chunks;
    g.chunks = chunks > blocks ? blocks : chunks;
    g.gen_chunks = g.chunks < kXhmGenChunks ? g.chunks : kXhmGenChunks;
    g.inter_chunks = g.chunks < kXhmInterChunks ? g.chunks : kXhmInterChunks;
    g.loss_chunks = (n + kXhmBlock - 1) / kXhmBlock;
    g.grid_groups = (nb + kXhmGridGroup - 1) / kXhmGridGroup;
    if (nb > (INT64_MAX / 128) / n) {
        g.bytes = -1;
        return g;
    }
    const int64_t row = (real == 8 ? (int64_t)sizeof(PowerRow<double>) : (int64_t)sizeof(PowerRow<float>));
    int64_t off = 0;
    g.table = off;
    off += align256(n * row);
    g.cuts = off;
    off += align256((kCutsPer * nb * kHmModes + 1) * (int64_t)sizeof(int));
    g.gh22 = off;
    off += align256(nb * n * 2 * real);
    g.partial = off;
    off += align256(nb * kXhmGradRows * g.chunks * kXhmGradSlots * 8);
    if (loss) {
        g.cot_plus = off;
        off += align256(nb * n * 2 * real);
        g.cot_cross = off;
        off += align256(nb * n * 2 * real);
        g.loss_partial = off;
        off += align256(nb * g.loss_chunks * 8);
    }
    if (grid) {
        g.gfm = off;
        off += align256(kHmModes * nb * n * real);
        g.grid_partial = off;
        off += align256(g.grid_groups * n * 8);
    }
    g.bytes = off;
    return g;
}

template <class T>
struct Vec2Of;
template <>
struct Vec2Of<float> {
    using type = float2;
};
template <>
struct Vec2Of<double> {
    using type = double2;
};
template <class T>
using vec2_t = typename Vec2Of<T>::type;

// a double summed over the wave (every lane active), in every lane
__device__ inline double wave_sum(double x) {
    for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, 64);
    return x;
}

// the sums over the wave of d[0..3] and the closed forms, written by lane 0
template <class T>
__device__ inline void xhm_grad_store(double* d, T v4, T v5, T v6, int lane, double* out) {
    double v[kXhmGradSlots] = {d[0], d[1], d[2], d[3], (double)v4, (double)v5, (double)v6, 0.0};
#pragma unroll
    for (int j = 0; j < 7; ++j) v[j] = wave_sum(v[j]);
    if (lane == 0)
        for (int j = 0; j < kXhmGradSlots; ++j) out[j] = v[j];
}

// mode h's cotangent from the polarizations' (F+, Fx the mode's factors)
template <class T, class T2>
__device__ inline void mode_cotangent(const T* fac, T2 cp, T2 cc, T& g_re, T& g_im) {
    g_re = fac[0] * cp.x + fac[1] * cp.y + fac[2] * cc.x + fac[3] * cc.y;
    g_im = fac[0] * cp.y - fac[1] * cp.x + fac[2] * cc.y - fac[3] * cc.x;
}

// the inclination's closed form: Re(conj(cot+) dF+ h + conj(cotx) dFx h)
template <class T, class T2>
__device__ inline T inclination_term(const T* dfac, T2 cp, T2 cc, T h_re, T h_im) {
    const T dp_re = dfac[0] * h_re - dfac[1] * h_im, dp_im = dfac[0] * h_im + dfac[1] * h_re;
    const T dc_re = dfac[2] * h_re - dfac[3] * h_im, dc_im = dfac[2] * h_im + dfac[3] * h_re;
    return cp.x * dp_re + cp.y * dp_im + cc.x * dc_re + cc.y * dc_im;
}

// a frequency's powers, as the split layout's forward takes them
__device__ inline PowF32 split_powers(const PowerRow<float>& row) {
    PowF32 q;
    q.f = row.f;
    q.t = row.t;
    q.m1 = row.m1;
    q.lf = row.lf;
    q.m76 = row.m76;
    q.m53 = row.m53;
    q.m53_hi = q.m53_lo = 0.0f;
    return q;
}
__device__ inline FreqPowers<double> split_powers(const PowerRow<double>& row) {
    return FreqPowers<double>{row.f, row.t, row.m1, row.lf, row.m76, row.m53};
}

// the cotangents 2 w (h - d) of the loss and its value per block (the
// evaluation's shape; loss_partial[b * chunks + c], summed in a fixed order)
template <class T>
__global__ __launch_bounds__(kXhmBlock) void xhm_residual_kernel(
    const T* __restrict__ rec22, const T* __restrict__ head, const T* __restrict__ modes, int64_t nb,
    const PowerRow<T>* __restrict__ table, int64_t n, int64_t chunks, const vec2_t<T>* __restrict__ data_plus,
    const vec2_t<T>* __restrict__ data_cross, const T* __restrict__ weights, vec2_t<T>* __restrict__ cot_plus,
    vec2_t<T>* __restrict__ cot_cross, double* __restrict__ loss_partial) {
    __shared__ double wave_sums[kXhmBlock / 64];
    const int64_t b = blockIdx.x / chunks;
    const int64_t c = blockIdx.x % chunks;
    const int64_t i = c * kXhmBlock + threadIdx.x;
    double l = 0.0;
    if (b < nb && i < n) {
        T h[5][2];
        xhm_modes_split_at(rec22 + b * kSlots, modes + b * kHmModes * kHmSlots, split_powers(table[i]), h);
        vec2_t<T> hp, hc;
        xhm_sum_split(head + b * kHeadSlots, h, hp, hc);
        const vec2_t<T> dp = data_plus[b * n + i], dc = data_cross[b * n + i];
        const T w = weights[i];
        const T rp_re = hp.x - dp.x, rp_im = hp.y - dp.y, rx_re = hc.x - dc.x, rx_im = hc.y - dc.y;
        l = (double)(w * (rp_re * rp_re + rp_im * rp_im + rx_re * rx_re + rx_im * rx_im));
        const T two_w = T(2) * w;
        cot_plus[b * n + i] = vec2_t<T>{two_w * rp_re, two_w * rp_im};
        cot_cross[b * n + i] = vec2_t<T>{two_w * rx_re, two_w * rx_im};
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) wave_sums[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0 && b < nb) {
        double total = 0.0;
        for (int w = 0; w < kXhmBlock / 64; ++w) total += wave_sums[w];
        loss_partial[b * chunks + c] = total;
    }
}

// Each higher mode's first ringdown bin (hm_ringdown_bin) where the grid
// ascends: the mode's M f then grows with the bin, and so does the
// predicate (each of its comparisons turns true once, a NaN bound leaves it
// constant), so that its general bins are those before the cut and its
// ringdown bins those from it. The general waves stop at the cut and the
// ringdown waves start at its block, instead of both visiting every bin;
// -1 (every wave visits every bin, testing each) where the grid does not
// ascend or M is not a positive finite number. Likewise a mode's
// intermediate bins, a range of their own waves that the general waves
// jump over.

// *ascends = 1 if the table's frequencies never decrease (and hold no NaN)
// and the bins' indices fit an int, else 0: one block of kOrderBlock threads
constexpr int kOrderBlock = 1024;
template <class T>
__global__ __launch_bounds__(kOrderBlock) void grid_order_kernel(const PowerRow<T>* __restrict__ table, int64_t n,
                                                                 int* __restrict__ ascends) {
    int descends = n > INT_MAX;
#pragma unroll 8
    for (int64_t i = threadIdx.x; i + 1 < n; i += kOrderBlock) descends |= !(table[i].f <= table[i + 1].f);
    descends = __syncthreads_or(descends);
    if (threadIdx.x == 0) *ascends = !descends;
}

// The first bin of [0, n) where pred (false, then true, along the bins)
// holds, n if none, by one wave: a 64-ary search (lane l probes lo + (l +
// 1) step - 1; three rounds for 8193 bins)
template <class Pred>
__device__ inline int64_t wave_first_true(int64_t n, int lane, Pred pred) {
    // pred is false below lo and true from hi
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t step = (hi - lo + 63) / 64;
        const int64_t p = lo + (lane + 1) * step - 1;
        const bool probe = p < hi;
        const uint64_t found = __ballot(probe && pred(p));
        const uint64_t probed = __ballot(probe);
        if (found) {
            const int64_t first = __ffsll((unsigned long long)found) - 1;
            hi = lo + (first + 1) * step - 1;
            lo += first * step;
        } else {
            lo += (64 - __clzll((unsigned long long)probed)) * step;
        }
    }
    return lo;
}

// One wave per (binary, mode), kCutsPer ints at cuts[kCutsPer x]: the first
// ringdown bin (-1: none known, every wave tests every bin); [lo, hi) its
// intermediate bins (hm_intermediate_bin: from the first bin past both
// inspiral bounds to the first bin at either ringdown bound); and the end of
// its inspiral bins (hm_inspiral_bin holds on the bins before it and on no
// bin from it). Empty ranges where the cut is unknown.
template <class T>
__global__ __launch_bounds__(64) void xhm_cuts_kernel(const T* __restrict__ modes, int64_t nb,
                                                      const PowerRow<T>* __restrict__ table, int64_t n,
                                                      const int* __restrict__ ascends, int* __restrict__ cuts) {
    const int64_t x = blockIdx.x;
    const int lane = threadIdx.x;
    if (x >= nb * kHmModes) return;
    const T* r = modes + x * kHmSlots;
    const T msec = r[kHmMsec];
    int rd = -1, lo = 0, hi = 0, ins = 0;
    if (*ascends && msec > T(0) && msec <= std::numeric_limits<T>::max()) {
        // the inspiral predicate is a conjunction of tests "M f below a
        // bound": true on the bins before some bin, false from it
        ins = (int)wave_first_true(n, lane, [&](int64_t i) { return !hm_inspiral_bin(r, msec * table[i].f); });
        rd = (int)wave_first_true(n, lane, [&](int64_t i) { return hm_ringdown_bin(r, msec * table[i].f); });
        lo = (int)wave_first_true(n, lane, [&](int64_t i) {
            const T mf = msec * table[i].f;
            return !(mf < r[kHmFAmpIn]) && !(mf < r[kHmFPhaseIn]);
        });
        hi = (int)wave_first_true(n, lane, [&](int64_t i) {
            const T mf = msec * table[i].f;
            return mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm];
        });
        hi = hi < lo ? lo : hi;
    }
    if (lane == 0) {
        cuts[kCutsPer * x] = rd;
        cuts[kCutsPer * x + 1] = lo;
        cuts[kCutsPer * x + 2] = hi;
        cuts[kCutsPer * x + 3] = ins;
    }
}

// the first bin of wave (c, lane) at or past the block of bin `from`: the
// wave takes blocks c, c + chunks, ... of 64 bins
__device__ inline int64_t first_bin_from(int64_t from, int64_t c, int64_t chunks, int lane) {
    const int64_t block = from / 64;
    const int64_t skip = block > c ? (block - c + chunks - 1) / chunks : 0;
    return (c + skip * chunks) * 64 + lane;
}

// one wave per (binary, chunk c < gen, higher mode k): the mode's adjoint
// (and, with gfm, its bins' dL/df); gen (<= chunks) waves take a mode's
// general bins, the short part of its grid before the ringdown, and the
// wave of chunk 0 zeroes the mode's partial sums of chunks gen .. chunks - 1
template <class T, bool Grid>
__device__ __forceinline__ void xhm_hm_adjoint_wave(
    const T* __restrict__ rec22, const T* __restrict__ head, const T* __restrict__ dhead, const T* __restrict__ modes,
    const T* __restrict__ dmodes, int64_t nb, const PowerRow<T>* __restrict__ table, int64_t n, int64_t chunks,
    int64_t gen, const vec2_t<T>* __restrict__ cot_plus, const vec2_t<T>* __restrict__ cot_cross,
    vec2_t<T>* __restrict__ gh22, double* __restrict__ partial, T* __restrict__ gfm, const int* __restrict__ cuts) {
    const int64_t x = blockIdx.x;
    const int k = (int)(x % kHmModes);
    const int64_t c = (x / kHmModes) % gen;
    const int64_t b = x / kHmModes / gen;
    if (b >= nb) return;
    const int lane = threadIdx.x;
    if (c == 0 && lane < kXhmGradSlots)
        for (int64_t rest = gen; rest < chunks; ++rest)
            partial[((b * kXhmGradRows + 1 + k) * chunks + rest) * kXhmGradSlots + lane] = 0.0;
    // the records in LDS: the loop's many record reads then cost no scalar
    // loads and waits (its registers cannot hold them all)
    __shared__ T s_mode[kHmSlots], s_22[kSlots];
    for (int j = lane; j < kHmSlots; j += 64) s_mode[j] = modes[(b * kHmModes + k) * kHmSlots + j];
    s_22[lane] = rec22[b * kSlots + lane];
    __syncthreads();
    const T* r = s_mode;
    const T* r22 = s_22;
    const T* fac = head + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T* dfac = dhead + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T emm = (T)kHmEmm[k];
    const bool on = r[kHmOn] != T(0);
    const bool mixing = r[kHmMixing] != T(0);
    const bool writes = k == kHmMixMode;
    const T amp0 = r[kHmAmp0];
    const T inv_amp0 = d_rcp(fabs(amp0));
    const T sign_inv = amp0 < T(0) ? -inv_amp0 : inv_amp0;
    T acc[kHmAcc];
#pragma unroll
    for (int j = 0; j < kHmAcc; ++j) acc[j] = T(0);
    T v4 = T(0), v5 = T(0), v6 = T(0);
    // the blocks of 64 bins c, c + gen, ... up to the mode's first ringdown
    // bin (all bins where the mode vanishes: this wave then writes each
    // bin's zero cotangents); the blocks wholly inside the mode's
    // intermediate range [lo, hi) (its own waves') jumped over, the bins at
    // its edges tested
    const int* cut = cuts + kCutsPer * (b * kHmModes + k);
    const int64_t end = (!on || cut[0] < 0) ? n : cut[0];
    const int64_t inter_lo = cut[1], inter_hi = cut[2];
    const int64_t skip_from = (inter_lo + 63) / 64, skip_to = inter_hi / 64;
    // from the block of the inspiral's end (the inspiral waves' bins before it)
    const int64_t ins_end = on ? cut[3] : 0;
    bool took = false;
    for (int64_t blk = first_bin_from(ins_end, c, gen, 0) / 64; blk * 64 < end; blk += gen) {
        if (on && blk >= skip_from && blk < skip_to) {
            blk += (skip_to - blk + gen - 1) / gen * gen;
            if (blk * 64 >= end) break;
        }
        const int64_t i = blk * 64 + lane;
        if (i >= end) continue;
        T gw_re = T(0), gw_im = T(0), gf = T(0);
        if (on) {
            const PowerRow<T> row = table[i];
            const BinPowers<T> p{row.f, row.t, row.m1, row.lf, row.m76, row.m53};
            const T mf = r[kHmMsec] * p.f;
            if (hm_ringdown_bin(r, mf)) continue;  // the ringdown waves' bin
            if (i < ins_end && hm_inspiral_bin(r, mf)) continue;  // the inspiral waves' bin
            if (i >= inter_lo && i < inter_hi && hm_intermediate_bin(r, mf)) continue;  // the intermediate waves' bin
            took = true;
            const vec2_t<T> cp = cot_plus[b * n + i], cc = cot_cross[b * n + i];
            T g_re, g_im;
            mode_cotangent(fac, cp, cc, g_re, g_im);
            T w_re = T(0), w_im = T(0);
            if (mixing && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm])) {
                // the (2,2) over |amp0|, as the forward computes it
                T a22, phi22, s22, c22;
                mode22_split(r22, split_powers(row), a22, phi22);
                m_sincos(phi22, &s22, &c22);
                w_re = a22 * c22 * inv_amp0;
                w_im = a22 * s22 * inv_amp0;
            }
            T h_re, h_im;
            hm_adjoint_split<Grid>(r, p, w_re, w_im, g_re, g_im, acc, h_re, h_im, gw_re, gw_im, gf);
            v4 += g_re * h_re + g_im * h_im;
            v6 += emm * (g_im * h_re - g_re * h_im);
            v5 += inclination_term(dfac, cp, cc, h_re, h_im);
            // w's dependence on amp0 (w = h_22 / |amp0|), for the mixing mode
            if (mixing) acc[aAmp0] -= (gw_re * w_re + gw_im * w_im) * sign_inv;
        }
        if (writes) gh22[b * n + i] = vec2_t<T>{gw_re * inv_amp0, gw_im * inv_amp0};
        if (Grid) gfm[(k * nb + b) * n + i] = gf;
    }
    if (!on || __ballot(took) == 0) {
        // a mode that vanishes (the odd modes of equal masses): nothing,
        // whatever its record's derivatives (which may be NaN there); a wave
        // that took no bin (on an ascending grid the other kinds of wave
        // usually take them all): zero sums, without the reductions
        if (lane < kXhmGradSlots) partial[((b * kXhmGradRows + 1 + k) * chunks + c) * kXhmGradSlots + lane] = 0.0;
        return;
    }
    // lane j: the totals of accumulators j and 64 + j
    T hi[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) hi[j] = 64 + j < kHmAcc ? acc[64 + j] : T(0);
    const double lo_total = (double)reduce_scatter64(acc, lane);
    const double hi_total = (double)reduce_scatter64(hi, lane);
    const T* t = dmodes + (b * kHmModes + k) * kHmSlots * 4;
    double d[4];
    for (int j = 0; j < 4; ++j) {
        const T tl = t[4 * kHmAccSlot[lane] + j];
        d[j] = tl == T(0) ? 0.0 : lo_total * (double)tl;
    }
    if (64 + lane < kHmAcc) {
        for (int j = 0; j < 4; ++j) {
            const T th = t[4 * kHmAccSlot[64 + lane] + j];
            d[j] += th == T(0) ? 0.0 : hi_total * (double)th;
        }
    }
    xhm_grad_store(d, v4, v5, v6, lane, partial + ((b * kXhmGradRows + 1 + k) * chunks + c) * kXhmGradSlots);
}

// FP32: three waves per SIMD (168 registers; the 86 accumulators in
// registers); FP64: the accumulators and the forward in doubles, one wave
template <bool Grid>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void xhm_hm_adjoint_kernel(
    const float* __restrict__ rec22, const float* __restrict__ head, const float* __restrict__ dhead,
    const float* __restrict__ modes, const float* __restrict__ dmodes, int64_t nb,
    const PowerRow<float>* __restrict__ table, int64_t n, int64_t chunks, int64_t gen,
    const float2* __restrict__ cot_plus, const float2* __restrict__ cot_cross, float2* __restrict__ gh22,
    double* __restrict__ partial, float* __restrict__ gfm, const int* __restrict__ cuts) {
    xhm_hm_adjoint_wave<float, Grid>(rec22, head, dhead, modes, dmodes, nb, table, n, chunks, gen, cot_plus,
                                     cot_cross, gh22, partial, gfm, cuts);
}
template <bool Grid>
__global__ __launch_bounds__(64) void xhm_hm_adjoint_f64_kernel(
    const double* __restrict__ rec22, const double* __restrict__ head, const double* __restrict__ dhead,
    const double* __restrict__ modes, const double* __restrict__ dmodes, int64_t nb,
    const PowerRow<double>* __restrict__ table, int64_t n, int64_t chunks, int64_t gen,
    const double2* __restrict__ cot_plus, const double2* __restrict__ cot_cross, double2* __restrict__ gh22,
    double* __restrict__ partial, double* __restrict__ gfm, const int* __restrict__ cuts) {
    xhm_hm_adjoint_wave<double, Grid>(rec22, head, dhead, modes, dmodes, nb, table, n, chunks, gen, cot_plus,
                                      cot_cross, gh22, partial, gfm, cuts);
}

// one wave per (binary, chunk, higher mode) on the mode's ringdown bins
// (hm_ringdown_bin): the plain modes (2,1), (3,3), (4,4) with 16 compact
// accumulators (Mixing false; blocks (b, c, 3 modes)), or the mixing (3,2)
// with 26 (Mixing true; blocks (b, c)), its (2,2) recomputed and the
// cotangent of the (2,2) written per bin as the general wave does
template <class T, bool Grid, bool Mixing>
__device__ __forceinline__ void xhm_hm_rd_adjoint_wave(
    const T* __restrict__ rec22, const T* __restrict__ head, const T* __restrict__ dhead, const T* __restrict__ modes,
    const T* __restrict__ dmodes, int64_t nb, const PowerRow<T>* __restrict__ table, int64_t n, int64_t chunks,
    const vec2_t<T>* __restrict__ cot_plus, const vec2_t<T>* __restrict__ cot_cross,
    vec2_t<T>* __restrict__ gh22, double* __restrict__ partial, T* __restrict__ gfm, const int* __restrict__ cuts) {
    constexpr int kModes = Mixing ? 1 : 3;
    constexpr int kAcc = Mixing ? (int)kMixAcc : (int)kRdAcc;
    const int64_t x = blockIdx.x;
    const int km = (int)(x % kModes);
    const int k = Mixing ? kHmMixMode : (km < 2 ? km : 3);
    const int64_t c = (x / kModes) % chunks;
    const int64_t b = x / kModes / chunks;
    if (b >= nb) return;
    const int lane = threadIdx.x;
    __shared__ T s_mode[kHmSlots], s_22[kSlots];
    for (int j = lane; j < kHmSlots; j += 64) s_mode[j] = modes[(b * kHmModes + k) * kHmSlots + j];
    if (Mixing) s_22[lane] = rec22[b * kSlots + lane];
    __syncthreads();
    const T* r = s_mode;
    double* out = partial + ((b * kXhmGradRows + 5 + k) * chunks + c) * kXhmGradSlots;
    if (r[kHmOn] == T(0)) {
        // the mode vanishes: the general wave writes its every bin
        if (lane < kXhmGradSlots) out[lane] = 0.0;
        return;
    }
    const T* fac = head + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T* dfac = dhead + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T emm = (T)kHmEmm[k];
    const T amp0 = r[kHmAmp0];
    const T inv_amp0 = d_rcp(fabs(amp0));
    const T sign_inv = amp0 < T(0) ? -inv_amp0 : inv_amp0;
    T acc[kAcc];
#pragma unroll
    for (int j = 0; j < kAcc; ++j) acc[j] = T(0);
    T v4 = T(0), v5 = T(0), v6 = T(0);
    // from the block of the mode's first ringdown bin
    const int cut = cuts[kCutsPer * (b * kHmModes + k)];
    // a plain mode's iteration loads the next one's powers and cotangents
    // before its own arithmetic, so that their latency overlaps it (the
    // mixing mode's registers have no room for them: 167 -> 184, two waves
    // per SIMD instead of three, 315 -> 380 us at 4096 x 8193)
    const int64_t step = chunks * 64;
    int64_t i = first_bin_from(cut < 0 ? 0 : cut, c, chunks, lane);
    PowerRow<T> next_row{};
    vec2_t<T> next_cp{}, next_cc{};
    if (!Mixing && i < n) {
        next_row = table[i];
        next_cp = cot_plus[b * n + i];
        next_cc = cot_cross[b * n + i];
    }
    for (; i < n; i += step) {
        PowerRow<T> row;
        vec2_t<T> cp, cc;
        if constexpr (Mixing) {
            row = table[i];
        } else {
            row = next_row;
            cp = next_cp;
            cc = next_cc;
            if (i + step < n) {
                next_row = table[i + step];
                next_cp = cot_plus[b * n + i + step];
                next_cc = cot_cross[b * n + i + step];
            }
        }
        const BinPowers<T> p{row.f, row.t, row.m1, row.lf, row.m76, row.m53};
        if (!hm_ringdown_bin(r, r[kHmMsec] * p.f)) continue;  // the general wave's bin
        if constexpr (Mixing) {
            cp = cot_plus[b * n + i];
            cc = cot_cross[b * n + i];
        }
        T g_re, g_im;
        mode_cotangent(fac, cp, cc, g_re, g_im);
        T h_re, h_im, gf = T(0);
        if constexpr (Mixing) {
            T a22, phi22, s22, c22;
            mode22_split(s_22, split_powers(row), a22, phi22);
            m_sincos(phi22, &s22, &c22);
            const T w_re = a22 * c22 * inv_amp0, w_im = a22 * s22 * inv_amp0;
            T gw_re, gw_im;
            hm_adjoint_ringdown_mixing<Grid>(r, p, w_re, w_im, g_re, g_im, acc, h_re, h_im, gw_re, gw_im, gf);
            // w's dependence on amp0 (w = h_22 / |amp0|)
            acc[xAmp0] -= (gw_re * w_re + gw_im * w_im) * sign_inv;
            gh22[b * n + i] = vec2_t<T>{gw_re * inv_amp0, gw_im * inv_amp0};
        } else {
            hm_adjoint_ringdown<Grid>(r, p, g_re, g_im, acc, h_re, h_im, gf);
        }
        v4 += g_re * h_re + g_im * h_im;
        v6 += emm * (g_im * h_re - g_re * h_im);
        v5 += inclination_term(dfac, cp, cc, h_re, h_im);
        if (Grid) gfm[(k * nb + b) * n + i] = gf;
    }
    // lane j: the total of accumulator j, contracted with its slot's
    // derivatives
    T full[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) full[j] = j < kAcc ? acc[j] : T(0);
    const double total = (double)reduce_scatter64(full, lane);
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    if (lane < kAcc) {
        const int slot = Mixing ? kMixAccSlot[lane] : kRdAccSlot[lane];
        const T* t = dmodes + ((b * kHmModes + k) * kHmSlots + slot) * 4;
        for (int j = 0; j < 4; ++j) d[j] = t[j] == T(0) ? 0.0 : total * (double)t[j];
    }
    xhm_grad_store(d, v4, v5, v6, lane, out);
}

template <bool Grid, bool Mixing>
__global__ __launch_bounds__(64) void xhm_hm_rd_adjoint_kernel(
    const float* __restrict__ rec22, const float* __restrict__ head, const float* __restrict__ dhead,
    const float* __restrict__ modes, const float* __restrict__ dmodes, int64_t nb,
    const PowerRow<float>* __restrict__ table, int64_t n, int64_t chunks,
    const float2* __restrict__ cot_plus, const float2* __restrict__ cot_cross, float2* __restrict__ gh22,
    double* __restrict__ partial, float* __restrict__ gfm, const int* __restrict__ cuts) {
    xhm_hm_rd_adjoint_wave<float, Grid, Mixing>(rec22, head, dhead, modes, dmodes, nb, table, n, chunks,
                                                cot_plus, cot_cross, gh22, partial, gfm, cuts);
}
template <bool Grid, bool Mixing>
__global__ __launch_bounds__(64) void xhm_hm_rd_adjoint_f64_kernel(
    const double* __restrict__ rec22, const double* __restrict__ head, const double* __restrict__ dhead,
    const double* __restrict__ modes, const double* __restrict__ dmodes, int64_t nb,
    const PowerRow<double>* __restrict__ table, int64_t n, int64_t chunks,
    const double2* __restrict__ cot_plus, const double2* __restrict__ cot_cross, double2* __restrict__ gh22,
    double* __restrict__ partial, double* __restrict__ gfm, const int* __restrict__ cuts) {
    xhm_hm_rd_adjoint_wave<double, Grid, Mixing>(rec22, head, dhead, modes, dmodes, nb, table, n, chunks,
                                                 cot_plus, cot_cross, gh22, partial, gfm, cuts);
}

// one wave per (binary, chunk, higher mode) on the mode's intermediate bins
// (hm_intermediate_bin in the cuts' range [lo, hi), the bins that the
// general waves skip; blocks (b, c, 4 modes)) with 24 compact accumulators;
// the (3,2)'s wave writes those bins' cotangent of the (2,2), zero (no
// rotation there), as the general wave does for its bins
template <class T, bool Grid>
__device__ __forceinline__ void xhm_hm_inter_adjoint_wave(
    const T* __restrict__ head, const T* __restrict__ dhead, const T* __restrict__ modes,
    const T* __restrict__ dmodes, int64_t nb, const PowerRow<T>* __restrict__ table, int64_t n, int64_t chunks,
    int64_t waves, const vec2_t<T>* __restrict__ cot_plus, const vec2_t<T>* __restrict__ cot_cross,
    vec2_t<T>* __restrict__ gh22, double* __restrict__ partial, T* __restrict__ gfm, const int* __restrict__ cuts) {
    // waves (<= chunks) per binary and mode; chunk 0's zeroes the mode's
    // partial sums of chunks waves .. chunks - 1
    const int64_t x = blockIdx.x;
    const int k = (int)(x % kHmModes);
    const int64_t c = (x / kHmModes) % waves;
    const int64_t b = x / kHmModes / waves;
    if (b >= nb) return;
    const int lane = threadIdx.x;
    if (c == 0 && lane < kXhmGradSlots)
        for (int64_t rest = waves; rest < chunks; ++rest)
            partial[((b * kXhmGradRows + 9 + k) * chunks + rest) * kXhmGradSlots + lane] = 0.0;
    __shared__ T s_mode[kHmSlots];
    for (int j = lane; j < kHmSlots; j += 64) s_mode[j] = modes[(b * kHmModes + k) * kHmSlots + j];
    __syncthreads();
    const T* r = s_mode;
    const bool writes = k == kHmMixMode;
    double* out = partial + ((b * kXhmGradRows + 9 + k) * chunks + c) * kXhmGradSlots;
    if (r[kHmOn] == T(0)) {
        // the mode vanishes: the general wave writes its every bin
        if (lane < kXhmGradSlots) out[lane] = 0.0;
        return;
    }
    const T* fac = head + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T* dfac = dhead + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T emm = (T)kHmEmm[k];
    T acc[kInterAcc];
#pragma unroll
    for (int j = 0; j < kInterAcc; ++j) acc[j] = T(0);
    T v4 = T(0), v5 = T(0), v6 = T(0);
    const int* cut = cuts + kCutsPer * (b * kHmModes + k);
    const int64_t lo = cut[1], hi = cut[2];
    // each iteration loads the next one's powers and cotangents before its
    // own arithmetic, so that their latency overlaps it
    const int64_t step = waves * 64;
    int64_t i = first_bin_from(lo, c, waves, lane);
    PowerRow<T> next_row{};
    vec2_t<T> next_cp{}, next_cc{};
    if (i < hi) {
        next_row = table[i];
        next_cp = cot_plus[b * n + i];
        next_cc = cot_cross[b * n + i];
    }
    for (; i < hi; i += step) {
        const PowerRow<T> row = next_row;
        const vec2_t<T> cp = next_cp, cc = next_cc;
        if (i + step < hi) {
            next_row = table[i + step];
            next_cp = cot_plus[b * n + i + step];
            next_cc = cot_cross[b * n + i + step];
        }
        if (i < lo) continue;
        const BinPowers<T> p{row.f, row.t, row.m1, row.lf, row.m76, row.m53};
        if (!hm_intermediate_bin(r, r[kHmMsec] * p.f)) continue;  // the general wave's bin
        T g_re, g_im;
        mode_cotangent(fac, cp, cc, g_re, g_im);
        T h_re, h_im, gf = T(0);
        hm_adjoint_intermediate<Grid>(r, p, g_re, g_im, acc, h_re, h_im, gf);
        v4 += g_re * h_re + g_im * h_im;
        v6 += emm * (g_im * h_re - g_re * h_im);
        v5 += inclination_term(dfac, cp, cc, h_re, h_im);
        if (writes) gh22[b * n + i] = vec2_t<T>{T(0), T(0)};
        if (Grid) gfm[(k * nb + b) * n + i] = gf;
    }
    // lane j: the total of accumulator j, contracted with its slot's
    // derivatives
    T full[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) full[j] = j < kInterAcc ? acc[j] : T(0);
    const double total = (double)reduce_scatter64(full, lane);
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    if (lane < kInterAcc) {
        const T* t = dmodes + ((b * kHmModes + k) * kHmSlots + kInterAccSlot[lane]) * 4;
        for (int j = 0; j < 4; ++j) d[j] = t[j] == T(0) ? 0.0 : total * (double)t[j];
    }
    xhm_grad_store(d, v4, v5, v6, lane, out);
}

template <bool Grid>
__global__ __launch_bounds__(64) void xhm_hm_inter_adjoint_kernel(
    const float* __restrict__ head, const float* __restrict__ dhead, const float* __restrict__ modes,
    const float* __restrict__ dmodes, int64_t nb, const PowerRow<float>* __restrict__ table, int64_t n,
    int64_t chunks, int64_t waves, const float2* __restrict__ cot_plus, const float2* __restrict__ cot_cross,
    float2* __restrict__ gh22, double* __restrict__ partial, float* __restrict__ gfm, const int* __restrict__ cuts) {
    xhm_hm_inter_adjoint_wave<float, Grid>(head, dhead, modes, dmodes, nb, table, n, chunks, waves, cot_plus,
                                           cot_cross, gh22, partial, gfm, cuts);
}
template <bool Grid>
__global__ __launch_bounds__(64) void xhm_hm_inter_adjoint_f64_kernel(
    const double* __restrict__ head, const double* __restrict__ dhead, const double* __restrict__ modes,
    const double* __restrict__ dmodes, int64_t nb, const PowerRow<double>* __restrict__ table, int64_t n,
    int64_t chunks, int64_t waves, const double2* __restrict__ cot_plus, const double2* __restrict__ cot_cross,
    double2* __restrict__ gh22, double* __restrict__ partial, double* __restrict__ gfm,
    const int* __restrict__ cuts) {
    xhm_hm_inter_adjoint_wave<double, Grid>(head, dhead, modes, dmodes, nb, table, n, chunks, waves, cot_plus,
                                            cot_cross,
                                            gh22, partial, gfm, cuts);
}

// one wave per (binary, chunk, higher mode) on the mode's inspiral bins
// (hm_inspiral_bin before the cuts' inspiral end, the bins that the general
// waves skip; blocks (b, c, 4 modes)) with 43 compact accumulators; waves (<=
// chunks) per binary and mode, chunk 0's zeroing the rest of its partial
// sums; the (3,2)'s wave writes those bins' zero cotangent of the (2,2)
template <class T, bool Grid>
__device__ __forceinline__ void xhm_hm_ins_adjoint_wave(
    const T* __restrict__ head, const T* __restrict__ dhead, const T* __restrict__ modes,
    const T* __restrict__ dmodes, int64_t nb, const PowerRow<T>* __restrict__ table, int64_t n, int64_t chunks,
    int64_t waves, const vec2_t<T>* __restrict__ cot_plus, const vec2_t<T>* __restrict__ cot_cross,
    vec2_t<T>* __restrict__ gh22, double* __restrict__ partial, T* __restrict__ gfm, const int* __restrict__ cuts) {
    const int64_t x = blockIdx.x;
    const int k = (int)(x % kHmModes);
    const int64_t c = (x / kHmModes) % waves;
    const int64_t b = x / kHmModes / waves;
    if (b >= nb) return;
    const int lane = threadIdx.x;
    if (c == 0 && lane < kXhmGradSlots)
        for (int64_t rest = waves; rest < chunks; ++rest)
            partial[((b * kXhmGradRows + 13 + k) * chunks + rest) * kXhmGradSlots + lane] = 0.0;
    __shared__ T s_mode[kHmSlots];
    for (int j = lane; j < kHmSlots; j += 64) s_mode[j] = modes[(b * kHmModes + k) * kHmSlots + j];
    __syncthreads();
    const T* r = s_mode;
    const bool writes = k == kHmMixMode;
    double* out = partial + ((b * kXhmGradRows + 13 + k) * chunks + c) * kXhmGradSlots;
    const int64_t end = cuts[kCutsPer * (b * kHmModes + k) + 3];
    if (r[kHmOn] == T(0) || c * 64 >= end) {
        // the mode vanishes (the general wave writes its every bin), or no
        // inspiral bin falls to this wave
        if (lane < kXhmGradSlots) out[lane] = 0.0;
        return;
    }
    const T* fac = head + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T* dfac = dhead + b * kHeadSlots + kHeadFactors + 4 * (k + 1);
    const T emm = (T)kHmEmm[k];
    T acc[kInsAcc];
#pragma unroll
    for (int j = 0; j < kInsAcc; ++j) acc[j] = T(0);
    T v4 = T(0), v5 = T(0), v6 = T(0);
    for (int64_t i = c * 64 + lane; i < end; i += waves * 64) {
        const PowerRow<T> row = table[i];
        const BinPowers<T> p{row.f, row.t, row.m1, row.lf, row.m76, row.m53};
        if (!hm_inspiral_bin(r, r[kHmMsec] * p.f)) continue;  // the general wave's bin
        const vec2_t<T> cp = cot_plus[b * n + i], cc = cot_cross[b * n + i];
        T g_re, g_im;
        mode_cotangent(fac, cp, cc, g_re, g_im);
        T h_re, h_im, gf = T(0);
        hm_adjoint_inspiral<Grid>(r, p, g_re, g_im, acc, h_re, h_im, gf);
        v4 += g_re * h_re + g_im * h_im;
        v6 += emm * (g_im * h_re - g_re * h_im);
        v5 += inclination_term(dfac, cp, cc, h_re, h_im);
        if (writes) gh22[b * n + i] = vec2_t<T>{T(0), T(0)};
        if (Grid) gfm[(k * nb + b) * n + i] = gf;
    }
    // lane j: the total of accumulator j, contracted with its slot's
    // derivatives
    T full[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) full[j] = j < kInsAcc ? acc[j] : T(0);
    const double total = (double)reduce_scatter64(full, lane);
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    if (lane < kInsAcc) {
        const T* t = dmodes + ((b * kHmModes + k) * kHmSlots + kInsAccSlot[lane]) * 4;
        for (int j = 0; j < 4; ++j) d[j] = t[j] == T(0) ? 0.0 : total * (double)t[j];
    }
    xhm_grad_store(d, v4, v5, v6, lane, out);
}

template <bool Grid>
__global__ __launch_bounds__(64) void xhm_hm_ins_adjoint_kernel(
    const float* __restrict__ head, const float* __restrict__ dhead, const float* __restrict__ modes,
    const float* __restrict__ dmodes, int64_t nb, const PowerRow<float>* __restrict__ table, int64_t n,
    int64_t chunks, int64_t waves, const float2* __restrict__ cot_plus, const float2* __restrict__ cot_cross,
    float2* __restrict__ gh22, double* __restrict__ partial, float* __restrict__ gfm, const int* __restrict__ cuts) {
    xhm_hm_ins_adjoint_wave<float, Grid>(head, dhead, modes, dmodes, nb, table, n, chunks, waves, cot_plus,
                                         cot_cross, gh22, partial, gfm, cuts);
}
template <bool Grid>
__global__ __launch_bounds__(64) void xhm_hm_ins_adjoint_f64_kernel(
    const double* __restrict__ head, const double* __restrict__ dhead, const double* __restrict__ modes,
    const double* __restrict__ dmodes, int64_t nb, const PowerRow<double>* __restrict__ table, int64_t n,
    int64_t chunks, int64_t waves, const double2* __restrict__ cot_plus, const double2* __restrict__ cot_cross,
    double2* __restrict__ gh22, double* __restrict__ partial, double* __restrict__ gfm,
    const int* __restrict__ cuts) {
    xhm_hm_ins_adjoint_wave<double, Grid>(head, dhead, modes, dmodes, nb, table, n, chunks, waves, cot_plus,
                                          cot_cross, gh22, partial, gfm, cuts);
}

// one wave per (binary, chunk): the (2,2)'s adjoint (the IMRPhenomXAS
// record's, with the mode's own cotangent plus the one the (3,2) passes);
// with Grid, each bin's dL/df of the five modes into gfm[0][b][i]
template <class T, bool Grid>
__global__ __launch_bounds__(64) void xhm_22_adjoint_kernel(
    const T* __restrict__ rec22, const double* __restrict__ jac22, const T* __restrict__ head,
    const T* __restrict__ dhead, int64_t nb, const PowerRow<T>* __restrict__ table, int64_t n, int64_t chunks,
    const vec2_t<T>* __restrict__ cot_plus, const vec2_t<T>* __restrict__ cot_cross,
    const vec2_t<T>* __restrict__ gh22, double* __restrict__ partial, T* __restrict__ gfm) {
    const int64_t b = blockIdx.x / chunks;
    const int64_t c = blockIdx.x % chunks;
    if (b >= nb) return;
    const int lane = threadIdx.x;
    T rec[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) rec[k] = rec22[b * kSlots + k];
    T acc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) acc[k] = T(0);
    const T* fac = head + b * kHeadSlots + kHeadFactors;
    const T* dfac = dhead + b * kHeadSlots + kHeadFactors;
    T v4 = T(0), v5 = T(0), v6 = T(0);
    // each iteration loads the next one's powers and cotangents before its
    // own arithmetic (236 -> 214 us at 4096 x 8193)
    const int64_t step = chunks * 64;
    int64_t i = c * 64 + lane;
    PowerRow<T> next_row{};
    vec2_t<T> next_cp{}, next_cc{}, next_extra{};
    if (i < n) {
        next_row = table[i];
        next_cp = cot_plus[b * n + i];
        next_cc = cot_cross[b * n + i];
        next_extra = gh22[b * n + i];
    }
    for (; i < n; i += step) {
        const PowerRow<T> row = next_row;
        const vec2_t<T> cp = next_cp, cc = next_cc, extra = next_extra;
        if (i + step < n) {
            next_row = table[i + step];
            next_cp = cot_plus[b * n + i + step];
            next_cc = cot_cross[b * n + i + step];
            next_extra = gh22[b * n + i + step];
        }
        const BinPowers<T> p{row.f, row.t, row.m1, row.lf, row.m76, row.m53};
        T g_re, g_im;
        mode_cotangent(fac, cp, cc, g_re, g_im);
        T amp, s, co;
        adjoint_point_k(rec, p, T(1), T(0), g_re + extra.x, g_im + extra.y, T(0), T(0), acc, amp, s, co);
        const T h_re = amp * co, h_im = amp * s;
        v4 += g_re * h_re + g_im * h_im;
        v6 += T(2) * (g_im * h_re - g_re * h_im);
        v5 += inclination_term(dfac, cp, cc, h_re, h_im);
        if (Grid) {
            T gf = grid_adjoint_point_k(rec, p, T(1), T(0), g_re + extra.x, g_im + extra.y, T(0), T(0));
            for (int k = 0; k < kHmModes; ++k) gf += gfm[(k * nb + b) * n + i];
            gfm[b * n + i] = gf;
        }
    }
    double total = (double)reduce_scatter64(acc, lane);
    // slots without a derivative here: the factors (the mode's are the
    // head's), T0's and A's low parts and the FP64 copies 60..63
    if (lane == kKp || lane == kKc || lane == kT0Lo || lane == kAnchorLo || lane >= 60) total = 0.0;
    const double* jrow = jac22 + (b * kSlots + lane) * 8;
    double d[4];
    for (int j = 0; j < 4; ++j) d[j] = jrow[j] == 0.0 ? 0.0 : total * jrow[j];
    xhm_grad_store(d, v4, v5, v6, lane, partial + (b * kXhmGradRows * chunks + c) * kXhmGradSlots);
}

// dL/df: partial[g][i] = the sum over binaries g * 64 .. g * 64 + 63 of
// gf[b][i] in order, then dldf[i] = the sum over the groups in order (no
// atomics: the same bits on every run)
template <class T>
__global__ __launch_bounds__(kBlock) void xhm_grid_sum_kernel(const T* __restrict__ gf, int64_t nb, int64_t n,
                                                              int64_t groups, double* __restrict__ grid_partial) {
    const int64_t tiles = (n + kBlock - 1) / kBlock;
    const int64_t g = blockIdx.x / tiles;
    const int64_t i = (blockIdx.x % tiles) * kBlock + threadIdx.x;
    if (g >= groups || i >= n) return;
    const int64_t b0 = g * kXhmGridGroup;
    const int64_t b1 = b0 + kXhmGridGroup < nb ? b0 + kXhmGridGroup : nb;
    double s = 0.0;
    for (int64_t b = b0; b < b1; ++b) s += (double)gf[b * n + i];
    grid_partial[g * n + i] = s;
}

// each binary's gradient (and loss) from the waves' sums, in a fixed order:
// d_L's derivative -Re conj(G) h / d_L, f_ref's half phi_ref's times the
// (2,2)'s phase turn per unit f_ref (its record's inspiral constant's
// derivative less its anchor's: zero where f_ref is zero or mass-scaled).
// Eight threads per binary, thread j summing column j of its rows (and the
// one of column 7, which is zero, the loss), each sum in row order.
constexpr int64_t kXhmFinishPerBlock = 64 / kXhmGradSlots;
__global__ __launch_bounds__(64) void xhm_adjoint_finish_kernel(const double* __restrict__ params,
                                                                const double* __restrict__ jac22, int64_t nb,
                                                                const double* __restrict__ partial, int64_t chunks,
                                                                const double* __restrict__ loss_partial,
                                                                int64_t loss_chunks, double* __restrict__ loss,
                                                                double* __restrict__ grads) {
    const int64_t b = (int64_t)blockIdx.x * kXhmFinishPerBlock + threadIdx.x / kXhmGradSlots;
    const int j = (int)(threadIdx.x % kXhmGradSlots);
    const bool live = b < nb;
    double s = 0.0, l = 0.0;
    if (live) {
        // unrolled so that the loads of a batch are in flight together
#pragma unroll 9
        for (int64_t rc = 0; rc < kXhmGradRows * chunks; ++rc)
            s += partial[(b * kXhmGradRows * chunks + rc) * kXhmGradSlots + j];
        if (loss && j == kXhmGradSlots - 1) {
#pragma unroll 8
            for (int64_t c = 0; c < loss_chunks; ++c) l += loss_partial[b * loss_chunks + c];
        }
    }
    // the binary's sums of columns 4 and 6, to every one of its threads
    const int first = (int)(threadIdx.x & ~(kXhmGradSlots - 1));
    const double s4 = __shfl(s, first + 4, 64), s6 = __shfl(s, first + 6, 64);
    if (!live) return;
    double v = s;
    if (j == 4) {
        v = -s4 / params[b * kParams + kDistance];
    } else if (j == 7) {
        const double turn =
            jac22[(b * kSlots + kInsConst) * 8 + kFRef] - jac22[(b * kSlots + kAnchor) * 8 + kFRef];
        v = 0.5 * s6 * turn;
        if (loss) loss[b] = l;
    }
    grads[b * 8 + j] = v;
}

// the reference phase with its derivatives, for model Model behind `name`
template <int Model>
int reference_phase_grad(const char* name, const double* params, int64_t n_binaries, const double* mf,
                         double* phase, double* grad, void* stream) {
    const std::string who(name);
    if (n_binaries < 0) return fail((who + ": negative size").c_str());
    if (n_binaries == 0) return 0;
    if (!params || !mf || !grad) return fail((who + ": null pointer").c_str());
    const int blocks = grid_blocks((5 * n_binaries + 63) / 64, (who + ": too many binaries").c_str());
    if (blocks < 0) return -1;
    reference_phase_grad_kernel<Model><<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, mf, phase, grad);
    return check_launch(name);
}

}  // namespace

// ------------------------------------------ the scratch-free fused loss

// The device scratch of wfx_xas_loss_grad_static: one call's records, their
// derivatives and the grid's power table, so that the call allocates
// nothing. Each GPU the library runs on holds its own copy until the
// process ends; every call on a GPU shares it, so calls that overlap on one
// GPU race. Host code never passes the array's name to a kernel (on the
// host it is a stand-in, not a device address): it passes the address that
// hipGetSymbolAddress gives for the current device.
constexpr size_t kStaticScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices];

// the scratch's address on the current device, looked up once per device
int scratch_address(const char* who, unsigned char** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail((std::string(who) + ": no current HIP device").c_str());
    if (dev < 0 || dev >= kMaxDevices) return fail((std::string(who) + ": device index beyond 63").c_str());
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr)
            return fail((std::string(who) + ": hipGetSymbolAddress(scratch): " + hipGetErrorString(err)).c_str());
        g_scratch_ptr[dev] = ptr;
    }
    *out = (unsigned char*)g_scratch_ptr[dev];
    return 0;
}

constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

// ================================================================= C ABI

extern "C" {

const char* wfx_last_error(void) { return g_error.c_str(); }

// bumped whenever an entry point's signature changes
int wfx_abi_version(void) { return 25; }

int wfx_xas_record_slots(void) { return kSlots; }

int wfx_xas_record_bytes(int precision) {
    if (precision == kPrecF64) return kSlots * 8;
    if (precision == kPrecF32 || precision == kPrecSplit || precision == kPrecSplitDF) return kSlots * 4;
    return fail("unknown precision");
}

int wfx_xas_setup(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                  double amp_scale, double reference_mf, int precision, void* records, uint32_t* flags,
                  void* stream) {
    if (n_binaries < 0) return fail("wfx_xas_setup: negative number of binaries");
    if (n_binaries == 0) return 0;
    if (!params || !f_min || !records) return fail("wfx_xas_setup: null pointer");
    if (f_min_stride != 0 && f_min_stride != 1) return fail("wfx_xas_setup: f_min_stride must be 0 or 1");
    const int blocks = grid_blocks((n_binaries + 63) / 64, "wfx_xas_setup: too many binaries");
    if (blocks < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    switch (precision) {
        case kPrecF64:
            setup_f64_kernel<<<blocks, 64, 0, s>>>(params, n_binaries, f_min, f_min_stride, amp_scale, reference_mf,
                                                   (double*)records, flags);
            break;
        case kPrecSplit:
        case kPrecSplitDF:
            setup_split_kernel<<<blocks, 64, 0, s>>>(params, n_binaries, f_min, f_min_stride, amp_scale,
                                                     reference_mf, (float*)records, flags);
            break;
        case kPrecF32:
            setup_f32_kernel<<<blocks, 64, 0, s>>>(params, n_binaries, f_min, f_min_stride, amp_scale, reference_mf,
                                                   (float*)records, flags);
            break;
        default:
            return fail("wfx_xas_setup: unknown precision");
    }
    return check_launch("wfx_xas_setup");
}

int wfx_xas_setup_jvp(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                      double amp_scale, double reference_mf, double* records, float* split_records, double* jacobian,
                      uint32_t* flags, void* stream) {
    return setup_jvp<XasRecord>("wfx_xas_setup_jvp", params, n_binaries, f_min, f_min_stride, amp_scale,
                                reference_mf, records, split_records, jacobian, flags, stream);
}

int wfx_xas_reference_phase(const double* params, int64_t n_binaries, const double* mf, double* phase,
                            void* stream) {
    if (n_binaries < 0) return fail("wfx_xas_reference_phase: negative size");
    if (n_binaries == 0) return 0;
    if (!params || !mf || !phase) return fail("wfx_xas_reference_phase: null pointer");
    const int blocks = grid_blocks((n_binaries + 63) / 64, "wfx_xas_reference_phase: too many binaries");
    if (blocks < 0) return -1;
    reference_phase_kernel<<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, mf, phase);
    return check_launch("wfx_xas_reference_phase");
}

int wfx_xas_fast_atan(const float* x, float* y, int64_t n, void* stream) {
    if (n < 0) return fail("wfx_xas_fast_atan: negative size");
    if (n == 0) return 0;
    if (!x || !y) return fail("wfx_xas_fast_atan: null pointer");
    const int blocks = grid_blocks((n + kBlock - 1) / kBlock, "wfx_xas_fast_atan: too large");
    if (blocks < 0) return -1;
    fast_atan_kernel<<<blocks, kBlock, 0, (hipStream_t)stream>>>(x, y, n);
    return check_launch("wfx_xas_fast_atan");
}

int wfx_xas_reference_phase_grad(const double* params, int64_t n_binaries, const double* mf, double* phase,
                                 double* grad, void* stream) {
    return reference_phase_grad<0>("wfx_xas_reference_phase_grad", params, n_binaries, mf, phase, grad, stream);
}

int wfx_xas_eval(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                 int precision, int group, int freqs_per_thread, void* hplus, void* hcross, void* stream) {
    return eval_shared<XasModel>("wfx_xas_eval", records, n_binaries, freqs, freqs_f64, n_freqs, precision, group,
                                 freqs_per_thread, hplus, hcross, stream);
}

int wfx_xas_tile_prefix(const int64_t* offsets, int64_t n_binaries, int bins_per_thread, int64_t* tiles,
                        void* stream) {
    if (n_binaries < 0) return fail("wfx_xas_tile_prefix: negative size");
    if (!offsets || !tiles) return fail("wfx_xas_tile_prefix: null pointer");
    if (bins_per_thread != 1 && bins_per_thread != 2 && bins_per_thread != 4 && bins_per_thread != 8 &&
        bins_per_thread != 16)
        return fail("wfx_xas_tile_prefix: bins_per_thread must be 1, 2, 4, 8 or 16");
    tile_prefix_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(offsets, n_binaries, (int64_t)kBlock * bins_per_thread,
                                                              tiles);
    return check_launch("wfx_xas_tile_prefix");
}

int wfx_xas_eval_ragged(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                        const double* grid_f0, const double* grid_df, const int64_t* offsets, const int64_t* tiles,
                        int precision, int bins_per_thread, int blocks, void* hplus, void* hcross, void* stream) {
    return eval_ragged<XasModel>("wfx_xas_eval_ragged", records, n_binaries, freqs, freqs_f64, grid_f0, grid_df,
                                 offsets, tiles, precision, bins_per_thread, blocks, hplus, hcross, stream);
}

// ------------------------------------------------------------- IMRPhenomD

int wfx_phd_setup(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                  double amp_scale, double reference_mf, int precision, void* records, uint32_t* flags,
                  void* stream) {
    if (n_binaries < 0) return fail("wfx_phd_setup: negative number of binaries");
    if (n_binaries == 0) return 0;
    if (!params || !f_min || !records) return fail("wfx_phd_setup: null pointer");
    if (f_min_stride != 0 && f_min_stride != 1) return fail("wfx_phd_setup: f_min_stride must be 0 or 1");
    if (precision != kPrecF64 && precision != kPrecSplit && precision != kPrecSplitDF)
        return fail("wfx_phd_setup: precision must be 0, 2 or 3 (the table of IMRPhenomD is FP64)");
    const int blocks = grid_blocks((n_binaries + 63) / 64, "wfx_phd_setup: too many binaries");
    if (blocks < 0) return -1;
    const bool f64 = precision == kPrecF64;
    setup_pd_kernel<<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, f_min, f_min_stride, amp_scale,
                                                            reference_mf, f64 ? (double*)records : nullptr,
                                                            f64 ? nullptr : (float*)records, flags);
    return check_launch("wfx_phd_setup");
}

int wfx_phd_reference_phase(const double* params, int64_t n_binaries, const double* mf, double* phase,
                            void* stream) {
    if (n_binaries < 0) return fail("wfx_phd_reference_phase: negative size");
    if (n_binaries == 0) return 0;
    if (!params || !mf || !phase) return fail("wfx_phd_reference_phase: null pointer");
    const int blocks = grid_blocks((n_binaries + 63) / 64, "wfx_phd_reference_phase: too many binaries");
    if (blocks < 0) return -1;
    reference_phase_pd_kernel<<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, mf, phase);
    return check_launch("wfx_phd_reference_phase");
}

int wfx_phd_reference_phase_grad(const double* params, int64_t n_binaries, const double* mf, double* phase,
                                 double* grad, void* stream) {
    return reference_phase_grad<1>("wfx_phd_reference_phase_grad", params, n_binaries, mf, phase, grad, stream);
}

int wfx_phd_eval(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                 int precision, int group, int freqs_per_thread, void* hplus, void* hcross, void* stream) {
    if (precision == kPrecF32) return fail("wfx_phd_eval: precision must be 0, 2 or 3 (no FP32 table)");
    return eval_shared<PhenomDModel>("wfx_phd_eval", records, n_binaries, freqs, freqs_f64, n_freqs, precision, group,
                                     freqs_per_thread, hplus, hcross, stream);
}

int wfx_phd_eval_ragged(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                        const double* grid_f0, const double* grid_df, const int64_t* offsets, const int64_t* tiles,
                        int precision, int bins_per_thread, int blocks, void* hplus, void* hcross, void* stream) {
    if (precision == kPrecF32) return fail("wfx_phd_eval_ragged: precision must be 0, 2 or 3 (no FP32 table)");
    return eval_ragged<PhenomDModel>("wfx_phd_eval_ragged", records, n_binaries, freqs, freqs_f64, grid_f0, grid_df,
                                     offsets, tiles, precision, bins_per_thread, blocks, hplus, hcross, stream);
}

int wfx_xas_power_table_bytes(int precision) {
    if (precision == kPrecF64) return (int)sizeof(PowerRow<double>);
    if (precision == kPrecF32 || precision == kPrecSplit || precision == kPrecSplitDF)
        return (int)sizeof(PowerRow<float>);
    return fail("wfx_xas_power_table_bytes: unknown precision");
}

int wfx_xas_power_table(const void* freqs, int freqs_f64, int64_t n_freqs, int precision, void* table,
                        void* stream) {
    if (n_freqs < 0) return fail("wfx_xas_power_table: negative size");
    if (n_freqs == 0) return 0;
    if (!freqs || !table) return fail("wfx_xas_power_table: null pointer");
    const int blocks = grid_blocks((n_freqs + kBlock - 1) / kBlock, "wfx_xas_power_table: too many frequencies");
    if (blocks < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    const bool f64 = freqs_f64 != 0;
    if (precision == kPrecF64) {
        if (f64)
            power_table_kernel<double, double><<<blocks, kBlock, 0, s>>>((const double*)freqs, n_freqs,
                                                                         (PowerRow<double>*)table);
        else
            power_table_kernel<double, float><<<blocks, kBlock, 0, s>>>((const float*)freqs, n_freqs,
                                                                        (PowerRow<double>*)table);
    } else if (precision == kPrecF32 || precision == kPrecSplit || precision == kPrecSplitDF) {
        if (f64)
            power_table_kernel<float, double><<<blocks, kBlock, 0, s>>>((const double*)freqs, n_freqs,
                                                                        (PowerRow<float>*)table);
        else
            power_table_kernel<float, float><<<blocks, kBlock, 0, s>>>((const float*)freqs, n_freqs,
                                                                       (PowerRow<float>*)table);
    } else {
        return fail("wfx_xas_power_table: unknown precision");
    }
    return check_launch("wfx_xas_power_table");
}

int wfx_xas_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs, int freqs_f64,
                    int64_t n_freqs, int precision, const void* cot_plus, const void* cot_cross, const void* powers,
                    double* grads, void* stream) {
    return adjoint<XasAdjoint, kFromCotangents>("wfx_xas_adjoint", records, jacobian, n_binaries, freqs, freqs_f64,
                                                n_freqs, precision, cot_plus, cot_cross, nullptr, powers, nullptr,
                                                grads, stream);
}

int64_t wfx_xas_grid_adjoint_workspace(int64_t n_binaries, int64_t n_freqs) {
    if (n_binaries <= 0 || n_freqs <= 0) return 0;
    return grid_adjoint_bytes(n_binaries, n_freqs);
}

int wfx_xas_grid_adjoint(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                         int precision, const void* cot_plus, const void* cot_cross, double* dldf, void* workspace,
                         int64_t workspace_bytes, void* stream) {
    return grid_adjoint<XasAdjoint>("wfx_xas_grid_adjoint", records, n_binaries, freqs, freqs_f64, n_freqs, precision,
                                    cot_plus, cot_cross, dldf, workspace, workspace_bytes, stream);
}

int wfx_xas_loss_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs,
                         int freqs_f64, int64_t n_freqs, int precision, const void* data_plus, const void* data_cross,
                         const void* weights, const void* powers, double* loss, double* grads, void* stream) {
    return adjoint<XasAdjoint, kLoss>("wfx_xas_loss_adjoint", records, jacobian, n_binaries, freqs, freqs_f64,
                                      n_freqs, precision, data_plus, data_cross, weights, powers, loss, grads,
                                      stream);
}

int wfx_xas_loss_grad_static(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                             double amp_scale, double reference_mf, const void* freqs, int freqs_f64,
                             int64_t n_freqs, int precision, const void* data_plus, const void* data_cross,
                             const void* weights, double* loss, double* grads, void* stream) {
    const char* who = "wfx_xas_loss_grad_static";
    if (n_binaries < 0 || n_freqs < 0) return fail("wfx_xas_loss_grad_static: negative size");
    if (n_binaries == 0) return 0;
    if (precision != kPrecF64 && precision != kPrecSplit && precision != kPrecSplitDF)
        return fail("wfx_xas_loss_grad_static: precision must be 0, 2 or 3");
    if (!params || !f_min || !loss || !grads || (n_freqs && (!freqs || !data_plus || !data_cross || !weights)))
        return fail("wfx_xas_loss_grad_static: null pointer");
    const bool f64 = precision == kPrecF64;
    // the scratch's three parts: records, derivatives, power table
    const size_t record_bytes = align256((size_t)n_binaries * kSlots * (f64 ? 8 : 4));
    const size_t jac_bytes = align256((size_t)n_binaries * kSlots * 8 * sizeof(double));
    const size_t row = f64 ? sizeof(PowerRow<double>) : sizeof(PowerRow<float>);
    const size_t power_bytes = align256((size_t)n_freqs * row);
    if (n_binaries > (int64_t)(kStaticScratchBytes / (kSlots * 8 * sizeof(double))) ||
        n_freqs > (int64_t)(kStaticScratchBytes / row) ||
        record_bytes + jac_bytes + power_bytes > kStaticScratchBytes)
        return fail("wfx_xas_loss_grad_static: B (4096 + 512) + N 48 bytes (FP64; B (4096 + 256) + N 32 "
                    "split) exceed the 48 MiB scratch: give fewer binaries per call");
    unsigned char* base = nullptr;
    if (scratch_address(who, &base) != 0) return -1;
    void* records = base;
    double* jacobian = (double*)(base + record_bytes);
    void* powers = base + record_bytes + jac_bytes;
    if (setup_jvp<XasRecord>(who, params, n_binaries, f_min, f_min_stride, amp_scale, reference_mf,
                             f64 ? (double*)records : nullptr, f64 ? nullptr : (float*)records, jacobian, nullptr,
                             stream) != 0)
        return -1;
    if (n_freqs > 0 && wfx_xas_power_table(freqs, freqs_f64, n_freqs, precision, powers, stream) != 0) return -1;
    return adjoint<XasAdjoint, kLoss>(who, records, jacobian, n_binaries, freqs, freqs_f64, n_freqs, precision,
                                      data_plus, data_cross, weights, powers, loss, grads, stream);
}

int wfx_phd_setup_jvp(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                      double amp_scale, double reference_mf, double* records, float* split_records, double* jacobian,
                      uint32_t* flags, void* stream) {
    return setup_jvp<PhenomDRecord>("wfx_phd_setup_jvp", params, n_binaries, f_min, f_min_stride, amp_scale,
                                    reference_mf, records, split_records, jacobian, flags, stream);
}

int wfx_phd_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs, int freqs_f64,
                    int64_t n_freqs, int precision, const void* cot_plus, const void* cot_cross, const void* powers,
                    double* grads, void* stream) {
    if (precision == kPrecF32) return fail("wfx_phd_adjoint: IMRPhenomD has no FP32-lead records");
    return adjoint<PhenomDAdjoint, kFromCotangents>("wfx_phd_adjoint", records, jacobian, n_binaries, freqs,
                                                    freqs_f64, n_freqs, precision, cot_plus, cot_cross, nullptr,
                                                    powers, nullptr, grads, stream);
}

int wfx_phd_grid_adjoint(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                         int precision, const void* cot_plus, const void* cot_cross, double* dldf, void* workspace,
                         int64_t workspace_bytes, void* stream) {
    if (precision == kPrecF32) return fail("wfx_phd_grid_adjoint: IMRPhenomD has no FP32-lead records");
    return grid_adjoint<PhenomDAdjoint>("wfx_phd_grid_adjoint", records, n_binaries, freqs, freqs_f64, n_freqs,
                                        precision, cot_plus, cot_cross, dldf, workspace, workspace_bytes, stream);
}

int wfx_phd_loss_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs,
                         int freqs_f64, int64_t n_freqs, int precision, const void* data_plus, const void* data_cross,
                         const void* weights, const void* powers, double* loss, double* grads, void* stream) {
    if (precision == kPrecF32) return fail("wfx_phd_loss_adjoint: IMRPhenomD has no FP32-lead records");
    return adjoint<PhenomDAdjoint, kLoss>("wfx_phd_loss_adjoint", records, jacobian, n_binaries, freqs, freqs_f64,
                                          n_freqs, precision, data_plus, data_cross, weights, powers, loss, grads,
                                          stream);
}


int wfx_xhm_record_slots(void) { return kHmSlots; }

int wfx_xhm_head_slots(void) { return kHeadSlots; }

int wfx_xhm_setup(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                  double amp_scale, double reference_mf, int precision, void* records22, void* head, void* modes,
                  uint32_t* flags, void* stream) {
    if (n_binaries < 0) return fail("wfx_xhm_setup: negative number of binaries");
    if (n_binaries == 0) return 0;
    if (!params || !f_min || !records22 || !head || !modes) return fail("wfx_xhm_setup: null pointer");
    if (f_min_stride != 0 && f_min_stride != 1) return fail("wfx_xhm_setup: f_min_stride must be 0 or 1");
    if (precision != kPrecF64 && precision != kPrecSplit)
        return fail("wfx_xhm_setup: precision must be 0 (FP64) or 2 (split)");
    const int blocks = grid_blocks((kHmModes + 1) * ((n_binaries + 63) / 64), "wfx_xhm_setup: too many binaries");
    if (blocks < 0) return -1;
    xhm_setup_kernel<<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, f_min, f_min_stride, amp_scale,
                                                             reference_mf, precision == kPrecSplit, records22, flags,
                                                             head, modes);
    return check_launch("wfx_xhm_setup");
}

int64_t wfx_xhm_setup_jvp_workspace(int64_t n_binaries) {
    if (n_binaries <= 0) return 0;
    if (n_binaries > INT64_MAX / (4 * (int64_t)sizeof(XasCoeffs22<Dual<1>>))) return -1;
    return 4 * n_binaries * (int64_t)sizeof(XasCoeffs22<Dual<1>>);
}

int wfx_xhm_setup_jvp(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                      double amp_scale, double reference_mf, int precision, void* records22, double* jacobian22,
                      uint32_t* flags, void* head, void* dhead, void* modes, void* dmodes, void* workspace,
                      int64_t workspace_bytes, void* stream) {
    if (n_binaries < 0) return fail("wfx_xhm_setup_jvp: negative number of binaries");
    if (n_binaries == 0) return 0;
    if (!params || !f_min || !records22 || !jacobian22 || !head || !dhead || !modes || !dmodes)
        return fail("wfx_xhm_setup_jvp: null pointer");
    if (f_min_stride != 0 && f_min_stride != 1) return fail("wfx_xhm_setup_jvp: f_min_stride must be 0 or 1");
    if (precision != kPrecF64 && precision != kPrecSplit)
        return fail("wfx_xhm_setup_jvp: precision must be 0 (FP64) or 2 (split)");
    const int64_t need = wfx_xhm_setup_jvp_workspace(n_binaries);
    if (need < 0) return fail("wfx_xhm_setup_jvp: too many binaries");
    if (!workspace || workspace_bytes < need || (uintptr_t)workspace % 256)
        return fail("wfx_xhm_setup_jvp: the workspace must be 256-aligned and hold wfx_xhm_setup_jvp_workspace() "
                    "bytes");
    const int blocks22 = grid_blocks((4 * n_binaries + 63) / 64, "wfx_xhm_setup_jvp: too many binaries");
    const int blocks =
        grid_blocks((4 * kHmModes + 1) * ((n_binaries + 63) / 64), "wfx_xhm_setup_jvp: too many binaries");
    if (blocks22 < 0 || blocks < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    XasCoeffs22<Dual<1>>* kept = (XasCoeffs22<Dual<1>>*)workspace;
    const bool split = precision == kPrecSplit;
    setup_jvp_kernel<XasRecord, true><<<blocks22, 64, 0, s>>>(
        params, n_binaries, f_min, f_min_stride, amp_scale, reference_mf, split ? nullptr : (double*)records22,
        split ? (float*)records22 : nullptr, jacobian22, flags, kept);
    xhm_setup_jvp_kernel<<<blocks, 64, 0, s>>>(params, n_binaries, f_min, f_min_stride, amp_scale, reference_mf,
                                               split, kept, head, dhead, modes, dmodes);
    return check_launch("wfx_xhm_setup_jvp");
}

int64_t wfx_xhm_adjoint_workspace(int64_t n_binaries, int64_t n_freqs, int precision, int loss, int grid) {
    if (n_binaries <= 0 || n_freqs <= 0) return 0;
    if (precision != kPrecF64 && precision != kPrecSplit) return -1;
    return xhm_grad_plan(n_binaries, n_freqs, loss != 0, grid != 0, precision == kPrecF64 ? 8 : 4).bytes;
}

}  // extern "C"

namespace {

// the launches of wfx_xhm_adjoint in T (float: split, double: FP64)
template <class T>
void xhm_adjoint_launch(const XhmGradPlan& g, const double* params, const T* records22, const double* jacobian22,
                        const T* head, const T* dhead, const T* modes, const T* dmodes, int64_t nb, const T* freqs,
                        int64_t n, const void* in_plus, const void* in_cross, const T* weights, char* ws,
                        double* loss, double* grads, double* dldf, hipStream_t s, int table_blocks, int hm_blocks,
                        int blocks22, int loss_blocks, int grid_blocks_1, int grid_blocks_2) {
    using V = vec2_t<T>;
    PowerRow<T>* table = (PowerRow<T>*)(ws + g.table);
    V* gh22 = (V*)(ws + g.gh22);
    double* partial = (double*)(ws + g.partial);
    T* gfm = dldf ? (T*)(ws + g.gfm) : nullptr;
    const V* cot_plus = (const V*)in_plus;
    const V* cot_cross = (const V*)in_cross;
    power_table_kernel<T, T><<<table_blocks, kBlock, 0, s>>>(freqs, n, table);
    // the cuts, then the grid's order flag after them
    int* cuts = (int*)(ws + g.cuts);
    int* ascends = cuts + kCutsPer * nb * kHmModes;
    grid_order_kernel<T><<<1, kOrderBlock, 0, s>>>(table, n, ascends);
    xhm_cuts_kernel<T><<<(int)(nb * kHmModes), 64, 0, s>>>(modes, nb, table, n, ascends, cuts);
    if (weights) {
        V* res_plus = (V*)(ws + g.cot_plus);
        V* res_cross = (V*)(ws + g.cot_cross);
        xhm_residual_kernel<T><<<loss_blocks, kXhmBlock, 0, s>>>(records22, head, modes, nb, table, n, g.loss_chunks,
                                                                 cot_plus, cot_cross, weights, res_plus, res_cross,
                                                                 (double*)(ws + g.loss_partial));
        cot_plus = res_plus;
        cot_cross = res_cross;
    }
    if (std::is_same<T, float>::value) {
        if (dldf)
            xhm_hm_adjoint_kernel<true><<<hm_blocks, 64, 0, s>>>(
                (const float*)records22, (const float*)head, (const float*)dhead, (const float*)modes,
                (const float*)dmodes, nb, (const PowerRow<float>*)table, n, g.chunks, g.gen_chunks,
                (const float2*)cot_plus,
                (const float2*)cot_cross, (float2*)gh22, partial, (float*)gfm, cuts);
        else
            xhm_hm_adjoint_kernel<false><<<hm_blocks, 64, 0, s>>>(
                (const float*)records22, (const float*)head, (const float*)dhead, (const float*)modes,
                (const float*)dmodes, nb, (const PowerRow<float>*)table, n, g.chunks, g.gen_chunks,
                (const float2*)cot_plus,
                (const float2*)cot_cross, (float2*)gh22, partial, nullptr, cuts);
    } else {
        if (dldf)
            xhm_hm_adjoint_f64_kernel<true><<<hm_blocks, 64, 0, s>>>(
                (const double*)records22, (const double*)head, (const double*)dhead, (const double*)modes,
                (const double*)dmodes, nb, (const PowerRow<double>*)table, n, g.chunks, g.gen_chunks,
                (const double2*)cot_plus, (const double2*)cot_cross, (double2*)gh22, partial, (double*)gfm, cuts);
        else
            xhm_hm_adjoint_f64_kernel<false><<<hm_blocks, 64, 0, s>>>(
                (const double*)records22, (const double*)head, (const double*)dhead, (const double*)modes,
                (const double*)dmodes, nb, (const PowerRow<double>*)table, n, g.chunks, g.gen_chunks,
                (const double2*)cot_plus, (const double2*)cot_cross, (double2*)gh22, partial, nullptr, cuts);
    }
    // the intermediate and the inspiral waves
    const int inter_blocks = (int)(kHmModes * nb * g.inter_chunks);
    if (std::is_same<T, float>::value) {
        const float *hd = (const float*)head, *dh = (const float*)dhead, *md = (const float*)modes,
                    *dm = (const float*)dmodes;
        const PowerRow<float>* tb = (const PowerRow<float>*)table;
        if (dldf)
            xhm_hm_inter_adjoint_kernel<true><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const float2*)cot_plus,
                (const float2*)cot_cross, (float2*)gh22, partial, (float*)gfm, cuts);
        else
            xhm_hm_inter_adjoint_kernel<false><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const float2*)cot_plus,
                (const float2*)cot_cross, (float2*)gh22, partial, nullptr, cuts);
        // the inspiral waves
        if (dldf)
            xhm_hm_ins_adjoint_kernel<true><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const float2*)cot_plus,
                (const float2*)cot_cross, (float2*)gh22, partial, (float*)gfm, cuts);
        else
            xhm_hm_ins_adjoint_kernel<false><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const float2*)cot_plus,
                (const float2*)cot_cross, (float2*)gh22, partial, nullptr, cuts);
    } else {
        const double *hd = (const double*)head, *dh = (const double*)dhead, *md = (const double*)modes,
                     *dm = (const double*)dmodes;
        const PowerRow<double>* tb = (const PowerRow<double>*)table;
        if (dldf)
            xhm_hm_inter_adjoint_f64_kernel<true><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const double2*)cot_plus,
                (const double2*)cot_cross, (double2*)gh22, partial, (double*)gfm, cuts);
        else
            xhm_hm_inter_adjoint_f64_kernel<false><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const double2*)cot_plus,
                (const double2*)cot_cross, (double2*)gh22, partial, nullptr, cuts);
        if (dldf)
            xhm_hm_ins_adjoint_f64_kernel<true><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const double2*)cot_plus,
                (const double2*)cot_cross, (double2*)gh22, partial, (double*)gfm, cuts);
        else
            xhm_hm_ins_adjoint_f64_kernel<false><<<inter_blocks, 64, 0, s>>>(
                hd, dh, md, dm, nb, tb, n, g.chunks, g.inter_chunks, (const double2*)cot_plus,
                (const double2*)cot_cross, (double2*)gh22, partial, nullptr, cuts);
    }
    // the ringdown waves: the plain modes, then the mixing (3,2) (its
    // cotangent of the (2,2) complete before the (2,2)'s waves)
    const int rd_blocks = 3 * blocks22, mix_blocks = blocks22;
    if (std::is_same<T, float>::value) {
        using V2 = float2;
        const float *r22 = (const float*)records22, *hd = (const float*)head, *dh = (const float*)dhead,
                    *md = (const float*)modes, *dm = (const float*)dmodes;
        const PowerRow<float>* tb = (const PowerRow<float>*)table;
        if (dldf) {
            xhm_hm_rd_adjoint_kernel<true, false><<<rd_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n, g.chunks,
                (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, (float*)gfm, cuts);
            xhm_hm_rd_adjoint_kernel<true, true><<<mix_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n, g.chunks,
                (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, (float*)gfm, cuts);
        } else {
            xhm_hm_rd_adjoint_kernel<false, false><<<rd_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n,
                g.chunks, (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, nullptr, cuts);
            xhm_hm_rd_adjoint_kernel<false, true><<<mix_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n,
                g.chunks, (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, nullptr, cuts);
        }
    } else {
        using V2 = double2;
        const double *r22 = (const double*)records22, *hd = (const double*)head, *dh = (const double*)dhead,
                     *md = (const double*)modes, *dm = (const double*)dmodes;
        const PowerRow<double>* tb = (const PowerRow<double>*)table;
        if (dldf) {
            xhm_hm_rd_adjoint_f64_kernel<true, false><<<rd_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n,
                g.chunks, (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, (double*)gfm, cuts);
            xhm_hm_rd_adjoint_f64_kernel<true, true><<<mix_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n,
                g.chunks, (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, (double*)gfm, cuts);
        } else {
            xhm_hm_rd_adjoint_f64_kernel<false, false><<<rd_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n,
                g.chunks, (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, nullptr, cuts);
            xhm_hm_rd_adjoint_f64_kernel<false, true><<<mix_blocks, 64, 0, s>>>(r22, hd, dh, md, dm, nb, tb, n,
                g.chunks, (const V2*)cot_plus, (const V2*)cot_cross, (V2*)gh22, partial, nullptr, cuts);
        }
    }
    if (dldf)
        xhm_22_adjoint_kernel<T, true><<<blocks22, 64, 0, s>>>(records22, jacobian22, head, dhead, nb, table, n,
                                                               g.chunks, cot_plus, cot_cross, gh22, partial,
                                                               gfm);
    else
        xhm_22_adjoint_kernel<T, false><<<blocks22, 64, 0, s>>>(records22, jacobian22, head, dhead, nb, table, n,
                                                                g.chunks, cot_plus, cot_cross, gh22, partial,
                                                                nullptr);
    xhm_adjoint_finish_kernel<<<(int)((nb + kXhmFinishPerBlock - 1) / kXhmFinishPerBlock), 64, 0, s>>>(
        params, jacobian22, nb, partial, g.chunks, weights ? (const double*)(ws + g.loss_partial) : nullptr,
        g.loss_chunks, weights ? loss : nullptr, grads);
    if (dldf) {
        double* grid_partial = (double*)(ws + g.grid_partial);
        xhm_grid_sum_kernel<T><<<grid_blocks_1, kBlock, 0, s>>>(gfm, nb, n, g.grid_groups, grid_partial);
        group_sum_kernel<<<grid_blocks_2, kBlock, 0, s>>>(grid_partial, g.grid_groups, n, dldf);
    }
}

}  // namespace

extern "C" {

int wfx_xhm_adjoint(const double* params, const void* records22, const double* jacobian22, const void* head,
                    const void* dhead, const void* modes, const void* dmodes, int64_t n_binaries, const void* freqs,
                    int64_t n_freqs, int precision, const void* in_plus, const void* in_cross, const void* weights,
                    void* workspace, int64_t workspace_bytes, double* loss, double* grads, double* dldf,
                    void* stream) {
    if (n_binaries < 0 || n_freqs < 0) return fail("wfx_xhm_adjoint: negative size");
    if (n_binaries == 0) return 0;
    if (!params || !records22 || !jacobian22 || !head || !dhead || !modes || !dmodes || !grads)
        return fail("wfx_xhm_adjoint: null pointer");
    if (precision != kPrecF64 && precision != kPrecSplit)
        return fail("wfx_xhm_adjoint: precision must be 0 (FP64) or 2 (split)");
    if (weights && !loss) return fail("wfx_xhm_adjoint: the loss needs its output");
    hipStream_t s = (hipStream_t)stream;
    if (n_freqs == 0) {
        // no frequency: a zero gradient and loss (a kernel, not a memset
        // node, should the caller capture the call in a graph)
        const int blocks = grid_blocks((n_binaries + kXhmFinishPerBlock - 1) / kXhmFinishPerBlock,
                                       "wfx_xhm_adjoint: grid too large");
        if (blocks < 0) return -1;
        xhm_adjoint_finish_kernel<<<blocks, 64, 0, s>>>(params, jacobian22, n_binaries, nullptr, 0, nullptr, 0,
                                                        weights ? loss : nullptr, grads);
        return check_launch("wfx_xhm_adjoint");
    }
    if (!freqs || !in_plus || !in_cross) return fail("wfx_xhm_adjoint: null pointer");
    const bool f64 = precision == kPrecF64;
    const XhmGradPlan g = xhm_grad_plan(n_binaries, n_freqs, weights != nullptr, dldf != nullptr, f64 ? 8 : 4);
    if (g.bytes < 0) return fail("wfx_xhm_adjoint: sizes too large");
    if (!workspace || workspace_bytes < g.bytes || (uintptr_t)workspace % 256)
        return fail("wfx_xhm_adjoint: the workspace must be 256-aligned and hold wfx_xhm_adjoint_workspace() bytes");
    const int64_t tiles = (n_freqs + kBlock - 1) / kBlock;
    const int table_blocks = grid_blocks(tiles, "wfx_xhm_adjoint: grid too large");
    const int hm_blocks = grid_blocks(n_binaries * g.gen_chunks * kHmModes, "wfx_xhm_adjoint: grid too large");
    const int blocks22 = grid_blocks(n_binaries * g.chunks, "wfx_xhm_adjoint: grid too large");
    const int loss_blocks = grid_blocks(n_binaries * g.loss_chunks, "wfx_xhm_adjoint: grid too large");
    const int grid_1 = grid_blocks(g.grid_groups * tiles, "wfx_xhm_adjoint: grid too large");
    if (table_blocks < 0 || hm_blocks < 0 || blocks22 < 0 || loss_blocks < 0 || grid_1 < 0) return -1;
    char* ws = (char*)workspace;
    if (f64)
        xhm_adjoint_launch<double>(g, params, (const double*)records22, jacobian22, (const double*)head,
                                   (const double*)dhead, (const double*)modes, (const double*)dmodes, n_binaries,
                                   (const double*)freqs, n_freqs, in_plus, in_cross, (const double*)weights, ws, loss,
                                   grads, dldf, s, table_blocks, hm_blocks, blocks22, loss_blocks, grid_1,
                                   table_blocks);
    else
        xhm_adjoint_launch<float>(g, params, (const float*)records22, jacobian22, (const float*)head,
                                  (const float*)dhead, (const float*)modes, (const float*)dmodes, n_binaries,
                                  (const float*)freqs, n_freqs, in_plus, in_cross, (const float*)weights, ws, loss,
                                  grads, dldf, s, table_blocks, hm_blocks, blocks22, loss_blocks, grid_1,
                                  table_blocks);
    return check_launch("wfx_xhm_adjoint");
}

int wfx_xhm_eval(const void* records22, const void* head, const void* modes, int64_t n_binaries, const void* freqs,
                 int freqs_f64, int64_t n_freqs, int precision, void* hplus, void* hcross, void* mode_out,
                 void* stream) {
    if (n_binaries < 0 || n_freqs < 0) return fail("wfx_xhm_eval: negative size");
    if (n_binaries == 0 || n_freqs == 0) return 0;
    if (!records22 || !head || !modes || !freqs || !hplus || !hcross) return fail("wfx_xhm_eval: null pointer");
    if (precision != kPrecF64 && precision != kPrecSplit)
        return fail("wfx_xhm_eval: precision must be 0 (FP64) or 2 (split)");
    if (precision == kPrecSplit && freqs_f64) return fail("wfx_xhm_eval: the split evaluation takes a float32 grid");
    const int64_t chunks = (n_freqs + kXhmBlock - 1) / kXhmBlock;
    if (n_binaries > INT64_MAX / chunks) return fail("wfx_xhm_eval: grid too large");
    const int blocks = grid_blocks(n_binaries * chunks, "wfx_xhm_eval: grid too large");
    if (blocks < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (precision == kPrecSplit)
        xhm_eval_split_kernel<<<blocks, kXhmBlock, 0, s>>>((const float*)records22, (const float*)head,
                                                           (const float*)modes, n_binaries, (const float*)freqs,
                                                           n_freqs, chunks, (float2*)hplus, (float2*)hcross,
                                                           (float2*)mode_out);
    else if (freqs_f64)
        xhm_eval_kernel<double><<<blocks, kXhmBlock, 0, s>>>((const double*)records22, (const double*)head,
                                                              (const double*)modes, n_binaries, (const double*)freqs,
                                                              n_freqs, chunks, (double2*)hplus, (double2*)hcross,
                                                              (double2*)mode_out);
    else
        xhm_eval_kernel<float><<<blocks, kXhmBlock, 0, s>>>((const double*)records22, (const double*)head,
                                                             (const double*)modes, n_binaries, (const float*)freqs,
                                                             n_freqs, chunks, (double2*)hplus, (double2*)hcross,
                                                             (double2*)mode_out);
    return check_launch("wfx_xhm_eval");
}

int wfx_xhm_eval_ragged(const void* records22, const void* head, const void* modes, int64_t n_binaries,
                        const void* freqs, int freqs_f64, const double* grid_f0, const double* grid_df,
                        const int64_t* offsets, const int64_t* tiles, int precision, int blocks, void* hplus,
                        void* hcross, void* mode_out, void* stream) {
    if (n_binaries < 0) return fail("wfx_xhm_eval_ragged: negative size");
    if (n_binaries == 0) return 0;
    if (!records22 || !head || !modes || !offsets || !tiles || !hplus || !hcross)
        return fail("wfx_xhm_eval_ragged: null pointer");
    if (!freqs && (!grid_f0 || !grid_df)) return fail("wfx_xhm_eval_ragged: give freqs or (grid_f0, grid_df)");
    if (precision != kPrecF64 && precision != kPrecSplit)
        return fail("wfx_xhm_eval_ragged: precision must be 0 (FP64) or 2 (split)");
    if (precision == kPrecSplit && freqs && freqs_f64)
        return fail("wfx_xhm_eval_ragged: the split evaluation takes a float32 grid");
    if (blocks <= 0) blocks = 8 * cu_count();
    hipStream_t s = (hipStream_t)stream;
    if (precision == kPrecSplit) {
        xhm_eval_split_ragged_kernel<float><<<blocks, kXhmBlock, 0, s>>>(
            (const float*)records22, (const float*)head, (const float*)modes, n_binaries, (const float*)freqs,
            grid_f0, grid_df, offsets, tiles, (float2*)hplus, (float2*)hcross, (float2*)mode_out);
    } else if (freqs && !freqs_f64) {
        xhm_eval_ragged_kernel<float><<<blocks, kXhmBlock, 0, s>>>(
            (const double*)records22, (const double*)head, (const double*)modes, n_binaries, (const float*)freqs,
            grid_f0, grid_df, offsets, tiles, (double2*)hplus, (double2*)hcross, (double2*)mode_out);
    } else {
        xhm_eval_ragged_kernel<double><<<blocks, kXhmBlock, 0, s>>>(
            (const double*)records22, (const double*)head, (const double*)modes, n_binaries, (const double*)freqs,
            grid_f0, grid_df, offsets, tiles, (double2*)hplus, (double2*)hcross, (double2*)mode_out);
    }
    return check_launch("wfx_xhm_eval_ragged");
}

int wfx_xhm_project(const double* records22, const double* head, int n_heads, const double* modes,
                    int64_t n_binaries, const double* freqs, int64_t n_freqs, const void* cot_plus,
                    const void* cot_cross, int64_t n_cot, double* partial, void* stream) {
    if (n_binaries < 0 || n_freqs < 0 || n_cot < 0) return fail("wfx_xhm_project: negative size");
    if (n_heads < 1 || n_heads > kMaxProjectHeads) return fail("wfx_xhm_project: n_heads must be 1 to 4");
    if (n_binaries == 0 || n_freqs == 0) return 0;
    if (n_cot == 0) return fail("wfx_xhm_project: no cotangent rows");
    if (!records22 || !head || !modes || !freqs || !cot_plus || !cot_cross || !partial)
        return fail("wfx_xhm_project: null pointer");
    const int64_t chunks = (n_freqs + kXhmBlock - 1) / kXhmBlock;
    if (n_binaries > INT64_MAX / chunks) return fail("wfx_xhm_project: grid too large");
    const int blocks = grid_blocks(n_binaries * chunks, "wfx_xhm_project: grid too large");
    if (blocks < 0) return -1;
    xhm_project_kernel<<<blocks, kXhmBlock, 0, (hipStream_t)stream>>>(
        records22, head, n_heads, modes, n_binaries, freqs, n_freqs, chunks, (const double2*)cot_plus,
        (const double2*)cot_cross, n_cot, partial);
    return check_launch("wfx_xhm_project");
}

}  // extern "C"
