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
#include "xas_point.h"

#include <atomic>
#include <cstdio>
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

// FP32 record from the FP64 setup (split precision): every slot rounded to
// FP32; T0 and A also as hi + lo and as doubles in slots 60..63
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
    float* out = records + b * kSlots;
    for (int k = 0; k < 60; ++k) out[k] = (float)rec[k];
    const double t0 = rec[kT0], anchor = rec[kAnchor];
    out[kT0] = (float)t0;
    out[kT0Lo] = (float)(t0 - (double)out[kT0]);
    out[kAnchor] = (float)anchor;
    out[kAnchorLo] = (float)(anchor - (double)out[kAnchor]);
    *reinterpret_cast<double*>(out + kT0F64) = t0;
    *reinterpret_cast<double*>(out + kAnchorF64) = anchor;
    if (flags) flags[b] = fl;
}

// plain FP32 record: the whole setup in FP32
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
    rec[kT0Lo] = 0.0f;
    rec[kAnchorLo] = 0.0f;
    float* out = records + b * kSlots;
    for (int k = 0; k < 60; ++k) out[k] = rec[k];
    *reinterpret_cast<double*>(out + kT0F64) = (double)rec[kT0];
    *reinterpret_cast<double*>(out + kAnchorF64) = (double)rec[kAnchor];
    if (flags) flags[b] = fl;
}

// the FP64 record and its derivatives with respect to the 8 parameters
// (m1, m2, chi1, chi2, distance, inclination, phi_ref, f_ref): one thread
// per (binary, direction); jac (B, 64, 8)
__global__ __launch_bounds__(64) void setup_jvp_kernel(const double* __restrict__ params, int64_t nb,
                                                       const double* __restrict__ f_min, int64_t f_min_stride,
                                                       double amp_scale, double reference_mf,
                                                       double* __restrict__ records, double* __restrict__ jac,
                                                       uint32_t* __restrict__ flags) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = gid >> 3;
    const int dir = (int)(gid & 7);
    if (b >= nb) return;
    using D = Dual<1>;
    D p[kParams];
    for (int k = 0; k < kParams; ++k) {
        p[k] = D(params[b * kParams + k]);
        p[k].d[0] = k == dir ? 1.0 : 0.0;
    }
    with_reference_mf(p, reference_mf);
    D rec[kSlots];
    const unsigned fl = xas_record<D>(p, D(f_min[b * f_min_stride]), amp_scale, rec);
    for (int k = 0; k < kSlots; ++k) jac[(b * kSlots + k) * 8 + dir] = k < 60 ? rec[k].d[0] : 0.0;
    if (dir == 0) {
        for (int k = 0; k < kSlots; ++k) records[b * kSlots + k] = rec[k].v;
        records[b * kSlots + kT0Lo] = 0.0;
        records[b * kSlots + kAnchorLo] = 0.0;
        if (flags) flags[b] = fl;
    }
}

// ------------------------------------------------------------ evaluation

constexpr int kBlock = 256;

template <class F>
__device__ inline double load_freq(const F* freqs, int64_t i) {
    return (double)freqs[i];
}

// shared grid, FP64
template <class F>
__global__ __launch_bounds__(kBlock) void eval_f64_shared(const double* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, int64_t n, int64_t ntiles,
                                                          int group, double2* __restrict__ hp,
                                                          double2* __restrict__ hc) {
    const int64_t w = xcd_work_id(blockIdx.x, gridDim.x);
    const int64_t tile = w % ntiles;
    const int64_t b0 = (w / ntiles) * group;
    const int64_t i = tile * kBlock + threadIdx.x;
    const bool valid = i < n;
    const FreqPowers<double> fp = freq_powers<double>(load_freq(freqs, valid ? i : n - 1));
    const int64_t b1 = b0 + group < nb ? b0 + group : nb;
    for (int64_t b = b0; b < b1; ++b) {
        double a, bb, c, d;
        point_f64(records + b * kSlots, fp, a, bb, c, d);
        if (valid) {
            hp[b * n + i] = make_double2(a, bb);
            hc[b * n + i] = make_double2(c, d);
        }
    }
}

// shared grid, FP32 paths
template <int L, class F>
__global__ __launch_bounds__(kBlock) void eval_f32_shared(const float* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, int64_t n, int64_t ntiles,
                                                          int group, float2* __restrict__ hp,
                                                          float2* __restrict__ hc) {
    const int64_t w = xcd_work_id(blockIdx.x, gridDim.x);
    const int64_t tile = w % ntiles;
    const int64_t b0 = (w / ntiles) * group;
    const int64_t i = tile * kBlock + threadIdx.x;
    const bool valid = i < n;
    const PowF32 p = pow_f32((float)freqs[valid ? i : n - 1]);
    const int64_t b1 = b0 + group < nb ? b0 + group : nb;
    for (int64_t b = b0; b < b1; ++b) {
        float2 a, c;
        point_f32<L>(records + b * kSlots, p, a, c);
        if (valid) {
            hp[b * n + i] = a;
            hc[b * n + i] = c;
        }
    }
}

// ragged grids: work item w -> (binary, tile) through the prefix sum of
// tiles per binary (tiles[b] .. tiles[b + 1])
__device__ inline int64_t owner(const int64_t* __restrict__ tiles, int64_t nb, int64_t w) {
    int64_t lo = 0, hi = nb;  // tiles[lo] <= w < tiles[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (tiles[mid] <= w)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <class F>
__device__ inline double ragged_freq(const F* freqs, const double* f0, const double* df, int64_t b, int64_t i,
                                     int64_t k) {
    return freqs ? (double)freqs[i] : f0[b] + (double)k * df[b];
}

template <class F>
__global__ __launch_bounds__(kBlock) void eval_f64_ragged(const double* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, const double* __restrict__ f0,
                                                          const double* __restrict__ df,
                                                          const int64_t* __restrict__ offsets,
                                                          const int64_t* __restrict__ tiles,
                                                          double2* __restrict__ hp, double2* __restrict__ hc) {
    const int64_t total = tiles[nb];
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int64_t b = owner(tiles, nb, w);
        const int64_t k = (w - tiles[b]) * kBlock + threadIdx.x;
        const int64_t i = offsets[b] + k;
        if (i >= offsets[b + 1]) continue;
        const FreqPowers<double> fp = freq_powers<double>(ragged_freq(freqs, f0, df, b, i, k));
        double a, bb, c, d;
        point_f64(records + b * kSlots, fp, a, bb, c, d);
        hp[i] = make_double2(a, bb);
        hc[i] = make_double2(c, d);
    }
}

template <int L, class F>
__global__ __launch_bounds__(kBlock) void eval_f32_ragged(const float* __restrict__ records, int64_t nb,
                                                          const F* __restrict__ freqs, const double* __restrict__ f0,
                                                          const double* __restrict__ df,
                                                          const int64_t* __restrict__ offsets,
                                                          const int64_t* __restrict__ tiles,
                                                          float2* __restrict__ hp, float2* __restrict__ hc) {
    const int64_t total = tiles[nb];
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int64_t b = owner(tiles, nb, w);
        const int64_t k = (w - tiles[b]) * kBlock + threadIdx.x;
        const int64_t i = offsets[b] + k;
        if (i >= offsets[b + 1]) continue;
        const PowF32 p = pow_f32((float)ragged_freq(freqs, f0, df, b, i, k));
        float2 a, c;
        point_f32<L>(records + b * kSlots, p, a, c);
        hp[i] = a;
        hc[i] = c;
    }
}

// tiles[b] = sum_{j < b} ceil(len_j / kBlock): one block, a running prefix
// over chunks of 1024 binaries
__global__ __launch_bounds__(1024) void tile_prefix_kernel(const int64_t* __restrict__ offsets, int64_t nb,
                                                           int64_t* __restrict__ tiles) {
    __shared__ int64_t scan[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t b = base + threadIdx.x;
        int64_t len = b < nb ? offsets[b + 1] - offsets[b] : 0;
        if (len < 0) len = 0;
        scan[threadIdx.x] = (len + kBlock - 1) / kBlock;
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

int grid_blocks(int64_t nblocks, const char* what) {
    if (nblocks <= 0) return 0;
    if (nblocks > 0x7fffffff) {
        fail(what);
        return -1;
    }
    return (int)nblocks;
}

}  // namespace

// ================================================================= C ABI

extern "C" {

const char* wfx_last_error(void) { return g_error.c_str(); }

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
                      double amp_scale, double reference_mf, double* records, double* jacobian, uint32_t* flags,
                      void* stream) {
    if (n_binaries < 0) return fail("wfx_xas_setup_jvp: negative number of binaries");
    if (n_binaries == 0) return 0;
    if (!params || !f_min || !records || !jacobian) return fail("wfx_xas_setup_jvp: null pointer");
    if (f_min_stride != 0 && f_min_stride != 1) return fail("wfx_xas_setup_jvp: f_min_stride must be 0 or 1");
    const int blocks = grid_blocks((8 * n_binaries + 63) / 64, "wfx_xas_setup_jvp: too many binaries");
    if (blocks < 0) return -1;
    setup_jvp_kernel<<<blocks, 64, 0, (hipStream_t)stream>>>(params, n_binaries, f_min, f_min_stride, amp_scale,
                                                             reference_mf, records, jacobian, flags);
    return check_launch("wfx_xas_setup_jvp");
}

int wfx_xas_eval(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                 int precision, int group, void* hplus, void* hcross, void* stream) {
    if (n_binaries < 0 || n_freqs < 0) return fail("wfx_xas_eval: negative size");
    if (n_binaries == 0 || n_freqs == 0) return 0;
    if (!records || !freqs || !hplus || !hcross) return fail("wfx_xas_eval: null pointer");
    if (group <= 0) group = precision == kPrecF64 ? 8 : 16;
    const int64_t ntiles = (n_freqs + kBlock - 1) / kBlock;
    const int64_t ngroups = (n_binaries + group - 1) / group;
    const int blocks = grid_blocks(ntiles * ngroups, "wfx_xas_eval: grid too large, raise the group size");
    if (blocks < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
#define WFX_F32_LAUNCH(L)                                                                                           \
    if (freqs_f64)                                                                                                  \
        eval_f32_shared<L, double><<<blocks, kBlock, 0, s>>>((const float*)records, n_binaries,                     \
                                                             (const double*)freqs, n_freqs, ntiles, group,          \
                                                             (float2*)hplus, (float2*)hcross);                      \
    else                                                                                                            \
        eval_f32_shared<L, float><<<blocks, kBlock, 0, s>>>((const float*)records, n_binaries, (const float*)freqs, \
                                                            n_freqs, ntiles, group, (float2*)hplus, (float2*)hcross);
    switch (precision) {
        case kPrecF64:
            if (freqs_f64)
                eval_f64_shared<double><<<blocks, kBlock, 0, s>>>((const double*)records, n_binaries,
                                                                  (const double*)freqs, n_freqs, ntiles, group,
                                                                  (double2*)hplus, (double2*)hcross);
            else
                eval_f64_shared<float><<<blocks, kBlock, 0, s>>>((const double*)records, n_binaries,
                                                                 (const float*)freqs, n_freqs, ntiles, group,
                                                                 (double2*)hplus, (double2*)hcross);
            break;
        case kPrecSplit:
            WFX_F32_LAUNCH(kLeadF64)
            break;
        case kPrecSplitDF:
            WFX_F32_LAUNCH(kLeadDF)
            break;
        case kPrecF32:
            WFX_F32_LAUNCH(kLeadF32)
            break;
        default:
            return fail("wfx_xas_eval: unknown precision");
    }
#undef WFX_F32_LAUNCH
    return check_launch("wfx_xas_eval");
}

int wfx_xas_tile_prefix(const int64_t* offsets, int64_t n_binaries, int64_t* tiles, void* stream) {
    if (n_binaries < 0) return fail("wfx_xas_tile_prefix: negative size");
    if (!offsets || !tiles) return fail("wfx_xas_tile_prefix: null pointer");
    tile_prefix_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(offsets, n_binaries, tiles);
    return check_launch("wfx_xas_tile_prefix");
}

int wfx_xas_eval_ragged(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                        const double* grid_f0, const double* grid_df, const int64_t* offsets, const int64_t* tiles,
                        int precision, int blocks, void* hplus, void* hcross, void* stream) {
    if (n_binaries < 0) return fail("wfx_xas_eval_ragged: negative size");
    if (n_binaries == 0) return 0;
    if (!records || !offsets || !tiles || !hplus || !hcross) return fail("wfx_xas_eval_ragged: null pointer");
    if (!freqs && (!grid_f0 || !grid_df)) return fail("wfx_xas_eval_ragged: give freqs or (grid_f0, grid_df)");
    if (blocks <= 0) blocks = 8 * cu_count();
    hipStream_t s = (hipStream_t)stream;
#define WFX_RAGGED_F32(L)                                                                                      \
    if (freqs_f64 || !freqs)                                                                                   \
        eval_f32_ragged<L, double><<<blocks, kBlock, 0, s>>>((const float*)records, n_binaries,                \
                                                             (const double*)freqs, grid_f0, grid_df, offsets,  \
                                                             tiles, (float2*)hplus, (float2*)hcross);          \
    else                                                                                                       \
        eval_f32_ragged<L, float><<<blocks, kBlock, 0, s>>>((const float*)records, n_binaries,                 \
                                                            (const float*)freqs, grid_f0, grid_df, offsets,    \
                                                            tiles, (float2*)hplus, (float2*)hcross);
    switch (precision) {
        case kPrecF64:
            if (freqs_f64 || !freqs)
                eval_f64_ragged<double><<<blocks, kBlock, 0, s>>>((const double*)records, n_binaries,
                                                                  (const double*)freqs, grid_f0, grid_df, offsets,
                                                                  tiles, (double2*)hplus, (double2*)hcross);
            else
                eval_f64_ragged<float><<<blocks, kBlock, 0, s>>>((const double*)records, n_binaries,
                                                                 (const float*)freqs, grid_f0, grid_df, offsets,
                                                                 tiles, (double2*)hplus, (double2*)hcross);
            break;
        case kPrecSplit:
            WFX_RAGGED_F32(kLeadF64)
            break;
        case kPrecSplitDF:
            WFX_RAGGED_F32(kLeadDF)
            break;
        case kPrecF32:
            WFX_RAGGED_F32(kLeadF32)
            break;
        default:
            return fail("wfx_xas_eval_ragged: unknown precision");
    }
#undef WFX_RAGGED_F32
    return check_launch("wfx_xas_eval_ragged");
}

}  // extern "C"
