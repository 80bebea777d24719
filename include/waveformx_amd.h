/*
 * waveformx_amd: the C ABI of the IMRPhenomXAS hot path on MI355X (gfx950),
 * exported by waveformx_amd/libwfx_xas.so.
 *
 * The reference crosses from PyTorch into its model through DLPack and a
 * jax.jit(jax.vmap(...)) of model.polarizations (src/waveformx/torch.py:
 * 58-106, src/waveformx/_api.py:122-129); these entry points take that
 * place: plain device pointers and sizes, caller-owned memory, work
 * enqueued on the caller's HIP stream (hipStream_t passed as void*).
 *
 * Conventions.
 *  - Every entry point returns 0 on success and -1 on an error that it
 *    detects before launching (bad size, null pointer, unknown precision,
 *    launch failure); wfx_last_error() then names it (per host thread).
 *    Nothing throws across the ABI. Kernel execution is asynchronous.
 *  - No entry point allocates or frees memory; buffers come from the caller.
 *  - Parameters are packed as double params[B][8] = {mass_1, mass_2 (solar
 *    masses), chi_1z, chi_2z, luminosity_distance (Mpc), inclination,
 *    phi_ref (rad), f_ref (Hz, 0 = the first frequency)}; in-plane spins
 *    do not enter (src/waveformx/models/xas/model.py:25-32).
 *  - precision: 0 = FP64 throughout (double records, complex128 out);
 *    1 = FP32 setup and evaluation (float records, complex64), no split;
 *    2 = FP64 setup, FP32 evaluation with the inspiral's leading phase term
 *    and the phase reduction in FP64 (float records, complex64);
 *    3 = as 2 with that term in double-float FP32 arithmetic.
 *  - Outputs are interleaved complex (re, im), row-major (B, N).
 *  - LAL's failures make the binary's row NaN and set flags (bit 0: mass
 *    ratio beyond 1000, 1: |spin| > 1, 2: final mass > 1, 3: |final spin|
 *    > 1, 4: f_MECO > f_ISCO; src/waveformx/models/xpnr/internals.py:136-149).
 *  - Calls on one stream are ordered; calls on different streams that share
 *    an output buffer race, as any kernel would.
 */
#ifndef WAVEFORMX_AMD_H
#define WAVEFORMX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The message of the last failed call on this host thread. */
const char* wfx_last_error(void);

/* The version of this ABI (25); a binding checks it before any call. */
int wfx_abi_version(void);

/* Slots of one binary's record (64) and its size in bytes for a precision
 * (512 for FP64, 256 otherwise). The record replaces the reference's
 * 86-entry table (src/waveformx/models/xas/core.py:68-101). */
int wfx_xas_record_slots(void);
int wfx_xas_record_bytes(int precision);

/* Per-binary setup: records[B][64] (double for precision 0, else float) and
 * flags[B] (may be NULL) of params[B][8] whose grids start at f_min
 * (f_min_stride 0: one value for all, 1: one per binary). reference_mf:
 * NaN for none, else f_ref <- reference_mf / M where f_ref was 0 (NaN
 * otherwise). Replaces core.table (src/waveformx/models/xas/core.py:
 * 257-379) with model.py:35-59's f_min and reference_mf handling. */
int wfx_xas_setup(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                  double amp_scale, double reference_mf, int precision, void* records, uint32_t* flags,
                  void* stream);

/* The records and their derivatives jacobian[B][64][8] (of the FP64
 * record's slots) with respect to the 8 packed parameters (forward mode
 * through the setup), for the adjoint. records: FP64 records as precision 0
 * writes them, split_records: FP32 records as precision 2 and 3 write them;
 * either may be NULL, not both. Replaces the forward half of the
 * reference's jax.vjp (src/waveformx/torch.py:87-93). */
int wfx_xas_setup_jvp(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                      double amp_scale, double reference_mf, double* records, float* split_records,
                      double* jacobian, uint32_t* flags, void* stream);

/* phase[B] = LAL's (2,2) phase of each binary at M f = mf[b] but for its
 * constant, phase_22 / eta + linb M f, in FP64: what the reference-
 * convention converters difference (src/waveformx/models/xas/reference.py:
 * 29-38, src/waveformx/models/taylorf2/reference.py:46-127). */
int wfx_xas_reference_phase(const double* params, int64_t n_binaries, const double* mf, double* phase,
                            void* stream);

/* wfx_xas_reference_phase with its derivatives grad[B][5] = d phase /
 * d(mass_1, mass_2, chi_1z, chi_2z, mf) (forward mode through the same FP64
 * code); phase may be NULL. The derivatives of the reference's converters,
 * which it takes with jax.grad (src/waveformx/models/xas/reference.py:29-38,
 * tests/test_xas.py:174-175). */
int wfx_xas_reference_phase_grad(const double* params, int64_t n_binaries, const double* mf, double* phase,
                                 double* grad, void* stream);

/* (h+, hx)[B][N] of records on one shared grid freqs[N] (float, or double
 * if freqs_f64). group: binaries per block (0: default, 4 on the FP32
 * paths, 8 for FP64); freqs_per_thread: 1, 2, 4 or 8 frequencies per thread
 * on the FP32 paths (0: default, 4); every
 * shape writes the same bits. Replaces core.frequency_powers +
 * core.evaluate under jax.vmap (src/waveformx/models/xas/core.py:382-507;
 * src/waveformx/torch.py:84-85). */
int wfx_xas_eval(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                 int precision, int group, int freqs_per_thread, void* hplus, void* hcross, void* stream);

/* tiles[B+1]: the prefix sum of the tiles (256 x bins_per_thread bins; 1, 2,
 * 4, 8 or 16) of each binary's range [offsets[b], offsets[b+1]), the work list of
 * wfx_xas_eval_ragged. No counterpart in the reference (its batches share
 * one grid). */
int wfx_xas_tile_prefix(const int64_t* offsets, int64_t n_binaries, int bins_per_thread, int64_t* tiles,
                        void* stream);

/* (h+, hx) of binaries on their own grids, flat: binary b owns
 * [offsets[b], offsets[b+1]) of freqs (float or double) and of the
 * outputs, or, with freqs NULL, the uniform grid grid_f0[b] + k
 * grid_df[b]. One persistent launch of `blocks` blocks (0: 8 per CU) over
 * the tiles from wfx_xas_tile_prefix (same bins_per_thread). Generalises the shared-grid batching
 * of src/waveformx/torch.py:84-85 (the MLX port's per-binary grids,
 * src/waveformx/models/xas/mlx/_kernels.py:523-525, need equal lengths). */
int wfx_xas_eval_ragged(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                        const double* grid_f0, const double* grid_df, const int64_t* offsets, const int64_t* tiles,
                        int precision, int bins_per_thread, int blocks, void* hplus, void* hcross,
                        void* stream);

/* A grid's table of frequency powers for the adjoint kernels:
 * wfx_xas_power_table_bytes(precision) bytes per frequency (32 for the FP32
 * paths, 48 for FP64) in the caller's buffer. No counterpart in the
 * reference (XLA shares core.frequency_powers across a vmap itself). */
int wfx_xas_power_table_bytes(int precision);
int wfx_xas_power_table(const void* freqs, int freqs_f64, int64_t n_freqs, int precision, void* table, void* stream);

/* grads[B][8] (double) of a scalar loss from its cotangents cot_plus,
 * cot_cross[B][N] (interleaved complex, dL/dRe + i dL/dIm, precision 0:
 * double, else float) on the shared grid: the adjoint kernel recomputes
 * each bin and contracts with jacobian (from wfx_xas_setup_jvp); records
 * are those of the forward precision; powers: the grid's table from
 * wfx_xas_power_table, or NULL (then computed per bin). Replaces the
 * reverse half of the reference's jax.vjp backward
 * (src/waveformx/torch.py:108-119). */
int wfx_xas_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs, int freqs_f64,
                    int64_t n_freqs, int precision, const void* cot_plus, const void* cot_cross, const void* powers,
                    double* grads, void* stream);

/* dldf[N] = dL/df (double), the gradient with respect to the shared
 * grid's frequencies (the first one's effect through f_min excluded, as the
 * reference's stop_gradient, src/waveformx/models/xas/model.py:59), from
 * the cotangents as wfx_xas_adjoint. workspace: device memory of at least
 * wfx_xas_grid_adjoint_workspace(B, N) bytes (a double per group of 64
 * binaries and frequency), 256-aligned; the groups are summed in a fixed
 * order, so the result is reproducible. Replaces the grid part of the
 * reference's vjp (src/waveformx/torch.py:116). */
int64_t wfx_xas_grid_adjoint_workspace(int64_t n_binaries, int64_t n_freqs);
int wfx_xas_grid_adjoint(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                         int precision, const void* cot_plus, const void* cot_cross, double* dldf, void* workspace,
                         int64_t workspace_bytes, void* stream);

/* wfx_xas_fast_atan: y[i] = the FP32 arctangent of the split and FP32
 * evaluations (atan(x) = x P(x^2) on [0, 1], pi/2 - atan(1/x) beyond, P the
 * float32 polynomial of waveformx_amd/csrc/gen_atan.py), for its accuracy
 * test. It replaces jnp.arctan in those arithmetics' phases (the
 * reference's src/waveformx/models/xas/core.py:441), which is no entry
 * point of the reference. */
int wfx_xas_fast_atan(const float* x, float* y, int64_t n, void* stream);

/* Fused forward + adjoint of L_b = sum_f weights[f] (|h+ - d+|^2 +
 * |hx - dx|^2) for data d (as the cotangents' layout) and weights[N]
 * (double for precision 0, else float): loss[B] and grads[B][8], both
 * double; nothing per bin is written. The consumer side of the path: the
 * reference computes such losses by composing model.polarizations with
 * jax.grad (src/waveformx/match.py:78-81 defines the weighted inner
 * product). */
int wfx_xas_loss_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs,
                         int freqs_f64, int64_t n_freqs, int precision, const void* data_plus, const void* data_cross,
                         const void* weights, const void* powers, double* loss, double* grads, void* stream);

/* wfx_xas_loss_grad_static: loss[B] and grads[B][8] as wfx_xas_loss_adjoint
 * gives them, straight from params[B][8] (with f_min, amp_scale and
 * reference_mf as wfx_xas_setup_jvp), for precision 0, 2 or 3: the same
 * derivative setup, power table and fused loss adjoint, with the records,
 * their derivatives and the table in a static device array of 48 MiB
 * instead of caller buffers. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, serving and
 * sampling loops that forbid hipMalloc). It allocates nothing; it needs
 * B x 4608 + N x 48 bytes (FP64) or B x 4352 + N x 32 (split), each part
 * rounded up to 256, within 48 MiB (8192 binaries x 4017 bins fit), and
 * fails otherwise. Each GPU holds one such array until the process ends,
 * shared by every call on it: calls that overlap on one GPU (two streams,
 * two host threads) race. Every call writes all of the array it reads, so
 * nothing needs clearing between calls. */
int wfx_xas_loss_grad_static(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                             double amp_scale, double reference_mf, const void* freqs, int freqs_f64,
                             int64_t n_freqs, int precision, const void* data_plus, const void* data_cross,
                             const void* weights, double* loss, double* grads, void* stream);

/* IMRPhenomD, on the same conventions. Its table is FP64 whatever the
 * evaluation's arithmetic (src/waveformx/models/phenomd/model.py:33-50):
 * precision 0, 2 or 3 only (1 fails). Records have the 64 slots of
 * IMRPhenomXAS's (wfx_xas_record_bytes) in IMRPhenomD's layout. LAL's flag:
 * bit 3, the final spin above one (phenomd/api.py:39-45); with
 * reference_mf the given f_ref must be zero, else the row is NaN.
 * wfx_phd_setup replaces phenomd core.table (core.py:274-377);
 * wfx_phd_eval and wfx_phd_eval_ragged replace core.frequency_powers +
 * core.evaluate under jax.vmap (core.py:395-485), with the launch shapes of
 * wfx_xas_eval and wfx_xas_eval_ragged (tiles from wfx_xas_tile_prefix);
 * every frequency is evaluated (no cutoff). wfx_phd_reference_phase:
 * phase[b] = -Psi + t0 M f at M f = mf[b] but for its constant, for the
 * reference-convention converters (phenomd/reference.py:18-26). */
int wfx_phd_setup(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                  double amp_scale, double reference_mf, int precision, void* records, uint32_t* flags,
                  void* stream);
int wfx_phd_reference_phase(const double* params, int64_t n_binaries, const double* mf, double* phase,
                            void* stream);
/* IMRPhenomD's reference phase with its derivatives, as
 * wfx_xas_reference_phase_grad (phenomd/reference.py:18-26). */
int wfx_phd_reference_phase_grad(const double* params, int64_t n_binaries, const double* mf, double* phase,
                                 double* grad, void* stream);
int wfx_phd_eval(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                 int precision, int group, int freqs_per_thread, void* hplus, void* hcross, void* stream);
int wfx_phd_eval_ragged(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64,
                        const double* grid_f0, const double* grid_df, const int64_t* offsets, const int64_t* tiles,
                        int precision, int bins_per_thread, int blocks, void* hplus, void* hcross,
                        void* stream);

/* IMRPhenomD's gradients, with the arguments and layouts of
 * wfx_xas_setup_jvp, wfx_xas_adjoint and wfx_xas_loss_adjoint (powers from
 * wfx_xas_power_table: both models read the same frequency powers); the
 * region bounds carry no derivative, as the reference's jnp.where
 * selections (phenomd/core.py:268-269, 426-440). Together with
 * wfx_phd_grid_adjoint (dL/df of the shared grid, as wfx_xas_grid_adjoint)
 * they replace the reference's jax.vjp backward for this model
 * (src/waveformx/torch.py:108-119). Precision 1 fails. */
int wfx_phd_setup_jvp(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                      double amp_scale, double reference_mf, double* records, float* split_records, double* jacobian,
                      uint32_t* flags, void* stream);
int wfx_phd_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs, int freqs_f64,
                    int64_t n_freqs, int precision, const void* cot_plus, const void* cot_cross, const void* powers,
                    double* grads, void* stream);
int wfx_phd_grid_adjoint(const void* records, int64_t n_binaries, const void* freqs, int freqs_f64, int64_t n_freqs,
                         int precision, const void* cot_plus, const void* cot_cross, double* dldf, void* workspace,
                         int64_t workspace_bytes, void* stream);
int wfx_phd_loss_adjoint(const void* records, const double* jacobian, int64_t n_binaries, const void* freqs,
                         int freqs_f64, int64_t n_freqs, int precision, const void* data_plus, const void* data_cross,
                         const void* weights, const void* powers, double* loss, double* grads, void* stream);

/* ------------------------------------------------------------ IMRPhenomXHM
 *
 * IMRPhenomXHM (release 122022, LAL's defaults: every mode, no
 * multibanding), FP64. Replaces the reference's XHM model behind
 * polarizations_function("IMRPhenomXHM") and batch_polarizations
 * (src/waveformx/models/xhm/api.py:63-81, model.py; set-up
 * xhm/reference.py:566-608, evaluation reference.py:620-677).
 *
 * wfx_xhm_record_slots / wfx_xhm_head_slots: the doubles of one
 * higher-mode record and of one binary's head.
 *
 * wfx_xhm_setup: for params (B, 8) and reference_mf as wfx_xas_setup's
 * (xhm/model.py:40-42) and precision 0 (FP64) or 2 (split), writes
 * records22 (B, 64) (the (2,2) mode: wfx_xas_setup's record of that
 * precision: double or float), head (B, wfx_xhm_head_slots()) (the factors
 * of the five modes (2,2), (2,1), (3,3), (3,2), (4,4) in the polarizations,
 * NaN where LAL fails), modes (B, 4, wfx_xhm_record_slots()) (the higher
 * modes) -- head and modes double for FP64, float for split -- and flags
 * (B,) (LAL's error bits, as wfx_xas_setup; may be null). The set-up is
 * FP64 for both.
 *
 * wfx_xhm_eval: (h+, hx) (B, N) on the shared grid freqs (N,) and, if
 * mode_out is not null, the modes h_{l,-m} (5, B, N) in the order above:
 * complex128 for precision 0 (freqs float64 if freqs_f64, else float32),
 * complex64 for precision 2 (split: FP32 with each mode's leading inspiral
 * phase in FP64, the reference's float32 design, xhm/core.py:13-21; freqs
 * float32). */
int wfx_xhm_record_slots(void);
int wfx_xhm_head_slots(void);
int wfx_xhm_setup(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                  double amp_scale, double reference_mf, int precision, void* records22, void* head, void* modes,
                  uint32_t* flags, void* stream);
int wfx_xhm_eval(const void* records22, const void* head, const void* modes, int64_t n_binaries, const void* freqs,
                 int freqs_f64, int64_t n_freqs, int precision, void* hplus, void* hcross, void* mode_out,
                 void* stream);

/* wfx_xhm_eval_ragged: the binaries on their own grids (the reference's
 * per-binary frequencies (B, N), src/waveformx/models/xhm/mlx/_kernels.py:
 * 565-576, generalised to a length per binary): binary b owns [offsets[b],
 * offsets[b + 1]) of the concatenated grid freqs (float32 or, for FP64,
 * float64 if freqs_f64) or, with freqs NULL, the uniform grid grid_f0[b] +
 * k grid_df[b] of that length (computed in FP64); hplus, hcross (total,)
 * and, if mode_out is not null, the modes (5, total), as wfx_xhm_eval
 * writes them for one binary on its grid (the same bits). tiles (B + 1,)
 * from wfx_xas_tile_prefix(offsets, B, 1, tiles) (tiles of 256 bins);
 * blocks the persistent grid's size (0: 8 per CU). offsets must start at
 * 0 and not decrease (checked on the host by the binding). */
int wfx_xhm_eval_ragged(const void* records22, const void* head, const void* modes, int64_t n_binaries,
                        const void* freqs, int freqs_f64, const double* grid_f0, const double* grid_df,
                        const int64_t* offsets, const int64_t* tiles, int precision, int blocks, void* hplus,
                        void* hcross, void* mode_out, void* stream);

/* wfx_xhm_setup_jvp: the gradient's set-up of IMRPhenomXHM (same arguments
 * as wfx_xhm_setup) with its derivatives, forward mode through the same
 * FP64 set-up: records22 (B, 64) the (2,2)'s IMRPhenomXAS record (double
 * for precision 0, the split float record for 2), jacobian22 (B, 64, 8)
 * double its derivatives and flags (B,) LAL's flag bits, as
 * wfx_xas_setup_jvp writes them; the head and the higher-mode records in
 * the split evaluation's layout (each mode's leading inspiral term and
 * inspiral constant as an FP64 slot, the later regions' constants reduced
 * modulo 2 pi, quotients folded): float for precision 2 (wfx_xhm_setup's
 * split records, the FP64 slots as two floats), double for precision 0 (the
 * same slots in FP64, an FP64 slot's second word zero; the FP64
 * evaluation's records are wfx_xhm_setup's, in LAL's form). dhead (B,
 * wfx_xhm_head_slots()) the head's derivative in the inclination, dmodes
 * (B, 4 modes, wfx_xhm_record_slots(), 4 directions) the records'
 * derivatives in (mass_1, mass_2, chi_1z, chi_2z) (an FP64 slot's in its
 * first word). The higher modes take the (2,2)'s coefficients from its
 * set-up through workspace: device memory of at least
 * wfx_xhm_setup_jvp_workspace(B) bytes, 256-aligned (-1: too many
 * binaries). The forward half of the reference's jax.vjp for IMRPhenomXHM
 * (src/waveformx/torch.py:87-93; the MLX port's tables_jvp,
 * src/waveformx/models/xhm/mlx/_kernels.py:551-603). */
int64_t wfx_xhm_setup_jvp_workspace(int64_t n_binaries);
int wfx_xhm_setup_jvp(const double* params, int64_t n_binaries, const double* f_min, int64_t f_min_stride,
                      double amp_scale, double reference_mf, int precision, void* records22, double* jacobian22,
                      uint32_t* flags, void* head, void* dhead, void* modes, void* dmodes, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* wfx_xhm_adjoint: grads (B, 8) double, the gradient in (mass_1, mass_2,
 * chi_1z, chi_2z, luminosity_distance, inclination, phi_ref, f_ref) of the
 * polarizations in the split layout's formula, FP32 (precision 2, freqs
 * float) or FP64 (precision 0, freqs double), on the grid freqs (N,), for
 * the binaries params (B, 8) set up by wfx_xas_setup_jvp (records22
 * (B, 64) the (2,2)'s record of that precision, jacobian22 (B, 64, 8) its
 * FP64 derivatives) and wfx_xhm_setup_jvp (head, dhead, modes, dmodes, the
 * same precision). With weights NULL, in_plus/in_cross are the cotangents
 * (B, N) complex64/complex128 (dL/dRe + i dL/dIm) and loss is not written;
 * else they are the data, weights (N,) float/double, and loss (B,) double
 * receives L = sum_f weights (|h+ - d+|^2 + |hx - dx|^2), the gradient being
 * L's. dldf (N,) double, if not NULL, receives dL/df of each frequency (the
 * grid's gradient, summed over the binaries). Reverse mode: one wave per
 * span of frequencies, mode and region (inspiral, intermediate, ringdown,
 * the rest: on an ascending grid each takes only its own blocks) of a
 * binary recomputes the mode, sums each bin's dL/d(record slot) in
 * registers and contracts the totals with the records' derivatives once
 * (on a grid in any other order every wave tests every bin: the same
 * gradient up to the sums' grouping); nothing per bin is kept but the
 * (3,2)'s
 * cotangent of the (2,2) it recycles (and, for the loss, the residuals; for
 * dldf, each bin's dL/df per mode). workspace: device memory of at least
 * wfx_xhm_adjoint_workspace(B, N, precision, weights != NULL, dldf != NULL)
 * bytes, 256-aligned (B x N x 8 bytes per real's 4, 3x that for the loss,
 * 4x more for dldf, and little more; -1 for an unknown precision). Each sum
 * runs in a fixed order: the result is reproducible. The reverse half of
 * the reference's jax.vjp for IMRPhenomXHM (src/waveformx/torch.py:108-119,
 * the grid's gradient at :116; its GPU port's cotangent_product,
 * src/waveformx/models/xhm/mlx/_kernels.py:300-441). */
int64_t wfx_xhm_adjoint_workspace(int64_t n_binaries, int64_t n_freqs, int precision, int loss, int grid);
int wfx_xhm_adjoint(const double* params, const void* records22, const double* jacobian22, const void* head,
                    const void* dhead, const void* modes, const void* dmodes, int64_t n_binaries, const void* freqs,
                    int64_t n_freqs, int precision, const void* in_plus, const void* in_cross, const void* weights,
                    void* workspace, int64_t workspace_bytes, double* loss, double* grads, double* dldf,
                    void* stream);

/* wfx_xhm_project: the FP64 polarizations of wfx_xhm_eval (FP64 records of
 * wfx_xhm_setup, float64 grid) contracted with cotangents without writing
 * them, for n_heads (1..4) heads (n_heads, B, wfx_xhm_head_slots()) sharing
 * the records and modes (the modes' factors in the polarizations; the head
 * of wfx_xhm_setup, or others): partial (n_heads, B, C), C = ceil(N / 256),
 * partial[j][b][c] = the sum over frequencies 256 c .. 256 c + 255 of
 * Re(conj(cot_plus) h+ + conj(cot_cross) hx) under head j, the cotangents
 * complex128 (n_cot, N), binary b taking row b % n_cot. The inner products
 * of the torch backward's vector-Jacobian product (src/waveformx/torch.py:
 * 108-119) for IMRPhenomXHM, whose gradient is differences of the set-up
 * and evaluation (waveformx_amd/xhm.py). */
int wfx_xhm_project(const double* records22, const double* head, int n_heads, const double* modes,
                    int64_t n_binaries, const double* freqs, int64_t n_freqs, const void* cot_plus,
                    const void* cot_cross, int64_t n_cot, double* partial, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WAVEFORMX_AMD_H */
