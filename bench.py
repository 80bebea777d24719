"""Throughput of the IMRPhenomXAS hot path on one MI355X.

A step is one pass of the hot path over one batch: the per-binary setup
kernel and the per-(binary, frequency) evaluation kernel for 65,536 seeded
binaries on linspace(20, 1024, 4017) Hz, FP32 evaluation with the
inspiral's leading term split off (FP64 setup), complex64 outputs written
to HBM. Inputs are resident on the device before the timed region.

Printed (one JSON line): waveforms/s of that step, the same for the
all-FP64 path and the double-float variant, the mismatch of the split path
against the FP64 path on the device, the dominant kernel's HBM roofline
(algorithmic bytes = 16 B per bin written, over the kernel's own
event-timed duration), and the CPU baseline: the FP64 oracle timed on the
host cores on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the FP32-split evaluation at 65,536 x 4017 in its
# default launch shape, from two rocprofv3 --pmc passes over
# `bench.py --only split` (summary under profiles/): WRITE_SIZE 4,118,276 KiB
# x 1.000 + FETCH_SIZE 16,624 KiB x 2.000, the factors that a 1 GiB device
# copy gives the two counters. Reported only for that workload and shape.
EVAL_TRAFFIC = {(65536, 4017): 4.251e9}
# VALU instructions per bin of the same launch (SQ_INSTS_VALU per 64 bins
# from rocprofv3 --pmc, summary under profiles/; the oracle's op count of the
# formula gives ~78 FP32 operation slots per bin, transcendentals at 4), and
# the non-packed FP32 VALU issue rate: 256 CUs x 4 SIMDs x 64 lanes per 4
# cycles at 2.4 GHz
EVAL_VALU_PER_BIN = {(65536, 4017): 77.5}
VALU_PEAK_TOPS = 256 * 4 * 64 / 4 * 2.4e9 / 1e12
METRIC = "XAS waveforms/s @4017 bins: FP32-split vs FP64 GPU (mismatch<=1e-6); fwd+grad/s"
# VALU instructions per bin of the IMRPhenomXHM evaluations at 4096 x 8193,
# FP64 and split (SQ_INSTS_VALU x 64 / bins from rocprofv3 --pmc, summaries
# under profiles/); FP64 FMA issues at the non-packed FP32 rate on this chip
XHM_VALU_PER_BIN = {(4096, 8193): 1822.4}
XHM_SPLIT_VALU_PER_BIN = {(4096, 8193): 629.9}
# The algorithmic work per bin of the same workload (the oracle's op
# counter, oracle_xhm_count, on its 4096 binaries: the branches each bin
# takes, the M f powers from the frequency's; pinned by
# tests/test_oracle_xhm.py::test_op_count): FP32 issue slots with the
# transcendentals at the quarter rate (xhm.op_slots) for the split
# evaluation; for FP64, its add, multiply, divide and compare operations
# (267.4) beside 17.6 elementary-function calls, which FP64 runs in software
XHM_SPLIT_SLOTS_PER_BIN = {(4096, 8193): 445.9}
XHM_FP64_OPS_PER_BIN = {(4096, 8193): 267.4}
XHM_FP64_CALLS_PER_BIN = {(4096, 8193): 17.6}


def draw_xhm(n, seed=4):
    """The multimode workload's binaries (B, 8): as :func:`draw` with
    q ~ U[0.1, 1]."""
    rng = np.random.default_rng(seed)
    m1 = rng.uniform(10.0, 80.0, n)
    q = rng.uniform(0.1, 1.0, n)
    c1 = rng.uniform(-0.9, 0.9, n)
    c2 = rng.uniform(-0.9, 0.9, n)
    dl = rng.uniform(100.0, 2000.0, n)
    inc = np.arccos(rng.uniform(-1.0, 1.0, n))
    phi = rng.uniform(0.0, 2 * np.pi, n)
    return np.stack([m1, m1 * q, c1, c2, dl, inc, phi, np.full(n, 20.0)], axis=1)


def draw(n, seed=1):
    """The workload's binaries (B, 8): m1 ~ U[10, 80], q ~ U[0.125, 1],
    chi ~ U[-0.9, 0.9], d_L ~ U[100, 2000] Mpc, inclination = arccos U[-1, 1],
    phi_ref ~ U[0, 2 pi], f_ref = 20 Hz."""
    rng = np.random.default_rng(seed)
    m1 = rng.uniform(10.0, 80.0, n)
    q = rng.uniform(0.125, 1.0, n)
    c1 = rng.uniform(-0.9, 0.9, n)
    c2 = rng.uniform(-0.9, 0.9, n)
    dl = rng.uniform(100.0, 2000.0, n)
    inc = np.arccos(rng.uniform(-1.0, 1.0, n))
    phi = rng.uniform(0.0, 2 * np.pi, n)
    return np.stack([m1, m1 * q, c1, c2, dl, inc, phi, np.full(n, 20.0)], axis=1)


#: the least time the card runs a leg's work before its timed calls: after a
#: change of load its clock dips and recovers over a few hundred
#: milliseconds, which a timed region inside that window would average in
WARM_SECONDS = 0.4


def timed(torch, fn, steps, warmup):
    """Average ms of ``fn`` over ``steps`` calls after ``warmup`` calls and
    at least WARM_SECONDS of them (at most 4000), by events on the current
    stream."""
    t0 = time.perf_counter()
    calls = 0
    while calls < warmup or (time.perf_counter() - t0 < WARM_SECONDS and calls < 4000):
        fn()
        calls += 1
        if calls % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def accuracy(torch, test, ref, w, rows=4096):
    """The worst over h+ and hx and over binaries of: the noise-weighted
    mismatch 1 - <a,b> / sqrt(<a,a><b,b>) (<a,b> = sum_f Re(a b*) w(f), no
    time or phase maximisation), the phase error |wrap(arg a - arg b)| over
    bins where b is not zero, and the relative amplitude error
    ||a| - |b|| / |b| over bins where |b| > 1e-6 max|b| of the row; test
    complex64, ref complex128, (B, N), in chunks of rows."""
    worst = {"mismatch": 0.0, "max_phase_error_rad": 0.0, "max_rel_amplitude_error": 0.0}
    for a_all, b_all in ((test[0], ref[0]), (test[1], ref[1])):
        for start in range(0, a_all.shape[0], rows):
            a = a_all[start:start + rows].to(torch.complex128)
            b = b_all[start:start + rows]
            ab = (a * b.conj()).real.mul(w).sum(1)
            aa = a.abs().square().mul(w).sum(1)
            bb = b.abs().square().mul(w).sum(1)
            mism = 1.0 - ab / (aa * bb).sqrt()
            babs = b.abs()
            dphi = torch.where(babs > 0, (a * b.conj()).angle().abs(), torch.zeros_like(babs))
            keep = babs > 1e-6 * babs.max(dim=1, keepdim=True).values
            damp = torch.where(keep, (a.abs() - babs).abs() / torch.where(keep, babs, torch.ones_like(babs)),
                               torch.zeros_like(babs))
            worst["mismatch"] = max(worst["mismatch"], float(mism.max()))
            worst["max_phase_error_rad"] = max(worst["max_phase_error_rad"], float(dphi.max()))
            worst["max_rel_amplitude_error"] = max(worst["max_rel_amplitude_error"], float(damp.max()))
    return worst


def bench_grad(torch, model, dev, nb, n, steps, warmup, approximant="IMRPhenomXAS", prefix=""):
    """The fused forward + adjoint of L = sum_f (|h+ - d+|^2 + |hx - dx|^2) / S(f)
    for nb binaries x n bins (FP32-split), d at parameters perturbed by 1%:
    binaries/s of the whole gradient step (the FP64 setup with derivatives,
    which also writes the FP32-split record; the grid's power table; the
    fused kernel), and its time over the forward step's. ``model``: the
    approximant's module (its setup, evaluate, polarizations); the keys of
    the result start with ``prefix``."""
    from waveformx_amd import adjoint

    freqs = np.linspace(20.0, 1024.0, n)
    p = draw(nb, seed=3)
    rng = np.random.default_rng(4)
    perturbed = p * (1 + 1e-2 * rng.normal(size=p.shape))
    perturbed[:, 7] = p[:, 7]
    params = torch.as_tensor(p, device=dev)
    f32 = torch.as_tensor(freqs, dtype=torch.float32, device=dev)
    d_plus, d_cross = model.polarizations(torch.as_tensor(perturbed, device=dev), f32, 1e21, None, "split")
    x = freqs / 100.0
    w = torch.as_tensor(1.0 / (x ** -4 + 1.0 + x ** 2), dtype=torch.float32, device=dev)
    out = (torch.empty((nb, n), dtype=torch.complex64, device=dev),
           torch.empty((nb, n), dtype=torch.complex64, device=dev))

    def forward():
        records, _ = model.setup(params, f32[:1], 1e21, None, "split")
        model.evaluate(records, f32, "split", out=out)

    def gradient():
        adjoint.loss_and_grad(params, f32, d_plus, d_cross, w, precision="split", approximant=approximant)

    ms_fwd = timed(torch, forward, steps, warmup)
    ms_grad = timed(torch, gradient, steps, warmup)
    # the split gradient against the FP64 adjoint (which matches the FP64
    # oracle's forward-mode gradient to 1e-8) at the same grid and data:
    # per parameter, max over binaries of |g - g64| / max over binaries |g64|
    loss, grads = adjoint.loss_and_grad(params, f32, d_plus, d_cross, w, precision="split", approximant=approximant)
    loss64, grads64 = adjoint.loss_and_grad(params, f32.to(torch.float64), d_plus.to(torch.complex128),
                                            d_cross.to(torch.complex128), w.to(torch.float64), precision="fp64",
                                            approximant=approximant)
    err = ((grads - grads64).abs().amax(0) / grads64.abs().amax(0))[:7]
    loss_err = float(((loss - loss64).abs() / loss64.abs()).max())
    result = {"fwd_grad_value": nb / (ms_grad * 1e-3), "fwd_grad_ms": ms_grad, "fwd_grad_binaries": nb,
              "fwd_grad_over_fwd": ms_grad / ms_fwd,
              "grad_rel_err_vs_fp64": dict(zip(("mass_1", "mass_2", "chi_1", "chi_2", "luminosity_distance",
                                                "inclination", "phi_ref"), map(float, err.cpu()))),
              "loss_rel_err_vs_fp64": loss_err}
    return {prefix + k: v for k, v in result.items()}


def ragged_grids(nb, seed=5):
    """Binaries each on its own uniform grid: f_lower ~ U[10, 30] Hz,
    duration T ~ log-U[4, 64] s (df = 1/T), f_max = min(2048, 0.3 / M),
    n = floor((f_max - f_lower) / df + 0.5) + 1; masses m1 ~ U[5, 100],
    q ~ U[0.125, 1], aligned spins ~ U[-0.9, 0.9]."""
    rng = np.random.default_rng(seed)
    m1 = rng.uniform(5.0, 100.0, nb)
    m2 = m1 * rng.uniform(0.125, 1.0, nb)
    p = draw(nb, seed=seed + 1)
    p[:, 0], p[:, 1] = m1, m2
    f_lower = rng.uniform(10.0, 30.0, nb)
    df = 1.0 / np.exp(rng.uniform(np.log(4.0), np.log(64.0), nb))
    f_max = np.minimum(2048.0, 0.3 / ((m1 + m2) * 4.925490947641267e-06))
    n = (np.floor((f_max - f_lower) / df + 0.5) + 1).astype(np.int64)
    p[:, 7] = f_lower
    return p, [(a, b, int(c)) for a, b, c in zip(f_lower, df, n)]


def bench_ragged(torch, xas, dev, nb, steps, warmup):
    """nb binaries on their own grids (FP32-split): the graph-replayed
    persistent launch (tuned tile and grid shape), against one launch per
    binary and against the batch padded to the longest grid."""
    from waveformx_amd.ragged import RaggedPlan

    p, grids = ragged_grids(nb)
    params = torch.as_tensor(p, device=dev)
    plan = RaggedPlan(grids, precision="split")
    ms_plan = timed(torch, lambda: plan(params), steps, warmup)
    lengths = np.array([g[2] for g in grids])
    # one launch per binary (the setup once for all)
    per = torch.as_tensor(np.stack([np.zeros(nb, np.int64), lengths], 1), device=dev)
    per_tiles = torch.as_tensor(np.stack([np.zeros(nb, np.int64), (lengths + 255) // 256], 1), device=dev)
    starts = np.concatenate([[0], np.cumsum(lengths)])

    blocks_per = np.maximum(np.minimum((lengths + 255) // 256, 2048), 1)

    def per_binary():
        records, _ = xas.setup(params, plan.f_min, 1e21, None, "split")
        for b in range(nb):
            xas.evaluate_ragged(records[b:b + 1], per[b], grid_f0=plan.grid_f0[b:b + 1],
                                grid_df=plan.grid_df[b:b + 1], precision="split", tiles=per_tiles[b],
                                blocks=int(blocks_per[b]), total=int(lengths[b]),
                                out=(plan.hplus[starts[b]:starts[b + 1]], plan.hcross[starts[b]:starts[b + 1]]))

    ms_per = timed(torch, per_binary, 2, 1)
    # every binary padded to the longest grid (its own f_lower and df)
    longest = int(lengths.max())
    padded = RaggedPlan([(g[0], g[1], longest) for g in grids], precision="split", tune=False,
                        bins_per_thread=plan.bins_per_thread, blocks=plan.blocks)
    ms_pad = timed(torch, lambda: padded(params), max(2, steps // 4), 1)
    total = int(lengths.sum())
    del padded
    return {"ragged_binaries": nb, "ragged_bins": total, "ragged_ms": ms_plan,
            "ragged_value": nb / (ms_plan * 1e-3), "ragged_bins_per_s": total / (ms_plan * 1e-3),
            "ragged_per_binary_launches_ms": ms_per, "ragged_padded_ms": ms_pad,
            "ragged_padded_bins": nb * longest, "ragged_shape": [plan.bins_per_thread, plan.blocks]}


XHM_MODE_NAMES = ("22", "21", "33", "32", "44")


def xhm_mode_gradients(torch, xhm, params, perturbed, f32, f64, w32, w64, rows):
    """Per mode, on the first ``rows`` binaries: the loss of that mode alone
    (the other modes' factors in the polarizations zeroed in the set-up's
    head, and in the data's), its split gradient against its FP64 one (both
    the exact adjoint; the (3,2)'s ringdown still recycles the (2,2)): the
    largest error per parameter over the binaries, relative to the
    parameter's largest gradient (max over the 7 parameters), and the worst
    fraction of the reference's float32 bar."""
    params = params[:rows]
    pert = torch.as_tensor(perturbed[:rows], device=params.device)
    out = {}
    for k, name in enumerate(XHM_MODE_NAMES):
        def only(head):
            kept = head[:, 4 * k:4 * k + 4].clone()
            head = head.clone()
            head[:, :20] = 0
            head[:, 4 * k:4 * k + 4] = kept
            return head

        records, head, modes_, _ = xhm.setup(pert, f64[:1])
        d_plus, d_cross = xhm.evaluate(records, only(head), modes_, f64)
        grads = {}
        for precision, grid, w, ctype in (("split", f32, w32, torch.complex64), ("fp64", f64, w64, torch.complex128)):
            j = xhm.setup_jvp(params, grid[:1].to(torch.float64), precision=precision)
            j["head"], j["dhead"] = only(j["head"]), only(j["dhead"])
            grads[precision] = xhm.adjoint_with(j, params, grid, d_plus.to(ctype), d_cross.to(ctype), w,
                                                precision=precision)[0][:, :7]
        g32, g64 = grads["split"], grads["fp64"]
        bar = 1e-2 * g64.abs() + 1e-3 * g64.abs().amax(0)
        out[name] = {"max_rel_err": float(((g32 - g64).abs().amax(0) / g64.abs().amax(0)).max()),
                     "worst_fraction_of_float32_bar": float(((g32 - g64).abs() / bar).max())}
    return out


def bench_xhm(torch, dev, nb, n, steps, warmup, cpu=True):
    """IMRPhenomXHM, the multimode workload: nb binaries on linspace(20,
    2048, n) Hz, the five modes summed into h+, hx, in FP64 (complex128) and
    in the split arithmetic (complex64): each step (setup of the (2,2) record
    and the four higher-mode records, then the evaluation), each evaluation
    kernel alone with its HBM and VALU rooflines, the split path's mismatch
    against FP64; the XHM oracle on the host cores on a bounded sample."""
    from waveformx_amd import xhm

    freqs = np.linspace(20.0, 2048.0, n)
    params = torch.as_tensor(draw_xhm(nb), device=dev)
    f64 = torch.as_tensor(freqs, device=dev)
    out = (torch.empty((nb, n), dtype=torch.complex128, device=dev),
           torch.empty((nb, n), dtype=torch.complex128, device=dev))
    state = {}

    def step():
        state["r"] = xhm.setup(params, f64[:1])
        xhm.evaluate(*state["r"][:3], f64, out=out)

    ms = timed(torch, step, steps, warmup)
    records, head, modes_, _ = state["r"]
    ms_eval = timed(torch, lambda: xhm.evaluate(records, head, modes_, f64, out=out), steps, warmup)
    ms_setup = timed(torch, lambda: xhm.setup(params, f64[:1]), steps, warmup)
    bytes_per_launch = 32.0 * nb * n
    achieved = bytes_per_launch / (ms_eval * 1e-3) / 1e9
    result = {"xhm_value": nb / (ms * 1e-3), "xhm_unit": "waveforms/s", "xhm_ms_per_step": ms,
              "xhm_eval_ms": ms_eval, "xhm_setup_ms": ms_setup,
              "xhm_config": {"workload": f"IMRPhenomXHM {nb} binaries x {n} bins linspace(20, 2048) Hz, FP64, "
                                         "modes (2,2), (2,1), (3,3), (3,2), (4,4)", "binaries": nb, "bins": n},
              "xhm_roofline_hbm": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": achieved / HBM_PEAK_GBS}}
    # the split arithmetic (FP32, each mode's leading inspiral phase in FP64)
    f32 = torch.as_tensor(freqs, dtype=torch.float32, device=dev)
    out32 = (torch.empty((nb, n), dtype=torch.complex64, device=dev),
             torch.empty((nb, n), dtype=torch.complex64, device=dev))

    def step_split():
        state["s"] = xhm.setup(params, f32[:1], precision="split")
        xhm.evaluate(*state["s"][:3], f32, out=out32)

    ms_split = timed(torch, step_split, steps, warmup)
    srec = state["s"]
    ms_split_eval = timed(torch, lambda: xhm.evaluate(*srec[:3], f32, out=out32), steps, warmup)
    # its accuracy: noise-weighted mismatch against the FP64 path on the same float32 grid, every binary
    f64r = f32.to(torch.float64)
    xhm.evaluate(*xhm.setup(params, f64r[:1])[:3], f64r, out=out)
    w = 1.0 / torch.as_tensor(((freqs / 100.0) ** -4 + 1.0 + (freqs / 100.0) ** 2), device=dev)
    worst = 0.0
    for a32, a64 in zip(out32, out):
        x = a32.to(torch.complex128)
        ab = (x * a64.conj()).real @ w
        worst = max(worst, float((1.0 - ab / torch.sqrt(((x.abs() ** 2) @ w) * ((a64.abs() ** 2) @ w))).max()))
    del out32
    # and per mode, on the first 512 binaries (rows where both modes vanish skipped)
    rows = min(nb, 512)
    m32 = xhm.evaluate(*(t[:rows] for t in srec[:3]), f32, with_modes=True)[2]
    m64 = xhm.evaluate(*(t[:rows] for t in xhm.setup(params[:rows], f64r[:1])[:3]), f64r, with_modes=True)[2]
    per_mode = []
    for a32, a64 in zip(m32, m64):
        x = a32.to(torch.complex128)
        n32, n64 = (x.abs() ** 2) @ w, (a64.abs() ** 2) @ w
        keep = (n32 > 0) & (n64 > 0)
        mm = 1.0 - ((x * a64.conj()).real @ w)[keep] / torch.sqrt(n32[keep] * n64[keep])
        per_mode.append(float(mm.max()) if bool(keep.any()) else 0.0)
    del m32, m64
    result.update({"xhm_split_value": nb / (ms_split * 1e-3), "xhm_split_ms_per_step": ms_split,
                   "xhm_split_eval_ms": ms_split_eval, "xhm_split_over_fp64": ms / ms_split,
                   "xhm_split_max_weighted_mismatch_vs_fp64": worst,
                   "xhm_split_mode_mismatch_vs_fp64": dict(zip(("22", "21", "33", "32", "44"), per_mode)),
                   "xhm_split_roofline_hbm": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                                              "achieved": 16.0 * nb * n / (ms_split_eval * 1e-3) / 1e9,
                                              "frac": 16.0 * nb * n / (ms_split_eval * 1e-3) / 1e9 / HBM_PEAK_GBS}})
    per_split = XHM_SPLIT_SLOTS_PER_BIN.get((nb, n))
    if per_split is not None:
        # algorithmic work (the formula's FP32 issue slots) per second against the VALU issue rate; the
        # instructions the kernel issues beside it
        tops = per_split * nb * n / (ms_split_eval * 1e-3) / 1e12
        issued = XHM_SPLIT_VALU_PER_BIN[(nb, n)] * nb * n / (ms_split_eval * 1e-3) / 1e12
        result["xhm_split_roofline"] = {"bound": "valu", "achieved": tops, "peak": VALU_PEAK_TOPS,
                                        "unit": "T op slots/s", "frac": tops / VALU_PEAK_TOPS,
                                        "per_bin": per_split, "issued_per_bin": XHM_SPLIT_VALU_PER_BIN[(nb, n)],
                                        "issue_frac": issued / VALU_PEAK_TOPS}
    # the gradient: the fused split loss and gradient (adjoint.loss_and_grad: the dual-number set-up, the
    # residuals, each mode's reverse-mode adjoint) of L = sum_f w(f) (|h+ - d+|^2 + |hx - dx|^2), d at
    # parameters perturbed by 1%, every binary; against the FP64 path's exact gradient of the same loss
    # (xhm.vjp, the same adjoint in FP64) on the same float32 grid; the FP64 fused loss and gradient timed too
    from waveformx_amd import adjoint

    rng = np.random.default_rng(7)
    perturbed = params.cpu().numpy() * (1.0 + 1e-2 * rng.normal(size=tuple(params.shape)))
    perturbed[:, 7] = params[:, 7].cpu().numpy()
    d_plus, d_cross = xhm.polarizations(torch.as_tensor(perturbed, device=dev), f64r)
    d_plus32, d_cross32 = d_plus.to(torch.complex64), d_cross.to(torch.complex64)
    w32 = w.to(torch.float32)

    def fwd_grad():
        state["g"] = adjoint.loss_and_grad(params, f32, d_plus32, d_cross32, w32, precision="split",
                                           approximant="IMRPhenomXHM")

    ms_grad = timed(torch, fwd_grad, steps, warmup)
    loss32, g32 = state["g"]
    w64 = w32.to(torch.float64)

    def fwd_grad64():
        state["g64"] = adjoint.loss_and_grad(params, f64r, d_plus, d_cross, w64, precision="fp64",
                                             approximant="IMRPhenomXHM")

    ms_grad64 = timed(torch, fwd_grad64, max(2, steps // 4), 1)
    del state["g64"]
    hp, hc = xhm.polarizations(params, f64r)
    g64 = xhm.vjp(params, f64r, 2.0 * w64 * (hp - d_plus), 2.0 * w64 * (hc - d_cross))
    loss64 = ((hp - d_plus).abs() ** 2 + (hc - d_cross).abs() ** 2) @ w64
    del hp, hc
    g_err = ((g32 - g64).abs().amax(0) / g64.abs().amax(0))[:7]
    bar = 1e-2 * g64.abs() + 1e-3 * g64.abs().amax(0)
    result.update({"xhm_fwd_grad_binaries": nb, "xhm_fwd_grad_ms": ms_grad,
                   "xhm_fwd_grad_value": nb / (ms_grad * 1e-3), "xhm_fwd_grad_over_split_fwd": ms_grad / ms_split,
                   "xhm_fp64_fwd_grad_ms": ms_grad64, "xhm_fp64_fwd_grad_over_fp64_fwd": ms_grad64 / ms,
                   "xhm_grad_finite": bool(torch.isfinite(g32[:, :7]).all()),
                   "xhm_grad_rel_err_split_vs_fp64": dict(zip(
                       ("mass_1", "mass_2", "chi_1", "chi_2", "luminosity_distance", "inclination", "phi_ref"),
                       map(float, g_err.cpu()))),
                   "xhm_grad_worst_fraction_of_float32_bar": float(((g32 - g64).abs() / bar)[:, :7].max()),
                   "xhm_loss_rel_err_split_vs_fp64": float(((loss32 - loss64).abs() / loss64).max())})
    del d_plus, d_cross
    result["xhm_grad_split_vs_fp64_per_mode"] = xhm_mode_gradients(torch, xhm, params, perturbed, f32, f64r, w32,
                                                                   w64, min(nb, 512))
    per_bin = XHM_FP64_OPS_PER_BIN.get((nb, n))
    if per_bin is not None:
        # the formula's FP64 arithmetic (add, multiply, divide, compare; its elementary functions apart) per
        # second against the FP64 VALU rate (the non-packed FP32 rate); the instructions issued beside it
        tops = per_bin * nb * n / (ms_eval * 1e-3) / 1e12
        issued = XHM_VALU_PER_BIN[(nb, n)] * nb * n / (ms_eval * 1e-3) / 1e12
        result["xhm_roofline"] = {"bound": "valu", "achieved": tops, "peak": VALU_PEAK_TOPS,
                                  "unit": "T FP64 ops/s", "frac": tops / VALU_PEAK_TOPS, "per_bin": per_bin,
                                  "calls_per_bin": XHM_FP64_CALLS_PER_BIN[(nb, n)],
                                  "issued_per_bin": XHM_VALU_PER_BIN[(nb, n)], "issue_frac": issued / VALU_PEAK_TOPS}
    del out, state
    if cpu:
        from oracle import xhm as oxhm

        sample = 8192
        cores = min(16, os.cpu_count() or 1)
        p = draw_xhm(sample, seed=5)
        oxhm.polarizations(p[:4], freqs, threads=cores)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            oxhm.polarizations(p, freqs, threads=cores)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        result["xhm_cpu_baseline"] = {"value": sample / t, "unit": "waveforms/s", "cores": cores, "kind": "port",
                                      "sample": f"{sample} binaries x {n} bins, the FP64 C++ XHM oracle, OpenMP "
                                                f"over binaries, median of 3 ({t:.2f} s each)"}
    return result


def valu_roofline(nb, n, ms_eval, group):
    """The evaluation's VALU instructions per second against the issue rate,
    for the workload and shape whose instruction count was measured."""
    per_bin = EVAL_VALU_PER_BIN.get((nb, n)) if group == 0 else None
    if per_bin is None:
        return None
    achieved = per_bin * nb * n / (ms_eval * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "T instructions/s",
            "frac": achieved / VALU_PEAK_TOPS, "per_bin": per_bin}


def host_cpu():
    """The host's CPU model and logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(n_binaries, freqs, cores=None, repeats=7):
    """The FP64 oracle on the host cores (16, the box's share, unless
    ``cores``) on ``n_binaries`` of the workload: median of ``repeats``
    timed runs after a warm-up."""
    from oracle import oracle

    cores = cores or min(16, os.cpu_count() or 1)
    params = draw(n_binaries, seed=2)
    oracle.polarizations(params[:8], freqs, threads=cores)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        oracle.polarizations(params, freqs, threads=cores)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": n_binaries / t, "unit": "waveforms/s", "cores": cores, "kind": "port",
            "sample": f"{n_binaries} binaries x {len(freqs)} bins, FP64 C++ oracle, OpenMP over binaries, "
                      f"median of {repeats} ({t:.2f} s each)", "host": host_cpu()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--binaries", type=int, default=65536)
    ap.add_argument("--bins", type=int, default=4017)
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--cpu-binaries", type=int, default=32768)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grad-binaries", type=int, default=8192,
                    help="binaries of the fused forward + adjoint measurement (0: skip)")
    ap.add_argument("--ragged-binaries", type=int, default=16384,
                    help="binaries of the ragged-grid measurement (0: skip)")
    ap.add_argument("--xhm-binaries", type=int, default=4096,
                    help="binaries of the IMRPhenomXHM leg (0: skip it)")
    ap.add_argument("--xhm-bins", type=int, default=8193)
    ap.add_argument("--only", choices=["split", "split-df", "fp64", "fp32", "grad", "ragged", "xhm"],
                    help="run just this arithmetic's step (setup + evaluation) K times, for a profiler")
    ap.add_argument("--sweep", action="store_true", help="time the evaluation's launch shapes and exit")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run a plain device copy of this many MiB K times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the binaries are independent, so the path shards by binary with no "
              "exchange; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from waveformx_amd import xas

    dev = torch.device("cuda", 0)
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        ms = timed(torch, lambda: dst.copy_(src), args.steps, args.warmup)
        print(json.dumps({"copy_mib": args.copy_mib, "ms": ms, "GB/s": 2 * src.numel() * 4 / ms / 1e6}), flush=True)
        ms = timed(torch, lambda: dst.fill_(1.5), args.steps, args.warmup)
        print(json.dumps({"fill_mib": args.copy_mib, "ms": ms, "GB/s": src.numel() * 4 / ms / 1e6}), flush=True)
        return
    if args.only == "grad":
        print(json.dumps(bench_grad(torch, xas, dev, args.grad_binaries, args.bins, args.steps, args.warmup)),
              flush=True)
        return
    if args.only == "xhm":
        print(json.dumps(bench_xhm(torch, dev, args.xhm_binaries, args.xhm_bins, args.steps, args.warmup,
                                   cpu=False)), flush=True)
        return
    if args.only == "ragged":
        # the ragged workload's graph-replayed plan alone (tuned shape)
        from waveformx_amd.ragged import RaggedPlan

        p, grids = ragged_grids(args.ragged_binaries)
        params = torch.as_tensor(p, device=dev)
        plan = RaggedPlan(grids, precision="split")
        ms = timed(torch, lambda: plan(params), args.steps, args.warmup)
        print(json.dumps({"only": "ragged", "ms_per_step": ms, "bins": plan.total,
                          "shape": [plan.bins_per_thread, plan.blocks]}), flush=True)
        return
    nb, n = args.binaries, args.bins
    freqs64 = np.linspace(20.0, 1024.0, n)
    params = torch.as_tensor(draw(nb), device=dev)
    f32 = torch.as_tensor(freqs64, dtype=torch.float32, device=dev)
    f64 = torch.as_tensor(freqs64, dtype=torch.float64, device=dev)
    if args.sweep:
        # launch shapes of the FP32-split evaluation: binaries per block x
        # frequencies per thread
        records, _ = xas.setup(params, f32[:1], 1e21, None, "split")
        out = (torch.empty((nb, n), dtype=torch.complex64, device=dev),
               torch.empty((nb, n), dtype=torch.complex64, device=dev))
        for group in (2, 3, 4, 6, 8, 16):
            for fpt in (2, 4, 8):
                ms = timed(torch, lambda: xas.evaluate(records, f32, "split", group, out=out, freqs_per_thread=fpt),
                           args.steps, args.warmup)
                print(json.dumps({"group": group, "freqs_per_thread": fpt, "eval_ms": ms,
                                  "GB/s": 16.0 * nb * n / ms / 1e6}), flush=True)
        return
    if args.only:
        grid = f64 if args.only == "fp64" else f32
        ctype = torch.complex128 if args.only == "fp64" else torch.complex64
        out = (torch.empty((nb, n), dtype=ctype, device=dev), torch.empty((nb, n), dtype=ctype, device=dev))

        def one():
            records, _ = xas.setup(params, grid[:1], 1e21, None, args.only)
            xas.evaluate(records, grid, args.only, args.group, out=out)

        ms = timed(torch, one, args.steps, args.warmup)
        print(json.dumps({"only": args.only, "ms_per_step": ms, "waveforms/s": nb / ms * 1e3}), flush=True)
        return
    out32 = (torch.empty((nb, n), dtype=torch.complex64, device=dev),
             torch.empty((nb, n), dtype=torch.complex64, device=dev))
    out64 = (torch.empty((nb, n), dtype=torch.complex128, device=dev),
             torch.empty((nb, n), dtype=torch.complex128, device=dev))
    rec = {}

    def step(precision, freqs, out):
        def run():
            records, _ = xas.setup(params, freqs[:1], 1e21, None, precision)
            rec[precision] = records
            xas.evaluate(records, freqs, precision, args.group, out=out)
        return run

    ms_split = timed(torch, step("split", f32, out32), args.steps, args.warmup)
    ms_df = timed(torch, step("split-df", f32, out32), args.steps, args.warmup)
    ms_fp32 = timed(torch, step("fp32", f32, out32), args.steps, args.warmup)
    ms_fp64 = timed(torch, step("fp64", f64, out64), max(3, args.steps // 2), 1)
    # the dominant kernel alone (the evaluation), on the stream it runs on
    ms_eval = timed(torch, lambda: xas.evaluate(rec["split"], f32, "split", args.group, out=out32), args.steps,
                    args.warmup)
    ms_setup = timed(torch, lambda: xas.setup(params, f32[:1], 1e21, None, "split"), args.steps, args.warmup)
    # IMRPhenomD on the same kernels, workload and buffers
    from waveformx_amd import phenomd

    def step_pd(precision, freqs, out):
        def run():
            records, _ = phenomd.setup(params, freqs[:1], 1e21, None, precision)
            phenomd.evaluate(records, freqs, precision, args.group, out=out)
        return run

    ms_pd_split = timed(torch, step_pd("split", f32, out32), args.steps, args.warmup)
    ms_pd_fp64 = timed(torch, step_pd("fp64", f64, out64), max(3, args.steps // 2), 1)
    pd_records, _ = phenomd.setup(params, f32[:1], 1e21, None, "split")
    ms_pd_eval = timed(torch, lambda: phenomd.evaluate(pd_records, f32, "split", args.group, out=out32), args.steps,
                       args.warmup)
    del pd_records

    # accuracy of the split path against the FP64 path on the device (which
    # matches the FP64 oracle to 1e-12): on the FP32-rounded grid the split
    # path sees, and on the exact grid; then the plain FP32 path's on the
    # rounded grid
    w = 1.0 / torch.as_tensor(((freqs64 / 100.0) ** -4 + 1.0 + (freqs64 / 100.0) ** 2), device=dev)
    step("fp64", f32.to(torch.float64), out64)()
    step("split", f32, out32)()
    rounded = accuracy(torch, out32, out64, w)
    step("fp32", f32, out32)()
    rounded_fp32 = accuracy(torch, out32, out64, w)
    step("split", f32, out32)()
    step("fp64", f64, out64)()
    exact = accuracy(torch, out32, out64, w)
    torch.cuda.synchronize()

    bytes_per_launch = 16.0 * nb * n
    achieved = bytes_per_launch / (ms_eval * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": nb / (ms_split * 1e-3),
        "unit": "waveforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_split,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (leading phase term f64)",
        "data": "synthetic: seeded parameter draws",
        "config": {"workload": f"IMRPhenomXAS {nb} binaries x {n} bins linspace(20, 1024) Hz, FP32-split forward "
                               "(setup + evaluation)", "binaries": nb, "bins": n},
        "fp64_value": nb / (ms_fp64 * 1e-3),
        "fp64_ms_per_step": ms_fp64,
        "split_over_fp64": ms_fp64 / ms_split,
        "split_df_value": nb / (ms_df * 1e-3),
        "fp32_value": nb / (ms_fp32 * 1e-3),
        "fp32_accuracy_vs_fp64_rounded_grid": rounded_fp32,
        "eval_ms": ms_eval,
        "setup_ms": ms_setup,
        "phenomd_split_value": nb / (ms_pd_split * 1e-3),
        "phenomd_fp64_value": nb / (ms_pd_fp64 * 1e-3),
        "phenomd_eval_ms": ms_pd_eval,
        "phenomd_roofline_frac": 16.0 * nb * n / (ms_pd_eval * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "max_weighted_mismatch_vs_fp64": rounded["mismatch"],
        "accuracy_vs_fp64_rounded_grid": rounded,
        "accuracy_vs_fp64_exact_grid": exact,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": EVAL_TRAFFIC.get((nb, n)) if args.group == 0 else None},
        # the same launch against the VALU issue rate (co-limiting: the
        # arithmetic runs under the output stream)
        "roofline_valu": valu_roofline(nb, n, ms_eval, args.group),
    }
    del out32, out64, rec
    torch.cuda.empty_cache()
    if args.grad_binaries:
        result.update(bench_grad(torch, xas, dev, args.grad_binaries, n, args.steps, args.warmup))
        result.update(bench_grad(torch, phenomd, dev, args.grad_binaries, n, args.steps, args.warmup,
                                 approximant="IMRPhenomD", prefix="phenomd_"))
    if args.ragged_binaries:
        result.update(bench_ragged(torch, xas, dev, args.ragged_binaries, args.steps, args.warmup))
    if args.xhm_binaries:
        result.update(bench_xhm(torch, dev, args.xhm_binaries, args.xhm_bins, args.steps, args.warmup,
                                cpu=not args.no_cpu))
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_binaries, freqs64)
        result["cpu_baseline_1core"] = cpu_baseline(max(64, args.cpu_binaries // 32), freqs64, cores=1)
    # literature, not measured here: the reference's JAX float64 jit(vmap) on
    # a 24-CPU M2 Ultra, 4096 binaries x 1000 bins in 18.97 ms
    # (report/data/xas.json:4132), and that figure scaled linearly to 4017 bins
    result["reference_published_cpu"] = {"value": 4096 / 18.97e-3, "unit": "waveforms/s @1000 bins",
                                         "scaled_to_bins": 4096 / 18.97e-3 * 1000 / n,
                                         "hardware": "Apple M2 Ultra CPU (24), JAX float64 jit(vmap)",
                                         "kind": "literature"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
