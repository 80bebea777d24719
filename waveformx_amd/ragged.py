"""Ragged batches: binaries each on its own uniform grid (own f_lower, df and
length), set up once and replayed as one HIP graph.

:class:`RaggedPlan` holds the static part of a ragged batch: the offsets of
each binary's bins, the prefix sum of work tiles, the output buffers, the
tile and persistent-grid shape chosen by a small autotuner (timed on the
device, cached on disk per device, precision and batch shape), and a HIP
graph of the setup and evaluation kernels. Calling it with new parameters
copies them into the graph's input buffer and replays the graph: one
persistent launch over all binaries, no host synchronisation.

The reference batches only binaries on one shared grid
(``src/waveformx/torch.py:84-85``); this is the extension the north star
names (ragged grids in one persistent, graph-captured launch).
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import xas

__all__ = ["RaggedPlan", "uniform_grid_length"]

BINS_PER_THREAD = (4, 8, 16)
# a persistent grid of 128-192 blocks per CU leaves the shortest tail of
# tiles on the 16,384-binary workload (~5% faster than 64 per CU)
BLOCKS_PER_CU = (16, 64, 128, 192)


def uniform_grid_length(f_min: float, f_max: float, delta_f: float) -> int:
    """The number of frequencies of ``FrequencyGrid.uniform(f_min, f_max,
    delta_f)`` (``src/waveformx/_results.py:53-57``)."""
    return int(np.floor((f_max - f_min) / delta_f + 0.5)) + 1


def check_descriptors(f0: np.ndarray, df: np.ndarray, lengths: np.ndarray) -> None:
    """Reject uniform-grid descriptors ``(f0, df, n)`` that no grid has:
    ``f0 > 0`` and ``df > 0`` (NaN fails both) and ``n >= 0``."""
    if not (np.all(f0 > 0) and np.all(df > 0) and np.all(lengths >= 0)):
        raise ValueError("grids need f0 > 0, df > 0 and n >= 0")


def _cache_path() -> Path:
    root = os.environ.get("WFX_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "waveformx_amd")
    return Path(root) / "ragged_tuning.json"


def _load_cache() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(cache: dict) -> None:
    try:
        path = _cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_name(f"{path.name}.{os.getpid()}.tmp")  # one per writer: concurrent tuners
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: tune again next time


class RaggedPlan:
    """A ragged batch of ``len(grids)`` binaries; ``grids[b] = (f0, df, n)``.

    ``plan(params)`` with ``params`` a (B, 8) float64 device tensor returns
    flat ``(hplus, hcross)`` (views of the plan's buffers, overwritten by
    the next call); ``plan.offsets`` (B+1,) locates each binary's bins.
    """

    def __init__(self, grids: Sequence[Tuple[float, float, int]], precision: str = "split",
                 amp_scale: float = 1e21, reference_mf: Optional[float] = None,
                 device: Optional[torch.device] = None, tune: bool = True, graph: bool = True,
                 bins_per_thread: Optional[int] = None, blocks: Optional[int] = None):
        from ._api import compute_device

        self.device = compute_device(device)
        self.precision = precision
        self.amp_scale = amp_scale
        self.reference_mf = reference_mf
        f0 = np.array([float(g[0]) for g in grids])
        df = np.array([float(g[1]) for g in grids])
        lengths = np.array([int(g[2]) for g in grids], dtype=np.int64)
        check_descriptors(f0, df, lengths)
        self.n_binaries = len(grids)
        offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        self.total = int(offsets[-1])
        dev = self.device
        self.offsets = torch.as_tensor(offsets, device=dev)
        self.grid_f0 = torch.as_tensor(f0, device=dev)
        self.grid_df = torch.as_tensor(df, device=dev)
        # the setup reads each binary's first frequency in the working precision
        self.f_min = self.grid_f0 if precision != "fp32" else self.grid_f0.to(torch.float32).to(torch.float64)
        ctype = torch.complex128 if precision == "fp64" else torch.complex64
        self.hplus = torch.empty((self.total,), dtype=ctype, device=dev)
        self.hcross = torch.empty((self.total,), dtype=ctype, device=dev)
        self.params = torch.zeros((self.n_binaries, 8), dtype=torch.float64, device=dev)
        self.params[:, 0] = 30.0
        self.params[:, 1] = 20.0
        self.params[:, 4] = 400.0
        self._tiles = {}
        if bins_per_thread is None or blocks is None:
            bpt, blk = self._tuned(lengths) if tune else (1, 0)
            bins_per_thread = bins_per_thread or bpt
            blocks = blocks if blocks is not None else blk
        self.bins_per_thread = int(bins_per_thread)
        self.blocks = int(blocks)
        self.graph = None
        if graph:
            self._capture()

    # ------------------------------------------------------------ kernels

    def tiles(self, bins_per_thread: int) -> torch.Tensor:
        if bins_per_thread not in self._tiles:
            self._tiles[bins_per_thread] = xas.tile_prefix(self.offsets, bins_per_thread)
        return self._tiles[bins_per_thread]

    def _run(self, bins_per_thread: int, blocks: int) -> None:
        records, _ = xas.setup(self.params, self.f_min, self.amp_scale, self.reference_mf, self.precision)
        xas.evaluate_ragged(records, self.offsets, grid_f0=self.grid_f0, grid_df=self.grid_df,
                            precision=self.precision, tiles=self.tiles(bins_per_thread), blocks=blocks,
                            total=self.total, bins_per_thread=bins_per_thread, out=(self.hplus, self.hcross))

    def _capture(self) -> None:
        self.tiles(self.bins_per_thread)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._run(self.bins_per_thread, self.blocks)  # warm-up outside the capture
        torch.cuda.current_stream(self.device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._run(self.bins_per_thread, self.blocks)

    def __call__(self, params: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        if params.shape != (self.n_binaries, 8):
            raise ValueError(f"params must be ({self.n_binaries}, 8)")
        self.params.copy_(params)
        if self.graph is not None:
            self.graph.replay()
        else:
            self._run(self.bins_per_thread, self.blocks)
        return self.hplus, self.hcross

    # ------------------------------------------------------------- tuning

    def _key(self, lengths: np.ndarray) -> str:
        name = torch.cuda.get_device_name(self.device)
        hist = np.histogram(np.log2(np.maximum(lengths, 1)), bins=8, range=(0, 24))[0]
        shapes = f"{'-'.join(map(str, BINS_PER_THREAD))}x{'-'.join(map(str, BLOCKS_PER_CU))}"
        return f"{name}|{shapes}|{self.precision}|{self.n_binaries}|{self.total}|{','.join(map(str, hist))}"

    def _tuned(self, lengths: np.ndarray) -> Tuple[int, int]:
        cache = _load_cache()
        key = self._key(lengths)
        if key in cache:
            return tuple(cache[key]["choice"])
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        timings = {}
        for bpt in BINS_PER_THREAD:
            for per_cu in BLOCKS_PER_CU:
                blocks = per_cu * cus
                self._run(bpt, blocks)
                torch.cuda.synchronize(self.device)
                start.record()
                for _ in range(3):
                    self._run(bpt, blocks)
                stop.record()
                torch.cuda.synchronize(self.device)
                timings[(bpt, blocks)] = start.elapsed_time(stop) / 3
        best = min(timings, key=timings.get)
        cache[key] = {"choice": list(best), "ms": {f"{b}x{k}": v for (b, k), v in timings.items()}}
        _store_cache(cache)
        return best
