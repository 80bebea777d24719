"""The C ABI's concurrency contract (include/waveformx_amd.h): entry points
called from two host threads at once, each on its own non-default HIP
stream, give the bits of sequential calls, and wfx_last_error() is per host
thread (ctypes releases the GIL around each call, so the calls overlap)."""

import threading

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def test_two_threads_on_two_streams_match_sequential_calls(cuda):
    import torch

    from waveformx_amd import adjoint, xas

    f32 = torch.as_tensor(np.linspace(20.0, 1024.0, 4017), dtype=torch.float32, device=cuda)
    f64 = f32.to(torch.float64)
    batches = [torch.as_tensor(oracle.draw(2048, seed=s), device=cuda) for s in (11, 12)]
    weights = torch.ones(4017, dtype=torch.float32, device=cuda)

    def work(params):
        out = []
        for precision, f in (("split", f32), ("fp64", f64)):
            records, _ = xas.setup(params, f[:1], 1e21, None, precision)
            hp, hc = xas.evaluate(records, f, precision)
            out += [hp, hc]
        data = out[0] * 0.99
        out += list(adjoint.loss_and_grad(params, f32, data, out[1], weights, precision="split"))
        return out

    sequential = [work(p) for p in batches]
    torch.cuda.synchronize()
    results = [None, None]
    errors = []
    start = threading.Barrier(2)

    def run(index):
        try:
            stream = torch.cuda.Stream(device=cuda)
            with torch.cuda.stream(stream):
                start.wait()
                for _ in range(5):
                    results[index] = work(batches[index])
            stream.synchronize()
        except Exception as error:  # noqa: BLE001 - reported below
            errors.append(error)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert all(not t.is_alive() for t in threads)
    torch.cuda.synchronize()
    for got, want in zip(results, sequential):
        for a, b in zip(got, want):
            assert torch.equal(a, b)


def test_last_error_is_per_thread(cuda):
    from waveformx_amd import _lib

    lib = _lib.lib()
    step = threading.Barrier(2)
    seen = {}

    def first():
        assert lib.wfx_xas_record_bytes(7) == -1  # unknown precision
        step.wait()
        step.wait()  # the other thread has failed in between
        seen["first"] = lib.wfx_last_error().decode()

    def second():
        step.wait()
        assert lib.wfx_xas_setup(None, 4, None, 0, 1.0, float("nan"), 0, None, None, None) == -1
        seen["second"] = lib.wfx_last_error().decode()
        step.wait()

    threads = [threading.Thread(target=first), threading.Thread(target=second)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=30)
    assert "precision" in seen["first"] and "null" not in seen["first"], seen
    assert "null pointer" in seen["second"], seen
