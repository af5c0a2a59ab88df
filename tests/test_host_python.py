"""Host-side Python logic that needs no GPU: error mapping of the C ABI and
the serving driver's latency / concurrency accounting."""
import numpy as np
import pytest

import csharp_amd as ca
from csharp_amd.serving import run_poisson


def test_jukes_cantor_code_raises_its_own_exception():
    with pytest.raises(ca.JukesCantorUndefined) as ei:
        ca._check(ca.SWG_E_JC_UNDEFINED)
    assert ei.value.code == ca.SWG_E_JC_UNDEFINED and isinstance(ei.value, ca.SwgError)
    with pytest.raises(ca.SwgError) as ei:
        ca._check(ca.SWG_E_UNSUPPORTED)
    assert not isinstance(ei.value, ca.JukesCantorUndefined)


class FakeServer:
    """Admits each query `wait_ms` after submission and finishes it
    `run_ms` later (server clock = submission index * 10 ms)."""

    def __init__(self, wait_ms, run_ms):
        self.wait_ms, self.run_ms = wait_ms, run_ms
        self.t = {}
        self.cells = 0.0

    def submit(self, q):
        k = len(self.t)
        self.t[k] = (10.0 * k, 10.0 * k + self.wait_ms, 10.0 * k + self.wait_ms + self.run_ms)
        self.cells += 1e9 * len(q)
        return k

    def wait(self, t, timeout_ms):
        return True

    def times(self, t):
        return self.t[t]

    def release(self, t):
        pass

    def stats(self):
        return len(self.t), self.cells, 0.0


def test_serving_accounting():
    r = run_poisson(FakeServer(5.0, 25.0), [b"A"] * 10, rate=1e6, seed=1)
    # latency = wait + run for every query
    assert r["p50_ms"] == 30.0 and r["mean_ms"] == 30.0 and r["queue_wait_p50_ms"] == 5.0
    # in flight from admission to completion: 25 ms each, a new one every 10 ms
    span_ms = 90.0 + 30.0
    assert r["admitted_concurrency_mean"] == round(10 * 25.0 / span_ms, 2)
    assert r["admitted_concurrency_max"] == 3
    assert r["gcups"] == round(10 * 1e9 / (span_ms / 1e3) / 1e9, 1)


def test_out_array_validation():
    # compute_block(out=...) takes the row pitch from the array itself and
    # refuses arrays that would place or overrun the block wrongly
    from csharp_amd import _out_ld
    full = np.zeros((50, 80), dtype=np.int16)
    assert _out_ld(full, np.int16, 50, 80) == 80
    assert _out_ld(full[10:20, 30:70], np.int16, 10, 40) == 80  # a window of a larger matrix
    with pytest.raises(ValueError):
        _out_ld(full, np.float64, 10, 10)          # wrong TDistance
    with pytest.raises(ValueError):
        _out_ld(full[:5, :5], np.int16, 10, 10)    # too small
    with pytest.raises(ValueError):
        _out_ld(full[:, ::2], np.int16, 10, 10)    # no unit column stride
    with pytest.raises(ValueError):
        _out_ld(np.zeros(100, np.int16), np.int16, 10, 10)  # not 2-D
    ro = np.zeros((10, 10), np.int16)
    ro.flags.writeable = False
    with pytest.raises(ValueError):
        _out_ld(ro, np.int16, 10, 10)
