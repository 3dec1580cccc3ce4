"""Batched multi-instance sweep on the GPU.

* the split-BF16 MFMA GEMM against the FP64 product W_in . X;
* the plain per-instance FP64 networks (CSR everywhere) against the CPU
  oracle's instance loop (orc_sweep_instance_run), 1e-9 relative;
* the batched FP32/BF16 form against the same oracle: max relative rate error
  <= 1e-4 after 300 steps.
"""

import math

import numpy as np
import pytest

from annarchy_amd.batched import RateSweep
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _sweep(seed=0):
    sizes = [150, 900, 40, 500, 1200, 75, 333]
    return RateSweep(sizes, n_in=100, p=0.05, seed=seed)


def test_gemm_split_bf16(gpu):
    sw = _sweep()
    b = sw.batched(lanes=16)
    b.run(2, graph_batch=2)
    y = b.input_sums()
    x0 = 1.0 + 0.5 * np.sin(sw.omega * 0.0 + sw.phase)  # X of step 0, used by step 1's GEMM
    ref = sw.w_in @ x0
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err < 2e-5, err


def _oracle(sw, i, steps):
    r, _ = orc.sweep_instance_run(sw.recurrent[i], sw.w_in[: min(sw.n_in, int(sw.sizes[i]))], sw.phase[:, i],
                                  sw.omega, float(sw.gains[i]), sw.tau, sw.dt, sw.r0[i], steps)
    return r


def test_plain_instances_match_oracle(gpu):
    sw = _sweep(1)
    steps = 150
    nets = sw.plain_networks()
    for i, (net, exc) in enumerate(nets):
        net.run(steps)
        ref = _oracle(sw, i, steps)
        got = exc.r.reshape(-1)
        assert np.max(np.abs(got - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref))), i


@pytest.mark.parametrize("mode,lanes", [(0, 8), (0, 32), (1, 8), (1, 32), (2, 16)])
def test_batched_matches_oracle(gpu, lanes, mode):
    sw = _sweep(2)
    steps = 300
    b = sw.batched(lanes=lanes)
    b.set_mode(mode)
    b.run(steps, graph_batch=20)
    got = b.rates()
    worst = 0.0
    for i in range(sw.n_inst):
        ref = _oracle(sw, i, steps)
        worst = max(worst, np.max(np.abs(got[i] - ref)) / max(1e-3, np.max(np.abs(ref))))
    assert worst <= 1e-4, worst


def test_sliced_mode_is_the_default_and_matches_lds_mode(gpu):
    # mode 2 (sliced chunks) is chosen when every instance fits; it sums each
    # row in one lane, so it agrees with the LDS kernel to FP32 rounding
    sw = _sweep(4)
    a = sw.batched(16)
    assert a.info()["mode"] == 2
    a.run(40, graph_batch=8)
    b = sw.batched(16)
    b.set_mode(1)
    b.run(40, graph_batch=8)
    for x, y in zip(a.rates(), b.rates()):
        assert np.max(np.abs(x - y)) <= 1e-5 * max(1e-3, np.max(np.abs(y)))


def test_batched_graph_remainder_parity(gpu):
    # odd step counts and graph/remainder mixes give the same state as one long run
    sw = _sweep(3)
    a = sw.batched(16)
    a.run(37, graph_batch=10)
    a.run(13, graph_batch=10)
    b = sw.batched(16)
    b.run(50, graph_batch=50)
    for x, y in zip(a.rates(), b.rates()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("sizes,n_in,steps", [([150, 900, 40, 500, 1200, 75, 333], 100, 5),
                                               ([300] * 40, 200, 9),
                                               ([90, 700], 48, 3)])
def test_chunked_gemm_last_step(gpu, sizes, n_in, steps):
    # one eager chunk of `steps` steps: the input psp of its last step is
    # W_in . X(t = steps - 2); shapes with partial tiles along every GEMM
    # dimension (instances not a multiple of 32, input neurons not a
    # multiple of 16 or 128, an odd number of row tiles)
    sw = RateSweep(sizes, n_in=n_in, p=0.05, seed=6)
    b = sw.batched(lanes=16)
    b.run(steps, graph_batch=steps + 1)
    assert b.info()["chunk_steps"] >= steps
    y = b.input_sums()
    x = 1.0 + 0.5 * np.sin(sw.omega * (steps - 2) * sw.dt + sw.phase)
    ref = sw.w_in @ x
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err < 2e-5, err
