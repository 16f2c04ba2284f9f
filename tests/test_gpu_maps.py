"""GPU parity of the beta <-> DFT maps through the C ABI.

Even shapes (the reference's unit-test shapes): bit-exact (IEEE ==) against
the golden fixtures and the CPU oracle.  Odd shapes (extension): bit-exact
against the oracle, which the CPU tests pin by invariants.
"""
import os

import numpy as np
import pytest
import torch

from oracle import maps

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(f for f in os.listdir(GOLD) if f.startswith("maps_"))


def _op(dims):
    from csipm.fft_operator import FFTOperator
    return FFTOperator(int(np.prod(dims)), len(dims), dims, [], rdft=True)


@pytest.mark.parametrize("fname", FIXTURES)
def test_maps_golden(cuda, fname):
    from csipm.fft_operator import DFT_to_beta, beta_to_DFT
    d = np.load(os.path.join(GOLD, fname))
    dims = tuple(int(v) for v in d["dims"])
    print("maps", dims, flush=True)
    op = _op(dims)
    v = beta_to_DFT(op, torch.from_numpy(d["beta_in"]).to(cuda)).cpu().numpy()
    assert np.array_equal(v, d["v_half"])
    b = DFT_to_beta(op, torch.from_numpy(d["v_in"]).to(cuda)).cpu().numpy()
    assert np.array_equal(b, d["beta_out"])
    if "v_full" in d:
        vf = beta_to_DFT(op, torch.from_numpy(d["beta_in"]).to(cuda), full=True).cpu().numpy()
        assert np.array_equal(vf, d["v_full"])
        bf = DFT_to_beta(op, torch.from_numpy(vf).to(cuda), full=True).cpu().numpy()
        assert np.array_equal(bf, maps.dft_to_beta(vf, dims))


@pytest.mark.parametrize("dims", [(1024,), (4096,), (255, 255), (64, 63), (33, 30, 26),
                                  (40, 32, 24), (17, 16, 15)])
def test_maps_vs_oracle(cuda, dims):
    from csipm.fft_operator import DFT_to_beta, beta_to_DFT
    print("maps", dims, flush=True)
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    beta = rng.standard_normal(n)
    op = _op(dims)
    v = beta_to_DFT(op, torch.from_numpy(beta).to(cuda)).cpu().numpy()
    assert np.array_equal(v, maps.beta_to_dft(beta, dims, rdft=True))
    x = rng.standard_normal(n)
    vh = np.fft.rfftn(x.reshape(dims[::-1])).reshape(-1) / np.sqrt(n)
    b = DFT_to_beta(op, torch.from_numpy(vh).to(cuda)).cpu().numpy()
    assert np.array_equal(b, maps.dft_to_beta(vh, dims))
    # round trip on the device
    rt = DFT_to_beta(op, beta_to_DFT(op, torch.from_numpy(beta).to(cuda))).cpu().numpy()
    assert np.abs(rt - beta).max() <= 4.5e-16 * np.abs(beta).max()
