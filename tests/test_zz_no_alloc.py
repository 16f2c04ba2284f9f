"""csipm_mask_index_missing_noalloc, the allocation-free variant of
csipm_mask_index_missing, called through the C ABI and compared bit for bit
with csipm_mask_index_missing (and the oracle's index_missing): two calls in
a row, the second on a smaller mask, so the second call runs on the count
workspace the first one left behind; then a mask above the variant's limit,
which is rejected on the host before any device work.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import realdata as ord_

pytestmark = pytest.mark.gpu


def _index(name, w, cuda):
    from csipm import _lib
    n = int(w.numel())
    out = torch.full((max(n, 1),), -1, dtype=torch.int64, device=cuda)
    cnt = ctypes.c_int64(-1)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    _lib.call(name, ctypes.c_void_p(w.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
              ctypes.byref(cnt), ctypes.c_void_p(stream))
    return out[:cnt.value].cpu().numpy(), cnt.value


def test_noalloc_matches_allocating_entry_point(cuda):
    print("mask index no-alloc: 300001 then 4097 samples", flush=True)
    rng = np.random.default_rng(7)
    for n in (300001, 4097):
        w_np = rng.random(n) * (rng.random(n) > 0.4)
        w = torch.from_numpy(w_np).to(cuda)
        ref, cref = _index("csipm_mask_index_missing", w, cuda)
        got, cgot = _index("csipm_mask_index_missing_noalloc", w, cuda)
        assert cgot == cref == len(ord_.index_missing(w_np)), n
        assert np.array_equal(got, ref), n
        assert np.array_equal(got, ord_.index_missing(w_np)), n


def test_noalloc_rejects_oversize(cuda):
    from csipm import _lib
    print("mask index no-alloc: size limit", flush=True)
    limit = (65536 - 1) * 4096
    w = torch.zeros(1, dtype=torch.float64, device=cuda)
    cnt = ctypes.c_int64(-1)
    # n beyond the limit: rejected by the host-side size check, no launch
    rc = _lib.lib().csipm_mask_index_missing_noalloc(
        ctypes.c_void_p(w.data_ptr()), limit + 1, ctypes.c_void_p(w.data_ptr()),
        ctypes.byref(cnt), None)
    assert rc == _lib.ERR_ARG
    assert "exceeds its maximum" in _lib.lib().csipm_last_error().decode()
    assert cnt.value == 0
