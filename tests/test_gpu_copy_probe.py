"""The measurement copy (ecsg_copy_probe) copies exactly, with and without
the non-temporal hints, disjoint and in place; other hints are rejected."""
import numpy as np
import pytest
import torch

import ecsg_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hints", [0, 3])
def test_copy_probe_copies(hints):
    L = ecsg_amd.load_library()
    n = 1_000_003
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.full_like(src, -1.0)
    assert L.ecsg_copy_probe(dst.data_ptr(), src.data_ptr(), n - 1, hints, None) == 0
    torch.cuda.synchronize()
    d = dst.cpu().numpy()
    assert np.array_equal(d[:-1], np.arange(n - 1, dtype=np.float32)) and d[-1] == -1.0
    assert L.ecsg_copy_probe(dst.data_ptr(), dst.data_ptr(), n, hints, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), d)


def test_copy_probe_rejects():
    L = ecsg_amd.load_library()
    x = torch.zeros(4, device="cuda")
    assert L.ecsg_copy_probe(x.data_ptr(), x.data_ptr(), 4, 1, None) == -1
    assert L.ecsg_copy_probe(None, x.data_ptr(), 4, 0, None) == -1
