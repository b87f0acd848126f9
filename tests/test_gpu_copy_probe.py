"""The measurement copy (ecsg_copy_probe) copies exactly, with and without
the non-temporal hints, disjoint and in place; other hints are rejected.
The streaming probe (ecsg_stream_probe) updates exactly the entries it
documents, in the stencil tiling and as a flat pass."""
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


def _layout(nx, ny, nz):
    """The documented device layout of a radius-1 FP32 grid."""
    line = 32
    P = (32 + nx + 2 + line - 1) // line * line
    S2 = P * (ny + 3)
    return P, S2, S2 + P + 32, (S2 * (nz + 3) + line - 1) // line * line


@pytest.mark.parametrize("pattern", [0, 1])
@pytest.mark.parametrize("by,zc", [(4, 2), (1, 1), (8, 8), (2, 0)])
def test_stream_probe_updates(by, zc, pattern):
    L = ecsg_amd.load_library()
    nx, ny, nz = 70, 9, 3                      # two tiles in x, a partial one in y
    F = L.ecsg_stream_probe_floats(nx, ny, nz)
    P, S2, o, astride = _layout(nx, ny, nz)
    assert F == 18 * astride
    rng = np.random.default_rng(by * 10 + zc)
    a = rng.standard_normal(F).astype(np.float32)
    buf = torch.from_numpy(a).cuda()
    assert L.ecsg_stream_probe(buf.data_ptr(), nx, ny, nz, by, zc, pattern, None) == 0
    torch.cuda.synchronize()
    got = buf.cpu().numpy().reshape(18, astride)
    ref = a.reshape(18, astride).copy()
    if zc == 0:
        idx = np.arange(S2, S2 + (nz + 1) * S2)
    else:
        iz, iy, ix = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
        idx = (o + iz * S2 + iy * P + ix).ravel()
    nr, nw = (18, 6) if pattern == 0 else (10, 3)
    for k in range(nw):
        u = ref[k, idx].copy()
        for j in range(nw + k, nr, nw):
            u = u + ref[j, idx]
        ref[k, idx] = u
    assert np.array_equal(got, ref)


def test_stream_probe_rejects():
    L = ecsg_amd.load_library()
    x = torch.zeros(L.ecsg_stream_probe_floats(2, 2, 2), device="cuda")
    assert L.ecsg_stream_probe(x.data_ptr(), 2, 2, 2, 3, 1, 0, None) == -1
    assert L.ecsg_stream_probe(x.data_ptr(), 2, 2, 2, 4, -1, 0, None) == -1
    assert L.ecsg_stream_probe(x.data_ptr(), 0, 2, 2, 4, 1, 0, None) == -1
    assert L.ecsg_stream_probe(x.data_ptr(), 2, 2, 2, 4, 1, 2, None) == -1
    assert L.ecsg_stream_probe(None, 2, 2, 2, 4, 1, 0, None) == -1
