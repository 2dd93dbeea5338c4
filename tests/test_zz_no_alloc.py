"""nan_nrc_infer_noalloc through the C ABI against nan_nrc_infer on the same
trained cache: the same kernels, so the same bits, on two calls in a row
(the second smaller, reading the scratch the first left behind), plus the
size limit, which is refused before anything is launched."""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg: str) -> None:
    print(f"[gpu-no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def _queries(rng, n, lo, hi, dev):
    pos = torch.from_numpy(rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)).to(dev)
    nrm = torch.nn.functional.normalize(torch.from_numpy(rng.standard_normal((n, 3)).astype(np.float32)), dim=1)
    alb = torch.from_numpy(rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32))
    return pos, nrm.to(dev).contiguous(), alb.to(dev)


def test_infer_without_allocation_equals_infer(device):
    from nan_amd import _native
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig

    dev = device.torch_device
    lo, hi = np.array([-1.0, 0.0, -1.0]), np.array([1.0, 2.0, 1.0])
    cache = RadianceCache(device, lo, hi, RadianceCacheConfig())
    rng = np.random.default_rng(5)
    # one training step so the table and weights are not the initial ones
    pos, nrm, alb = _queries(rng, 8192, lo, hi, dev)
    cache.train_step(pos, nrm, alb, torch.rand((8192, 3), device=dev))
    lib = _native.lib()
    stream = device.stream_handle
    for n in (50_000, 1_234):
        _say(f"{n} queries: nan_nrc_infer vs nan_nrc_infer_noalloc")
        pos, nrm, alb = _queries(rng, n, lo, hi, dev)
        want = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        got = torch.full((n, 3), -1.0, dtype=torch.float32, device=dev)
        _native.check(lib.nan_nrc_infer(cache.handle, pos.data_ptr(), nrm.data_ptr(), alb.data_ptr(),
                                        want.data_ptr(), n, stream), "nan_nrc_infer")
        _native.check(lib.nan_nrc_infer_noalloc(cache.handle, pos.data_ptr(), nrm.data_ptr(), alb.data_ptr(),
                                                got.data_ptr(), n, stream), "nan_nrc_infer_noalloc")
        torch.cuda.synchronize()
        w = want.cpu().numpy()
        g = got.cpu().numpy()
        assert (w > 0).mean() > 0.2  # the trained cache predicts something
        np.testing.assert_array_equal(g.view(np.uint32), w.view(np.uint32))
    _say("65,537 queries are refused")
    n = 65_537
    pos, nrm, alb = _queries(rng, n, lo, hi, dev)
    out = torch.full((n, 3), 7.0, dtype=torch.float32, device=dev)
    rc = lib.nan_nrc_infer_noalloc(cache.handle, pos.data_ptr(), nrm.data_ptr(), alb.data_ptr(), out.data_ptr(), n,
                                   stream)
    assert rc != 0 and b"65536" in lib.nan_last_error()
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
