"""naga_knn_noalloc: the kNN with device pointers and the library's static
device scratch (no device allocation) against naga_knn, bit for bit, on
two calls in a row (the second with fewer reference and query points, so
stale scratch contents from the first call would show), and its size
limit.  Device buffers come from the HIP runtime the library is linked
against (naga_amd.devmem)."""
import numpy as np
import pytest

import naga_amd
from naga_amd import devmem

pytestmark = pytest.mark.gpu


def _noalloc(ref, query, k):
    D = ref.shape[1]
    dref, dq = devmem.DeviceArray(ref), devmem.DeviceArray(query)
    fill = np.full((len(query), k), -1, np.int32)
    dd2, didx = devmem.DeviceArray(fill.view(np.float32)), devmem.DeviceArray(fill)
    st = naga_amd.load().naga_knn_noalloc(dref.ptr, len(ref), dq.ptr, len(query), D, k,
                                          dd2.ptr, didx.ptr, None)
    assert st == 0, naga_amd.load().naga_last_error()
    devmem.synchronize()
    out = (dd2.to_host(np.float32, (len(query), k)), didx.to_host(np.uint32, (len(query), k)))
    for a in (dref, dq, dd2, didx):
        a.free()
    return out


@pytest.mark.parametrize("D,k", [(3, 32), (2, 16)])
def test_knn_noalloc_matches_knn(D, k):
    rng = np.random.default_rng(7 + D)
    for n_ref, n_q in ((200_000, 50_000), (30_000, 7_000)):
        ref = rng.random((n_ref, D), dtype=np.float32)
        query = rng.random((n_q, D), dtype=np.float32)
        print(f"noalloc kNN D={D} k={k} n_ref={n_ref} n_q={n_q}", flush=True)
        want_d2, want_idx = naga_amd.batched_nearest_neighbors(k, query, ref)
        got_d2, got_idx = _noalloc(ref, query, k)
        assert np.array_equal(got_idx, want_idx)
        assert np.array_equal(got_d2, want_d2)


def test_knn_noalloc_rejects_oversize():
    n = (1 << 20) + 1
    dref = devmem.DeviceArray(nbytes=n * 3 * 4)
    out = devmem.DeviceArray(nbytes=64)
    st = naga_amd.load().naga_knn_noalloc(dref.ptr, n, dref.ptr, 1, 3, 4, out.ptr, out.ptr, None)
    assert st == 1
    assert "no-allocation limit" in naga_amd.load().naga_last_error().decode()
    dref.free()
    out.free()
