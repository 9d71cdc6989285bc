"""vx_topn_noalloc (top-N with no device allocation: its key buffers are a
fixed array in the library's code object) against vx_topn through the C-ABI:
two calls in a row, the second on a smaller input, each bit-identical to
vx_topn; sizes above its limit are rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import vesper_ref as ref
from vesper_python_amd import _native as nat

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 18


def _rows(fn, scores, nrows, stride, col, k):
    out = torch.full((k,), -1, dtype=torch.int32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(nat.load(), fn)(ctypes.c_void_p(scores.data_ptr()), nrows, stride, col, k,
                                 ctypes.c_void_p(out.data_ptr()), s)
    nat.check(rc, fn)
    torch.cuda.synchronize()
    return out[:min(k, nrows)].cpu().numpy()


def test_noalloc_topn_equals_topn_twice_then_smaller(gpu):
    rng = np.random.default_rng(77)
    cases = [(MAX_ROWS, 1, 0, 1024), (24624, 3, 2, 10)]      # second call: smaller input
    for nrows, stride, col, k in cases:
        print("[case] noalloc topn", nrows, stride, col, k, flush=True)
        s = rng.standard_normal(nrows * stride).astype(np.float32)
        s[rng.integers(0, nrows * stride, nrows // 8)] = 0.25          # ties resolve by row
        sd = torch.from_numpy(s).cuda()
        want = _rows("vx_topn", sd, nrows, stride, col, k)
        got = _rows("vx_topn_noalloc", sd, nrows, stride, col, k)
        assert np.array_equal(got, want)
        if nrows < 100000:
            assert list(got) == ref.stable_topn(list(s[col::stride]), k)


def test_noalloc_topn_rejects_oversize(gpu):
    sd = torch.zeros(MAX_ROWS + 1, device="cuda")
    out = torch.empty(10, dtype=torch.int32, device="cuda")
    rc = nat.load().vx_topn_noalloc(ctypes.c_void_p(sd.data_ptr()), MAX_ROWS + 1, 1, 0, 10,
                                    ctypes.c_void_p(out.data_ptr()), None)
    assert rc == -1
    assert b"at most" in nat.load().vx_last_error()
