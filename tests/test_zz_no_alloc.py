"""bfd_eval_data_noalloc: the allocation-free form of bfd_eval_data on
device-resident evaluations, called through the C ABI.  Two calls in a row
(the second smaller, so stale scratch contents from the first would show)
must equal bfd_eval_data's values and the restatement's MLEs; sizes beyond
the fixed scratch are refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(bfd, ctx, dev_ptr, height, width, point):
    pt = np.ascontiguousarray(point if len(point) else np.zeros((1, 2), np.uint64))
    out = np.zeros((width, 2), np.uint64)
    bfd._check(bfd.lib().bfd_eval_data_noalloc(ctx.h, ctypes.c_void_p(dev_ptr), height, width, bfd._p(pt),
                                               bfd._p(out)))
    return out


def test_eval_data_noalloc_equals_eval_data(bfd, oracle):
    import torch
    ctx = bfd.default_context()
    cases = [(18, 3), (11, 1), (0, 4), (9, 4099), (20, 1), (5, 2)]
    for n, w in cases:
        print(f"noalloc eval n={n} w={w}", flush=True)
        data, point = bfd.random_vec(700 + n + w, (1 << n) * w), bfd.random_vec(800 + n, n)
        dev = torch.from_numpy(data.view(np.int64).reshape(-1)).to("cuda:0")
        torch.cuda.synchronize()
        got = _noalloc(bfd, ctx, dev.data_ptr(), 1 << n, w, point)
        exp = bfd.eval_data(None, point, w, ctx, device_ptr=dev.data_ptr(), height=1 << n)
        assert (got == exp).all(), (n, w)
        if w <= 4:
            mat = data.reshape(1 << n, w, 2)
            ref = np.concatenate([oracle.mle_evaluate(np.ascontiguousarray(mat[:, c]), point) for c in range(w)])
            assert (got == ref).all(), (n, w)
        assert (dev.cpu().numpy() == data.view(np.int64).reshape(-1)).all()


def test_eval_data_noalloc_refuses_oversize(bfd):
    import torch
    ctx = bfd.default_context()
    dev = torch.zeros(64, dtype=torch.int64, device="cuda:0")
    pt = np.zeros((33, 2), np.uint64)
    out = np.zeros((65536, 2), np.uint64)
    L = bfd.lib()
    # the checks run before any device work, so the small buffer is never read
    assert L.bfd_eval_data_noalloc(ctx.h, ctypes.c_void_p(dev.data_ptr()), 1 << 33, 1, bfd._p(pt),
                                   bfd._p(out)) == bfd.BFD_ERR_ARG
    assert L.bfd_eval_data_noalloc(ctx.h, ctypes.c_void_p(dev.data_ptr()), 1, 65536, bfd._p(pt),
                                   bfd._p(out)) == bfd.BFD_ERR_ARG
    assert L.bfd_eval_data_noalloc(ctx.h, ctypes.c_void_p(dev.data_ptr()), 3, 1, bfd._p(pt),
                                   bfd._p(out)) == bfd.BFD_ERR_ARG
