"""fecsim_multi_decode_static (the allocation-free mixed-queue decode: its workspace is a
static device array) against fecsim_multi_decode on the same queues: posteriors,
iteration counts and summed iterations equal, on two calls in a row, the second with a
smaller queue; a queue beyond the static bound is rejected."""
import sys

import numpy as np
import pytest

from fecsim import lib as L
from fecsim.multi import MultiDecoder
from test_gpu_multi import _queue

pytestmark = pytest.mark.gpu


def _run(m, q, static):
    import torch
    code, offs, packed, parts, css, decs = q
    n = len(code)
    ct = torch.from_numpy(code).cuda()
    ot = torch.from_numpy(offs).cuda()
    post = torch.zeros_like(packed)
    iters = torch.zeros(n, dtype=torch.int32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    (m.decode_static if static else m.decode)(ct, ot, packed, post, iters=iters, tot_iter=tot)
    torch.cuda.synchronize()
    return post.cpu().numpy(), iters.cpu().numpy(), int(tot.item())


def test_multi_decode_static_equals_multi_decode(cuda):
    import torch
    for call, n in enumerate((3000, 700)):
        print("no-alloc call", call, "queue", n, file=sys.stderr, flush=True)
        q = _queue(n, seed=40 + call)
        m = MultiDecoder(q[5], max_cw=n)
        want = _run(m, q, static=False)
        got = _run(m, q, static=True)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        assert got[2] == want[2] == int(want[1].sum())
    # beyond the static workspace: rejected before any launch
    code = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = L.lib().fecsim_multi_decode_static(m._m, 1048577, code.data_ptr(),
                                            code.data_ptr(), code.data_ptr(), code.data_ptr(), None, None, None)
    assert rc == -1
