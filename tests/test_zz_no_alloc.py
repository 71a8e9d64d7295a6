"""mkf_fit_adjoint_noalloc (fixed device scratch, no allocation) against mkf_fit_adjoint.

Both run the same kernels on the same inputs, so chi2 and the gradient must agree bit for bit: two calls in a
row, the second on fewer tracks (the fixed scratch keeps the first call's contents, none of which the second
may read), and a batch beyond the scratch's 64 MiB is refused without touching the device."""
import ctypes

import numpy as np
import pytest

import mkfit_amd as mk
from mkfit_amd.alignment import make_event

pytestmark = pytest.mark.gpu


def _data(ctx, n, seed):
    import torch
    ev, geo, th = make_event(n, seed=seed)
    return mk.AlignmentData(ctx, ev, "cuda:0", checkpoint_every=5), torch.from_numpy(th.astype(np.float32)).cuda()


def test_noalloc_matches_allocating_entry_point(gpu_device):
    import torch
    ctx = mk.Context(bfield=3.8)
    for n, seed in ((60000, 31), (17000, 32)):            # 50 MB, then 14 MB of the 64 MiB scratch
        data, theta = _data(ctx, n, seed)
        print("noalloc adjoint: %d tracks" % n, flush=True)
        c2n, gn = data.run(theta, noalloc=True)
        torch.cuda.synchronize()
        c2a, ga = data.run(theta)
        torch.cuda.synchronize()
        assert torch.isfinite(gn).all() and torch.isfinite(c2n).all()
        assert torch.equal(c2n, c2a), n
        assert torch.equal(gn, ga), (n, (gn - ga).abs().max().item())
    ctx.close()


def test_noalloc_refuses_batch_beyond_scratch(gpu_device):
    ctx = mk.Context(bfield=3.8)
    data, theta = _data(ctx, 90000, 33)                    # 75.6 MB > 64 MiB
    with pytest.raises(mk.MkfError, match="adjoint scratch"):
        data.run(theta, noalloc=True)
    ctx.close()
