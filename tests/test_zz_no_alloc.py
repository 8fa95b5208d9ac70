"""dsv_mlp_grad_noalloc (static device scratch, no allocation) against dsv_mlp_grad (caller
workspace), through the C-ABI: the same gradient and loss bits on two calls in a row, the
second with fewer samples (the scratch keeps the first call's partials), and a clear refusal of
a sample count beyond the scratch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grad_with_workspace(lib, params, X, y, n, stream):
    import torch
    ws = torch.empty(int(lib.dsv_mlp_grad_workspace(n)), dtype=torch.uint8, device=X.device)
    g = torch.empty(4737, dtype=torch.float32, device=X.device)
    loss = torch.empty(1, dtype=torch.float32, device=X.device)
    from dsv import _lib
    _lib.check(lib.dsv_mlp_grad(params.data_ptr(), X.data_ptr(), y.data_ptr(), n, g.data_ptr(),
                                loss.data_ptr(), ws.data_ptr(), ws.numel(), stream))
    return g, loss


def test_noalloc_gradient_matches_workspace_gradient(gpu):
    import torch
    from dsv import _lib
    from dsv.core import stream_ptr
    from dsv.mlp import glorot_params
    lib = _lib.load()
    stream = stream_ptr(gpu)
    rng = np.random.default_rng(5)
    params = torch.tensor(glorot_params(3), device=gpu)
    for n in (16384, 3001):  # the largest accepted, then a smaller one
        X = torch.tensor(rng.normal(size=(n, 7)).astype(np.float32), device=gpu)
        y = torch.tensor(rng.normal(size=n).astype(np.float32), device=gpu)
        g_ref, l_ref = _grad_with_workspace(lib, params, X, y, n, stream)
        g = torch.full((4737,), float("nan"), dtype=torch.float32, device=gpu)
        loss = torch.full((1,), float("nan"), dtype=torch.float32, device=gpu)
        _lib.check(lib.dsv_mlp_grad_noalloc(params.data_ptr(), X.data_ptr(), y.data_ptr(), n,
                                            g.data_ptr(), loss.data_ptr(), stream))
        torch.cuda.synchronize()
        assert torch.equal(g, g_ref) and torch.equal(loss, l_ref), n
        assert bool(torch.isfinite(g).all()) and float(g.abs().max()) > 0
    # beyond the scratch: refused on the host, nothing launched
    X = torch.zeros((16385, 7), dtype=torch.float32, device=gpu)
    y = torch.zeros(16385, dtype=torch.float32, device=gpu)
    rc = lib.dsv_mlp_grad_noalloc(params.data_ptr(), X.data_ptr(), y.data_ptr(), 16385,
                                  g.data_ptr(), None, stream)
    assert rc == _lib.DSV_E_ARG and b"exceeds the static scratch" in lib.dsv_last_error()
