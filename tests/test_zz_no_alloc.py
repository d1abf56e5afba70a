"""The no-allocation gradient entry point (ssmd_mlp_grad_noalloc) against
the workspace one (ssmd_mlp_grad), through the C ABI: bit-identical
gradient and loss on two calls in a row (the second smaller, so the shared
static workspace holds the first call's rows beyond what the second uses),
at the size limit, and a clean SSMD_EINVAL one sample past it."""
import numpy as np
import pytest
import torch

from ssmdlbm_amd import abi
from ssmdlbm_amd.closure import Network, bf16_bits, init_params

pytestmark = pytest.mark.gpu


def _batch(n, seed, weighted, device):
    g = torch.Generator().manual_seed(seed)
    x = bf16_bits(torch.randn(n, 8, generator=g)).to(device).contiguous()
    t = torch.randn(n, generator=g).to(device)
    w = torch.rand(n, generator=g).to(device) if weighted else None
    return x, t, w


def _with_workspace(lib, net, x, t, w, stream):
    n = x.shape[0]
    need = int(lib.ssmd_mlp_grad_work_bytes(n))
    work = torch.empty(need, dtype=torch.uint8, device=x.device)
    grad = torch.empty(abi.MLP_PARAMS, device=x.device)
    loss = torch.empty(1, device=x.device)
    abi.check(lib.ssmd_mlp_grad(net.packed.data_ptr(), x.data_ptr(), t.data_ptr(), None if w is None else w.data_ptr(),
                                n, grad.data_ptr(), loss.data_ptr(), work.data_ptr(), need, stream), "ssmd_mlp_grad")
    return grad, loss


def _no_alloc(lib, net, x, t, w, stream, n=None):
    n = x.shape[0] if n is None else n
    grad = torch.full((abi.MLP_PARAMS,), float("nan"), device=x.device)
    loss = torch.full((1,), float("nan"), device=x.device)
    rc = lib.ssmd_mlp_grad_noalloc(net.packed.data_ptr(), x.data_ptr(), t.data_ptr(),
                                   None if w is None else w.data_ptr(), n, grad.data_ptr(), loss.data_ptr(), stream)
    return rc, grad, loss


def test_noalloc_gradient_matches_the_workspace_entry_point(gpu):
    lib = abi.load()
    net = Network(init_params(7))
    stream = torch.cuda.current_stream().cuda_stream
    for n, seed, weighted in ((40000, 1, True), (1000, 2, False)):
        x, t, w = _batch(n, seed, weighted, gpu)
        rc, g1, l1 = _no_alloc(lib, net, x, t, w, stream)
        assert rc == 0, abi.load().ssmd_error_string(rc)
        g0, l0 = _with_workspace(lib, net, x, t, w, stream)
        torch.cuda.synchronize()
        assert torch.equal(g1, g0), (n, float((g1 - g0).abs().max()))
        assert torch.equal(l1, l0), (n, float(l1), float(l0))
        assert float(g0.abs().max()) > 0 and np.isfinite(float(l0))


def test_noalloc_gradient_size_limit(gpu):
    lib = abi.load()
    net = Network(init_params(8))
    stream = torch.cuda.current_stream().cuda_stream
    top = abi.MLP_GRAD_NOALLOC_MAX
    x, t, w = _batch(top + 1, 3, True, gpu)
    xs, ts, ws = x[:top].contiguous(), t[:top].contiguous(), w[:top].contiguous()
    rc, g1, l1 = _no_alloc(lib, net, xs, ts, ws, stream)
    assert rc == 0
    g0, l0 = _with_workspace(lib, net, xs, ts, ws, stream)
    torch.cuda.synchronize()
    assert torch.equal(g1, g0) and torch.equal(l1, l0)
    # one sample past the limit: rejected on the host before any launch
    rc, g2, _ = _no_alloc(lib, net, x, t, w, stream, n=top + 1)
    assert rc == abi.EINVAL
    torch.cuda.synchronize()
    assert torch.isnan(g2).all()
