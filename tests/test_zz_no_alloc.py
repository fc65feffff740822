"""n3_mlp_grad_noalloc (the gradient whose partial sums live in the library's
file-scope scratch array) against n3_mlp_grad, called through the C ABI:
identical bits on two calls in a row, the second with fewer tuples (so the
scratch still holds the first call's partials past the second call's
blocks)."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(pkg, dev, rng, pairs):
    def features(rows):
        f = pkg.learned.point_features(rng.integers(0, 40, rows), np.log(rng.uniform(0.05, 2.0, rows)),
                                       rng.integers(0, 600, rows))
        return torch.from_numpy(pkg.learned.to_bf16(f).view(np.int16)).to(dev)

    y = torch.from_numpy(rng.random(pairs).astype(np.float32)).to(dev)
    return features(pairs), features(pairs), y


def _grad(pkg, dev, entry, params, xu, xv, y):
    lib = pkg._lib.lib()
    grad = torch.zeros(lib.n3_mlp_params(), dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    pkg._lib.check(getattr(lib, entry)(params.data_ptr(), xu.data_ptr(), xv.data_ptr(), y.data_ptr(), y.numel(),
                                       grad.data_ptr(), loss.data_ptr(), pkg._lib.stream_handle(dev)))
    return grad, loss


def test_noalloc_gradient_equals_the_allocating_one(pkg, dev):
    rng = np.random.default_rng(21)
    w1 = rng.normal(0, math.sqrt(2 / 8), (128, 8))
    master = np.concatenate([w1.reshape(-1), rng.normal(0, 0.1, 128), rng.normal(0, math.sqrt(1 / 128), 128),
                             [0.3]]).astype(np.float32)
    params = torch.from_numpy(pkg.learned.to_bf16(master).view(np.int16)).to(dev)
    # 70,000 tuples take all 128 partial columns; 1,000 take 8 of them
    for pairs in (70000, 1000):
        xu, xv, y = _inputs(pkg, dev, rng, pairs)
        ref_grad, ref_loss = _grad(pkg, dev, "n3_mlp_grad", params, xu, xv, y)
        got_grad, got_loss = _grad(pkg, dev, "n3_mlp_grad_noalloc", params, xu, xv, y)
        torch.cuda.synchronize(dev)
        print(f"no-alloc gradient: pairs={pairs}", flush=True)
        assert torch.equal(got_grad, ref_grad), pairs
        assert torch.equal(got_loss, ref_loss), pairs
        assert torch.count_nonzero(ref_grad) > 1000


def test_noalloc_gradient_rejects_bad_arguments(pkg, dev):
    lib = pkg._lib.lib()
    x = torch.zeros(8, dtype=torch.int16, device=dev)
    y = torch.zeros(1, dtype=torch.float32, device=dev)
    grad = torch.zeros(lib.n3_mlp_params(), dtype=torch.float32, device=dev)
    with pytest.raises(ValueError):
        pkg._lib.check(lib.n3_mlp_grad_noalloc(x.data_ptr(), x.data_ptr(), x.data_ptr(), y.data_ptr(), 0,
                                               grad.data_ptr(), y.data_ptr(), pkg._lib.stream_handle(dev)))
