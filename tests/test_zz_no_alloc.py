"""num_dot_noalloc (the dot product that allocates nothing) through the C ABI
against num_dot, the entry point it is a variant of: the same kernels, so
the results must be identical bit for bit, on consecutive calls whose
second one is smaller (the static scratch keeps the first call's partials),
and on an empty vector."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from num_amd import _lib
from num_amd._lib import call

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def _dot_ref(x, y, stream):
    n = x.shape[0]
    ws = torch.empty(int(_lib.load().num_dot_workspace_bytes(n)) // 8 + 1,
                     dtype=torch.float64, device="cuda")
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    call("num_dot", n, _p(x), _p(y), _p(out), _p(ws), stream)
    return out


def _dot_noalloc(x, y, stream):
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    call("num_dot_noalloc", x.shape[0], _p(x), _p(y), _p(out), stream)
    return out


def test_dot_noalloc_equals_dot(ks):
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(11)
    # the largest first (every one of the 2048 partials written), then smaller
    # ones that use fewer partials than the previous call left behind
    for n in (5_000_000, 3_001, 1, 0):
        xa = rng.uniform(-1, 1, n)
        ya = rng.uniform(-1, 1, n)
        x = torch.from_numpy(xa).cuda()
        y = torch.from_numpy(ya).cuda()
        ref = _dot_ref(x, y, st)
        got = _dot_noalloc(x, y, st)
        torch.cuda.synchronize()
        print(f"no-alloc dot n={n}", flush=True)
        assert torch.equal(got, ref), (n, got.item(), ref.item())
        exact = math.fsum((xa * ya).tolist())
        assert abs(got.item() - exact) <= 1e-12 * max(1.0, float(np.abs(xa * ya).sum()))


def test_dot_noalloc_errors(ks):
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        call("num_dot_noalloc", -1, None, None, _p(out), None)
    with pytest.raises(ValueError):
        call("num_dot_noalloc", 10, None, None, _p(out), None)
    with pytest.raises(ValueError):
        call("num_dot_noalloc", 0, None, None, None, None)
