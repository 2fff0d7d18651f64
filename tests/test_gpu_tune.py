"""The MODWT plan tuner: after wh_modwt_tune the transforms of that shape
run with the fastest plan found, and their results are bit-identical to the
default plan's (every plan computes each coefficient with the same
operations)."""
import pytest
import torch

from wavecuda_amd import device as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("filt", ["Haar", "LA8"])
def test_tuned_plan_same_results(dtype, filt):
    n, J = 1 << 20, 10
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(n, dtype=dtype, device="cuda", generator=g)
    w0 = D.modwt_fwd(x, J, filt)
    x0 = D.modwt_inv(w0, n, J, filt)
    before = (D.modwt_plan(False, n, J, filt, dtype), D.modwt_plan(True, n, J, filt, dtype))
    D.modwt_tune(n, J, filt, dtype)
    after = (D.modwt_plan(False, n, J, filt, dtype), D.modwt_plan(True, n, J, filt, dtype))
    print(filt, dtype, before, "->", after, flush=True)
    w1 = D.modwt_fwd(x, J, filt)
    x1 = D.modwt_inv(w1, n, J, filt)
    assert torch.equal(w0, w1)
    assert torch.equal(x0, x1)


def test_tune_rejects_bad_shapes():
    with pytest.raises(Exception):
        D.modwt_tune(1000, 4, "LA8")
