"""Beyond the benchmark size: 64-bit lengths and grids at n = 2^26 and
2^28, checked through properties that do not depend on n -- perfect
reconstruction, Parseval (DWT), the per-level energy identity
||S_j||^2 + ||D_j||^2 = 2 ||S_{j-1}||^2 (MODWT), the exact MAD against a
full sort, and FP32 against FP64."""
import pytest
import torch

from wavecuda_amd import device as D

pytestmark = pytest.mark.gpu


def _x(n, seed, dtype=torch.float64):
    return torch.randn(n, dtype=dtype, device="cuda", generator=torch.Generator("cuda").manual_seed(seed))


def test_dwt_2p28_full_depth():
    n = 1 << 28
    print("DWT LA8 n=2^28 full depth", flush=True)
    x = _x(n, 1)
    y = D.dwt_fwd(x, 0, "LA8")
    ex, ey = torch.dot(x, x).item(), torch.dot(y, y).item()
    assert abs(ey - ex) / ex <= 1e-12
    xr = D.dwt_inv(y, 0, "LA8")
    del y
    assert (xr - x).abs().max().item() / x.abs().max().item() <= 1e-12


def test_modwt_2p26_energy_and_reconstruction():
    n, J = 1 << 26, 10
    print("MODWT LA8 n=2^26 J=10", flush=True)
    x = _x(n, 2)
    w = D.modwt_fwd(x, J, "LA8")
    prev = torch.dot(x, x).item()
    for j in range(J):
        s, d = w[2 * n * j: 2 * n * j + n], w[2 * n * j + n: 2 * n * (j + 1)]
        es = torch.dot(s, s).item()
        e = es + torch.dot(d, d).item()
        assert abs(e - 2 * prev) / (2 * prev) <= 1e-12, j
        prev = es
    xr = D.modwt_inv(w, n, J, "LA8")
    assert (xr - x).abs().max().item() / x.abs().max().item() <= 1e-12


def test_mad_2p26_exact():
    n = 1 << 26
    print("MAD n=2^26", flush=True)
    v = _x(n, 3)
    s = torch.sort(v).values
    med = 0.5 * (s[(n - 1) // 2] + s[n // 2])
    dev = torch.sort((v - med).abs()).values
    mad = 1.4826 * (0.5 * (dev[(n - 1) // 2] + dev[n // 2]))
    assert D.median(v).item() == med.item()
    assert D.mad(v).item() == mad.item()


def test_modwt_fp32_2p26():
    n, J = 1 << 26, 10
    print("MODWT FP32 n=2^26 J=10", flush=True)
    x = _x(n, 4, torch.float32)
    w32 = D.modwt_fwd(x, J, "D4")
    w64 = D.modwt_fwd(x.double(), J, "D4")
    rel = (torch.linalg.vector_norm(w32.double() - w64) / torch.linalg.vector_norm(w64)).item()
    assert rel <= 1e-5
    del w64
    xr = D.modwt_inv(w32, n, J, "D4")
    assert (torch.linalg.vector_norm(xr.double() - x.double())
            / torch.linalg.vector_norm(x.double())).item() <= 1e-5


def test_mad_fp32_2p26_exact():
    """FP32 selections beyond 2^22 values take three grid digits, so the
    keys left for the finishing workgroup fit its LDS: the median is exact
    (a full sort of the same values), the MAD (deviations formed in FP32)
    within FP32 rounding of the FP64 one."""
    n = 1 << 26
    print("MAD FP32 n=2^26", flush=True)
    v = _x(n, 5, torch.float32)
    s = torch.sort(v.double()).values
    med = 0.5 * (s[(n - 1) // 2] + s[n // 2])
    dev = torch.sort((v.double() - med).abs()).values
    mad = 1.4826 * (0.5 * (dev[(n - 1) // 2] + dev[n // 2]))
    assert D.median(v).item() == med.item()
    assert abs(D.mad(v).item() - mad.item()) <= 1e-6 * mad.item()
