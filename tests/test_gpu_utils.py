"""GPU checks of the measurement utilities behind bench.py's ceilings: the
known-size copy kernel (byte-counter calibration, read+write ceiling) and
the streaming-read kernel (read ceiling), through the C ABI."""
import math

import pytest
import torch

import spmv_amd as S

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 5_000_001])
def test_copy_f64_copies_every_entry(n):
    g = torch.Generator(device="cpu").manual_seed(n)
    src = torch.rand(n, dtype=torch.float64, generator=g).to(DEV)
    dst = torch.full((n,), -1.0, dtype=torch.float64, device=DEV)
    S.copy_f64(dst, src)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 1000, 1_000_003, 5_000_001])
def test_read_f64_sums_every_entry_once(n):
    g = torch.Generator(device="cpu").manual_seed(7 + n)
    h = torch.rand(n, dtype=torch.float64, generator=g)
    src = h.to(DEV)
    out = torch.full((1,), 7.0, dtype=torch.float64, device=DEV)  # cleared by the call
    S.read_f64(src, out)
    torch.cuda.synchronize()
    exact = math.fsum(h.tolist())
    got = out.item()
    if n == 0:
        assert got == 0.0
    else:
        # positive terms summed as about ten per thread, a 256-lane tree and
        # at most 1024 workgroup sums added in turn: relative error below
        # ~1100 u = 1.2e-13.  A missed or doubled entry (>= 1e-7 relative
        # here for the largest n) would fail this
        assert abs(got - exact) <= 1e-12 * exact
    # a second call starts from zero again
    S.read_f64(src, out)
    torch.cuda.synchronize()
    assert abs(out.item() - got) <= 1e-12 * max(exact, 1.0)


def test_utilities_reject_misaligned_vectors():
    base = torch.zeros(1025, dtype=torch.float64, device=DEV)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    with pytest.raises(S.SpmvError, match="aligned"):
        S.read_f64(base[1:], out)  # 8 bytes past a 16-byte boundary
    with pytest.raises(S.SpmvError, match="aligned"):
        S.copy_f64(base[1:], base[:1024])
    # the failed calls leave the library usable
    S.read_f64(base[:1024], out)
    torch.cuda.synchronize()
    assert out.item() == 0.0
