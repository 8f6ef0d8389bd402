"""qm_reduce_noalloc (fixed device scratch, no allocation) against qm_reduce through the C-ABI:
the same bits for every operation and dtype, on two calls in a row with a smaller input
second, and the size limit refused before anything runs."""
import ctypes

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _both(op, dtype, a, b, n):
    from qmtools import _native as nat
    lib, ctx, s = nat.library(), nat.context(), nat.stream()
    r1, r2 = ctypes.c_double(0.0), ctypes.c_double(0.0)
    pa = nat.dptr(a)
    pb = nat.dptr(b) if b is not None else None
    assert lib.qm_reduce(ctx, s, op, dtype, pa, pb, n, ctypes.byref(r1)) == nat.QM_OK
    rc = lib.qm_reduce_noalloc(ctx, s, op, dtype, pa, pb, n, ctypes.byref(r2))
    assert rc == nat.QM_OK, nat.library().qm_last_error(ctx)
    return r1.value, r2.value


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_reduce_noalloc_equals_reduce(dtype):
    import torch
    from qmtools import _native as nat
    code = nat.QM_F64 if dtype == "f64" else nat.QM_F32
    tdt = torch.float64 if dtype == "f64" else torch.float32
    g = torch.Generator(device="cpu").manual_seed(3)
    # 9.4M elements (2048 blocks of > 4096 each), then 70001 (18 blocks), then 1 and 0
    for n in (9_400_001, 70_001, 1, 0):
        a = torch.randn(max(n, 1), generator=g, dtype=torch.float64).to(tdt).cuda()
        b = (torch.rand(max(n, 1), generator=g, dtype=torch.float64) + 1.0).to(tdt).cuda()
        for op in (nat.QM_RED_SUM, nat.QM_RED_ABSDIFF, nat.QM_RED_RELDIFF):
            _say("reduce_noalloc", dtype, n, op)
            r1, r2 = _both(op, code, a, b if op != nat.QM_RED_SUM else None, n)
            assert r2 == r1, (n, op, r1, r2)
            if n == 0:
                assert r2 == 0.0
    # the oracle's sum of the last large input, as a check that both are right, not only equal
    from oracle import reduce as orr
    x = torch.randn(123_457, generator=g, dtype=torch.float64).cuda()
    r1, r2 = _both(nat.QM_RED_SUM, nat.QM_F64, x, None, x.numel())
    assert r2 == pytest.approx(orr.total(x.cpu().numpy()), rel=1e-13, abs=1e-12)


def test_reduce_noalloc_refuses_size_beyond_limit():
    import torch
    from qmtools import _native as nat
    lib, ctx = nat.library(), nat.context()
    a = torch.zeros(8, dtype=torch.float64, device="cuda")
    out = ctypes.c_double(-1.0)
    _say("reduce_noalloc size limit")
    rc = lib.qm_reduce_noalloc(ctx, nat.stream(), nat.QM_RED_SUM, nat.QM_F64, nat.dptr(a), None, 1 << 34,
                               ctypes.byref(out))
    assert rc == nat.QM_ELIMIT
    assert b"2^34" in lib.qm_last_error(ctx)
    assert out.value == -1.0
