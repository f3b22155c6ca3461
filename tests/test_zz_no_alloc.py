"""pyqed_spectral_sum_static, the entry point that allocates nothing (its weights live in a static
device array), against pyqed_spectral_sum with a caller-provided workspace, called through the C ABI:
bitwise equal on two calls in a row, the second smaller than the first (so the static array still
holds the first call's weights), on both kernel forms (MFMA for >= 48 output columns, VALU below) and
both kinds; plus the size limit, which must be rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(rng, nrow, k, ncol, ns):
    A = rng.normal(size=(nrow, k)) + 1j * rng.normal(size=(nrow, k))
    B = rng.normal(size=(k, ncol)) + 1j * rng.normal(size=(k, ncol))
    lam = -rng.uniform(0.05, 1.0, k) + 1j * rng.normal(size=k)
    s = np.linspace(0.0, 3.0, ns)
    return A, lam, B, s


def _call(static, kind, A, lam, B, s, tfast):
    import torch
    from pyqed_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    nrow, k = A.shape
    ncol, ns = B.shape[1], s.size
    At, Bt = torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev)
    lt, st = torch.as_tensor(lam, device=dev), torch.as_tensor(s, device=dev)
    shape = (nrow, ncol, ns) if tfast else (ns, nrow, ncol)
    strides = (1, ncol * ns, ns) if tfast else (nrow * ncol, ncol, 1)
    out = torch.full(shape, float("nan"), dtype=torch.complex128, device=dev)
    stream = _lib.stream_ptr(torch)
    if static:
        rc = L.pyqed_spectral_sum_static(kind, nrow, k, ncol, ns, At.data_ptr(), lt.data_ptr(), Bt.data_ptr(),
                                         st.data_ptr(), out.data_ptr(), *strides, stream)
    else:
        wb = int(L.pyqed_spectral_work_bytes(k, ns))
        work = torch.full((wb // 16,), float("nan"), dtype=torch.complex128, device=dev)  # garbage, not zeros
        rc = L.pyqed_spectral_sum(kind, nrow, k, ncol, ns, At.data_ptr(), lt.data_ptr(), Bt.data_ptr(),
                                  st.data_ptr(), out.data_ptr(), *strides, work.data_ptr(), wb, stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return np.moveaxis(o, 2, 0) if tfast else o


def _numpy(kind, A, lam, B, s):
    f = np.exp(np.outer(s, lam)) if kind == 0 else -1.0 / (lam[None, :] + 1j * s[:, None])
    return np.einsum("aj,tj,jb->tab", A, f, B, optimize=True)


@pytest.mark.parametrize("kind", [0, 1])
def test_static_scratch_matches_workspace_entry_point(kind):
    rng = np.random.default_rng(7 + kind)
    cases = [(3, 600, 64, 1500, False),   # 900000 weights, MFMA form, [t][a][b]
             (2, 40, 5, 77, True),        # smaller second call: VALU form, [a][b][t]
             (4, 256, 50, 1024, True)]    # MFMA form with t contiguous
    for nrow, k, ncol, ns, tfast in cases:
        A, lam, B, s = _inputs(rng, nrow, k, ncol, ns)
        ref = _call(False, kind, A, lam, B, s, tfast)
        got = _call(True, kind, A, lam, B, s, tfast)
        assert np.array_equal(got, ref), (nrow, k, ncol, ns)
        want = _numpy(kind, A, lam, B, s)
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), (nrow, k, ncol, ns)


def test_static_scratch_rejects_sizes_above_its_limit():
    import torch
    from pyqed_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    k, ns = 1025, 1024  # one weight row more than the 2^20 the static array holds
    assert k * ns > _lib.SPECTRAL_STATIC_MAX
    dummy = torch.zeros(16, dtype=torch.complex128, device=dev)
    p = dummy.data_ptr()
    rc = L.pyqed_spectral_sum_static(0, 1, k, 1, ns, p, p, p, p, p, 1, 1, 1, _lib.stream_ptr(torch))
    assert rc != 0
    assert b"exceed the static scratch" in L.pyqed_last_error()
    torch.cuda.synchronize()
    assert (k - 1) * ns <= _lib.SPECTRAL_STATIC_MAX  # the largest accepted row count for this ns
