"""Spectral sums of a diagonalised generator on the GPU (pyqed_spectral_sum of libpyqed_amd).

    out[t, a, b] = sum_j A[a, j] f(lam_j, s_t) B[j, b],   f = exp(lam s) or -1/(lam + i s)

The eigen-decomposition propagators and correlation functions (pyqed/superoperator.py:455-772,
pyqed/oqs.py:160-362) are all of this form once the generator is diagonalised on the host.
"""
import numpy as np

from . import _lib

TIME, FREQ = 0, 1


def spectral_sum(kind, A, lam, B, s, layout="tab"):
    """A [nrow, k], lam [k], B [k, ncol], s [ns] -> host array [ns, nrow, ncol] (layout 'tab') or
    [nrow, ncol, ns] (layout 'abt')."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    A = np.array(A, np.complex128, order="C", ndmin=2)
    B = np.array(B, np.complex128, order="C", ndmin=2)
    lam = np.array(lam, np.complex128, order="C").reshape(-1)
    s = np.array(s, np.float64, order="C").reshape(-1)
    nrow, k = A.shape
    ncol = B.shape[1]
    ns = s.size
    if B.shape[0] != k or lam.size != k:
        raise ValueError(f"spectral_sum: A {A.shape}, lam {lam.shape} and B {B.shape} do not agree")
    if layout not in ("tab", "abt"):
        raise ValueError("layout must be 'tab' or 'abt'")
    if ns == 0:
        return np.zeros((0, nrow, ncol) if layout == "tab" else (nrow, ncol, 0), np.complex128)
    shape = (ns, nrow, ncol) if layout == "tab" else (nrow, ncol, ns)
    strides = (nrow * ncol, ncol, 1) if layout == "tab" else (1, ncol * ns, ns)
    out = torch.empty(shape, dtype=torch.complex128, device=dev)
    wb = int(L.pyqed_spectral_work_bytes(k, ns))
    work = torch.empty(max(wb, 16), dtype=torch.uint8, device=dev)
    At, Bt = torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev)
    lt, st = torch.as_tensor(lam, device=dev), torch.as_tensor(s, device=dev)
    row_chunk = 65535
    for r0 in range(0, nrow, row_chunk):
        nr = min(row_chunk, nrow - r0)
        o = out[r0:r0 + nr] if layout == "abt" else out[:, r0:r0 + nr]
        _lib.check(L.pyqed_spectral_sum(kind, nr, k, ncol, ns, At[r0:r0 + nr].data_ptr(), lt.data_ptr(), Bt.data_ptr(),
                                        st.data_ptr(), o.data_ptr(), strides[0], strides[1], strides[2],
                                        work.data_ptr(), wb, _lib.stream_ptr(torch)))
    torch.cuda.synchronize()
    return out.cpu().numpy()
