"""Packed Hermitian state layout (host side, CPU): the helpers the engine uses to write and read y
agree with the kernels' index formula, element (u, v), u <= v, of an ADO at u(2d-u-1)/2 + v."""
import numpy as np
import pytest


@pytest.mark.parametrize("d", [1, 2, 5, 16, 32])
def test_pack_order_matches_kernel_index(d):
    from pyqed_amd.heom.engine import pack_upper
    a = np.arange(d * d, dtype=float).reshape(d, d) + 0j  # a[u, v] = u*d + v
    v = pack_upper(a[None])
    assert v.size == d * (d + 1) // 2
    for u in range(d):
        for w in range(u, d):
            assert v[(u * (2 * d - u - 1)) // 2 + w].real == u * d + w


@pytest.mark.parametrize("d", [1, 3, 12])
def test_unpack_is_hermitian_completion(d):
    from pyqed_amd.heom.engine import pack_upper, unpack_upper
    rng = np.random.default_rng(d)
    A = rng.normal(size=(4, d, d)) + 1j * rng.normal(size=(4, d, d))
    H = A + np.conj(np.transpose(A, (0, 2, 1)))
    back = unpack_upper(pack_upper(H), d)
    assert np.array_equal(back, H)
    # a non-Hermitian input keeps its upper triangle and diagonal; the lower triangle is the conjugate
    g = unpack_upper(pack_upper(A), d)
    iu = np.triu_indices(d)
    assert np.array_equal(g[:, iu[0], iu[1]], A[:, iu[0], iu[1]])
    assert np.array_equal(g, np.triu(g) + np.conj(np.transpose(np.triu(g, 1), (0, 2, 1))))
