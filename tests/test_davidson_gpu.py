"""Device Davidson (src/davidson.c restated on HBM-resident vectors) against
numpy's dense symmetric eigensolver (LAPACK syevd) on random symmetric
operators given as device matvec callbacks: lowest eigenvalue to 1e-9
relative and the eigenvector's residual below the requested tolerance."""
import numpy as np
import pytest

import t3ns_amd as T


class _DevArray:
    """Zero-copy view of raw device memory for torch.as_tensor."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"data": (int(ptr), False), "shape": (n,),
                                         "typestr": "<f8", "version": 2}


def _problem(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * 0.1
    A = (M + M.T) / 2 + np.diag(np.linspace(-spread, spread * 3, n))
    return A


@pytest.mark.gpu
@pytest.mark.parametrize("n,max_vecs,keep", [(64, 30, 2), (700, 30, 2), (700, 8, 3), (3000, 16, 2)])
def test_davidson_dense_symmetric(gpu, n, max_vecs, keep):
    import torch
    print("davidson n=%d mv=%d keep=%d" % (n, max_vecs, keep), flush=True)
    A = _problem(n, seed=n)
    w = np.linalg.eigvalsh(A)
    dA = torch.from_numpy(A).cuda()
    x0 = torch.from_numpy(np.random.default_rng(1).standard_normal(n)).cuda()
    dd = torch.from_numpy(np.diag(A).copy()).cuda()

    def mv(xp, yp, sp):
        x = torch.as_tensor(_DevArray(xp, n), device="cuda")
        y = torch.as_tensor(_DevArray(yp, n), device="cuda")
        y.copy_(dA @ x)
        torch.cuda.synchronize()

    torch.cuda.synchronize()
    e, its, rc = T.davidson_device(n, x0.data_ptr(), dd.data_ptr(), mv, 0, max_vecs, keep,
                                   tol=1e-9, max_its=400)
    assert rc == 0, (rc, its)
    assert abs(e - w[0]) <= 1e-9 * max(1.0, abs(w[0])), (e, w[0])
    v = x0.cpu().numpy()
    assert abs(np.linalg.norm(v) - 1) < 1e-10
    assert np.linalg.norm(A @ v - e * v) < 1e-8


@pytest.mark.gpu
def test_davidson_stops_at_max_its(gpu):
    import torch
    n = 500
    A = _problem(n, seed=3, spread=0.01)
    dA = torch.from_numpy(A).cuda()
    x0 = torch.from_numpy(np.random.default_rng(2).standard_normal(n)).cuda()
    dd = torch.from_numpy(np.diag(A).copy()).cuda()

    def mv(xp, yp, sp):
        y = torch.as_tensor(_DevArray(yp, n), device="cuda")
        y.copy_(dA @ torch.as_tensor(_DevArray(xp, n), device="cuda"))
        torch.cuda.synchronize()

    e, its, rc = T.davidson_device(n, x0.data_ptr(), dd.data_ptr(), mv, 0, 10, 2, tol=1e-14,
                                   max_its=3)
    assert rc == 1 and its == 3
