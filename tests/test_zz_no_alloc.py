"""t3ns_davidson_dev_noalloc (the allocation-free variant of
t3ns_davidson_dev) against t3ns_davidson_dev itself, through the C ABI, on
the same operator, guess and diagonal: the same eigenvalue (1e-12
relative), the same eigenvector up to sign (1e-10) and the same iteration
count, on two calls in a row, the second on a smaller problem (the fixed
workspace keeps the first call's contents).  Sizes beyond the fixed
workspace are refused with an error."""
import numpy as np
import pytest

from t3ns_amd.heff import davidson_device, davidson_device_noalloc


class _DevArray:
    """Zero-copy view of raw device memory for torch.as_tensor."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"data": (int(ptr), False), "shape": (n,),
                                         "typestr": "<f8", "version": 2}


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * 0.1
    return (M + M.T) / 2 + np.diag(np.linspace(-1.0, 3.0, n))


def _solve(fn, A, guess, max_vecs=30, keep=2, tol=1e-10, max_its=300):
    import torch
    n = A.shape[0]
    dA = torch.from_numpy(A).cuda()
    x = torch.from_numpy(guess.copy()).cuda()
    dd = torch.from_numpy(np.diag(A).copy()).cuda()

    def mv(xp, yp, sp):
        y = torch.as_tensor(_DevArray(yp, n), device="cuda")
        y.copy_(dA @ torch.as_tensor(_DevArray(xp, n), device="cuda"))
        torch.cuda.synchronize()

    torch.cuda.synchronize()
    e, its, rc = fn(n, x.data_ptr(), dd.data_ptr(), mv, 0, max_vecs, keep, tol, max_its)
    torch.cuda.synchronize()
    return e, x.cpu().numpy(), its, rc


@pytest.mark.gpu
def test_noalloc_davidson_matches_allocating_entry_point(gpu):
    for n, seed in ((2500, 1), (900, 2)):  # the second call is smaller
        print("noalloc davidson n=%d" % n, flush=True)
        A = _problem(n, seed)
        guess = np.random.default_rng(seed + 10).standard_normal(n)
        e0, x0, i0, r0 = _solve(davidson_device, A, guess)
        e1, x1, i1, r1 = _solve(davidson_device_noalloc, A, guess)
        assert r0 == 0 and r1 == 0, (r0, r1)
        assert i1 == i0
        assert abs(e1 - e0) <= 1e-12 * abs(e0), (e0, e1)
        sign = np.sign(np.dot(x0, x1))
        assert np.max(np.abs(x1 - sign * x0)) < 1e-10
        w = np.linalg.eigvalsh(A)[0]
        assert abs(e1 - w) <= 1e-9 * max(1.0, abs(w))


@pytest.mark.gpu
def test_noalloc_davidson_refuses_oversized_problem(gpu):
    import torch
    n = 200_000  # (2 x 30 + 3) x n doubles exceed the fixed 64 MiB
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    d = torch.ones(n, dtype=torch.float64, device="cuda")

    def mv(xp, yp, sp):  # never called: the size check comes first
        raise AssertionError("matvec called on a refused problem")

    with pytest.raises(RuntimeError, match="workspace"):
        davidson_device_noalloc(n, x.data_ptr(), d.data_ptr(), mv, 0, 30, 2, 1e-8, 10)
