"""The SVD as the kernels run it on the GPU (mpm_device_svd: the Jacobi
rotations with the hardware square root, reciprocal and reciprocal square
root estimates, which the host build of the same code does not use) against LAPACK (numpy.linalg.svd) and
scipy.linalg.polar, on near-identity, random, diagonal, rotated, rank-
deficient, repeated-singular-value and reflected matrices.  Bars as the
host check of the same code: reconstruction 2e-5 of the matrix's scale,
U and V orthonormal to 2e-6 with det +1, singular values 2e-5 of the
largest, and U V^T within 1e-4 of the polar rotation where det F > 1e-2."""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu


def _rot(rng, dim):
    Q = np.linalg.qr(rng.normal(size=(dim, dim)))[0]
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    return Q


def _cases(dim, rng):
    out = []
    for t in range(3000):
        k = t % 8
        if k == 0:
            A = np.eye(dim) + 10.0 ** rng.uniform(-6, -1.3) * rng.normal(size=(dim, dim))
        elif k == 1:
            A = rng.normal(size=(dim, dim))
        elif k == 2:
            A = np.diag(rng.uniform(0.5, 1.5, dim))
        elif k == 3:  # F = R diag(1 +- 0.05) Q^T, the solver's working range
            A = _rot(rng, dim) @ np.diag(1 + 0.05 * rng.uniform(-1, 1, dim)) @ _rot(rng, dim).T
        elif k == 4:  # rank deficient
            s = rng.uniform(0.5, 1.5, dim)
            s[rng.integers(dim):] = 0.0
            A = _rot(rng, dim) @ np.diag(s) @ _rot(rng, dim).T
        elif k == 5:  # repeated singular values
            s = np.full(dim, rng.uniform(0.5, 1.5))
            if dim == 3 and t % 16 == 5:
                s[2] = rng.uniform(0.5, 1.5)
            A = _rot(rng, dim) @ np.diag(s) @ _rot(rng, dim).T
        elif k == 6:  # reflections (det < 0)
            A = _rot(rng, dim) @ np.diag(rng.uniform(0.8, 1.2, dim) * np.r_[np.ones(dim - 1), -1])
        else:  # a pure rotation, and the zero matrix now and then
            A = _rot(rng, dim) if t % 64 != 7 else np.zeros((dim, dim))
        out.append(A)
    return np.asarray(out, np.float32)


@pytest.mark.parametrize("dim", [2, 3])
def test_device_svd_against_lapack(dim):
    from taichi_elements_amd import _lib
    from taichi_elements_amd.mpm_solver import check
    L = _lib.lib()
    A = _cases(dim, np.random.default_rng(60 + dim))
    n = len(A)
    U = np.zeros_like(A)
    V = np.zeros_like(A)
    s = np.zeros((n, dim), np.float32)
    check(L.mpm_device_svd(0, dim, n, _lib.fptr(A), _lib.fptr(U), _lib.fptr(s), _lib.fptr(V)))
    worst = dict(rec=0.0, orth=0.0, det=0.0, sig=0.0, polar=0.0)
    for i in range(n):
        A64, U64, V64, s64 = (A[i].astype(float), U[i].astype(float), V[i].astype(float),
                              s[i].astype(float))
        scale = max(1.0, np.abs(A64).max())
        worst["rec"] = max(worst["rec"], np.abs((U64 * s64) @ V64.T - A64).max() / scale)
        worst["orth"] = max(worst["orth"], np.abs(U64.T @ U64 - np.eye(dim)).max(),
                            np.abs(V64.T @ V64 - np.eye(dim)).max())
        worst["det"] = max(worst["det"], abs(np.linalg.det(U64) - 1), abs(np.linalg.det(V64) - 1))
        ref = np.linalg.svd(A64, compute_uv=False)
        worst["sig"] = max(worst["sig"], np.abs(np.abs(s64) - ref).max() / max(ref[0], 1e-30)
                           if ref[0] > 0 else np.abs(s64).max())
        if np.linalg.det(A64) > 1e-2:
            R, _ = scipy.linalg.polar(A64)
            worst["polar"] = max(worst["polar"], np.abs(U64 @ V64.T - R).max())
    print(dim, {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    assert worst["rec"] <= 2e-5
    assert worst["orth"] <= 2e-6
    assert worst["det"] <= 1e-5
    assert worst["sig"] <= 2e-5
    assert worst["polar"] <= 1e-4
    # singular values ordered by magnitude, the last one carrying det's sign
    assert np.all(np.abs(s[:, :-1]) >= np.abs(s[:, 1:]) * (1 - 1e-5) - 1e-6)
