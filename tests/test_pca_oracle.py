"""Pins the FP64 PCA oracle: randomized SVD on the implicit standardized operator must
reproduce the dense eigendecomposition of the standardized GRM (Xstd Xstd^T / M)."""

import numpy as np

import grgl_amd
from helpers import dense_x
from oracle.pca_oracle import OracleStdOperator, randomized_pca


def test_std_operator_matches_dense():
    a = grgl_amd.grg_from_coalescent(300, 1500, 3, seed=17).arrays()
    op = OracleStdOperator(a)
    Xs = op.dense_std(dense_x(a, by_individual=True))
    rng = np.random.default_rng(0)
    B = rng.standard_normal((op.shape[1], 4))
    C = rng.standard_normal((op.shape[0], 4))
    np.testing.assert_allclose(op.matmat(B), Xs @ B, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(op.rmatmat(C), Xs.T @ C, rtol=1e-10, atol=1e-9)
    # mu/sigma definition of the reference tutorial: sigma = sqrt(2 f (1 - f)), f = freq.
    f = dense_x(a).mean(axis=0)
    np.testing.assert_allclose(op.freq, f, rtol=0, atol=1e-15)


def test_randomized_pca_matches_dense_eigh():
    a = grgl_amd.grg_from_coalescent(400, 3000, 4, seed=5).arrays()
    op = OracleStdOperator(a)
    Xs = op.dense_std(dense_x(a, by_individual=True))
    grm = Xs @ Xs.T / op.shape[1]
    w, v = np.linalg.eigh(grm)
    order = np.argsort(w)[::-1][:10]
    w, v = w[order], v[:, order]
    # With enough power iterations the implementation converges to the dense answer to
    # machine precision (a wrong transpose or a dropped rank-1 term would not).
    evals, evecs = randomized_pca(op, k=10, oversample=20, power_iters=15)
    assert (np.abs(evals - w) / w).max() < 1e-9
    assert (1 - np.abs(np.sum(evecs * v, axis=0))).max() < 1e-9
    # The configured method (block 20, 5 power iterations) is an approximation whose error
    # is set by the spectral gap; it stays within 1e-2 here and improves monotonically.
    e5, _ = randomized_pca(op, k=10, oversample=10, power_iters=5)
    assert (np.abs(e5 - w) / w).max() < 1e-2


class _DenseCpuOp:
    """The device PCA's dense steps (grgl_amd.pca.TallSkinnyOps) over an explicit Xstd on the CPU."""

    def __init__(self, Xs):
        import torch

        from grgl_amd.pca import TallSkinnyOps

        self.ops = TallSkinnyOps("cpu")
        self.X = torch.from_numpy(Xs)
        self.shape = Xs.shape

    def __getattr__(self, name):
        return getattr(self.ops, name)

    @property
    def qr_mode(self):
        return self.ops.qr_mode

    @qr_mode.setter
    def qr_mode(self, v):
        self.ops.qr_mode = v

    def matmat(self, B):
        return self.X @ B

    def rmatmat(self, C):
        return self.X.T @ C


def test_tall_skinny_steps_on_cpu():
    import torch

    from grgl_amd.pca import TallSkinnyOps

    ops = TallSkinnyOps("cpu")
    rng = np.random.default_rng(8)
    for rows in (5, 1023, 1024, 70_001):
        Y = rng.standard_normal((rows, 5 if rows == 5 else 20))
        np.testing.assert_allclose(ops.gram(torch.from_numpy(Y)).numpy(), Y.T @ Y, rtol=1e-12, atol=1e-9 * rows)
    ops.begin()
    Q = ops.qr(torch.from_numpy(Y)).numpy()
    assert ops.factorizations_ok() and np.abs(Q.T @ Q - np.eye(20)).max() < 1e-13
    _, s = ops.svd_small(torch.from_numpy(Y))
    np.testing.assert_allclose(s.numpy(), np.linalg.svd(Y, compute_uv=False), rtol=1e-10)
    # A rank-deficient block must be flagged (singular Gram or R2 far from I).
    Yd = np.concatenate([Y[:, :10], Y[:, :10] @ rng.standard_normal((10, 10))], axis=1)
    ops.begin()
    ops.qr(torch.from_numpy(Yd))
    assert not ops.factorizations_ok()


def test_device_pca_algorithm_on_cpu_matches_oracle():
    """randomized_pca of the product (CholeskyQR2 + Gram eigh) against the oracle's numpy
    Householder QR + SVD with the same random block, and the fallback on a rank-deficient case."""
    from grgl_amd.pca import eigen_errors
    from grgl_amd.pca import randomized_pca as product_pca

    a = grgl_amd.grg_from_coalescent(400, 3000, 4, seed=5).arrays()
    op = OracleStdOperator(a)
    Xs = op.dense_std(dense_x(a, by_individual=True))
    dop = _DenseCpuOp(Xs)
    e, U = product_pca(dop, k=10)
    re, rU = randomized_pca(op, k=10, omega=dop.randn(op.shape[1], 20, 20261203).numpy())
    rel, ang = eigen_errors(e, U, re, rU)
    assert rel.max() < 1e-10 and ang.max() < 1e-6
    b = grgl_amd.grg_from_coalescent(60, 12, 1, seed=9).arrays()
    opb = OracleStdOperator(b)
    dop = _DenseCpuOp(opb.dense_std(dense_x(b, by_individual=True)))
    e, U = product_pca(dop, k=4, oversample=16)
    assert dop.qr_mode == "householder"
    re, rU = randomized_pca(opb, k=4, oversample=16, omega=dop.randn(opb.shape[1], 20, 20261203).numpy())
    rel, ang = eigen_errors(e, U, re, rU)
    assert rel.max() < 1e-10 and ang.max() < 1e-6


def test_oracle_operator_against_arpack():
    """Cross-check with ARPACK (scipy eigsh) on the GRM operator x -> Xstd (Xstd^T x) / M, applied
    implicitly through the oracle's traversals (no dense X)."""
    from scipy.sparse.linalg import LinearOperator, eigsh

    a = grgl_amd.grg_from_coalescent(2000, 8000, 6, seed=31).arrays()
    op = OracleStdOperator(a)
    n, m = op.shape
    grm = LinearOperator((n, n), matvec=lambda x: op.matmat(op.rmatmat(x.reshape(n, 1))).ravel() / m,
                         dtype=np.float64)
    w = np.sort(eigsh(grm, k=10, which="LA", tol=1e-12)[0])[::-1]
    e, _ = randomized_pca(op, k=10, oversample=20, power_iters=25)
    assert (np.abs(e - w) / w).max() < 1e-8
