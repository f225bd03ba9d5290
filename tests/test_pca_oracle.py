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
