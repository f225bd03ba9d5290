"""Device PCA (FP32 traversals, FP64 statistics and QR on the GPU) against the FP64 CPU
oracle running the same randomized SVD with the same random block.

Bars: top-10 eigenvalues within 1e-4 relative, eigenvectors within 1e-2 rad; the
difference comes only from FP32 rounding inside the traversals."""

import numpy as np
import pytest

import grgl_amd
from grgl_amd import capi
from grgl_amd.pca import DeviceStdOperator, eigen_errors, randomized_pca
from oracle.pca_oracle import OracleStdOperator
from oracle.pca_oracle import randomized_pca as oracle_pca

pytestmark = pytest.mark.gpu


def test_device_pca_matches_fp64_oracle(gpu):
    a = grgl_amd.grg_from_coalescent(4000, 20000, 5, seed=23).arrays()
    cg = capi.Graph(a)
    op = DeviceStdOperator(cg)
    ref = OracleStdOperator(a)
    # Allele counts are exact (int32 traversal); the FP64 division may differ by one ulp.
    np.testing.assert_allclose(op.freq.cpu().numpy(), ref.freq, rtol=1e-15, atol=0)
    e, U = randomized_pca(op, k=10)
    omega = op.randn(op.shape[1], 20, 20261203).cpu().numpy()
    re, rU = oracle_pca(ref, k=10, omega=omega)
    rel, ang = eigen_errors(e, U, re, rU)
    assert rel.max() < 1e-4, rel
    assert ang.max() < 1e-2, ang
    cg.close()


def test_dense_steps_against_numpy(gpu):
    """Chunked Gram, CholeskyQR2 and the Gram-based small SVD on tall FP64 blocks."""
    import torch

    a = grgl_amd.grg_from_coalescent(200, 1000, 1, seed=2).arrays()
    cg = capi.Graph(a)
    op = DeviceStdOperator(cg)
    rng = np.random.default_rng(5)
    for rows in (1, 1023, 5000, 300_001):
        Y = rng.standard_normal((rows, 3 if rows == 1 else 20))
        Yd = torch.from_numpy(Y).cuda()
        np.testing.assert_allclose(op.gram(Yd).cpu().numpy(), Y.T @ Y, rtol=1e-12, atol=1e-9 * rows)
        if rows < 20:
            continue
        op.begin()
        Q = op.qr(Yd).cpu().numpy()
        assert op.factorizations_ok()
        assert np.abs(Q.T @ Q - np.eye(20)).max() < 1e-13
        P = Q @ (Q.T @ Y)  # projection onto range(Q) reproduces Y
        assert np.abs(P - Y).max() < 1e-10 * np.abs(Y).max()
        U, s = op.svd_small(Yd)
        s_ref = np.linalg.svd(Y, compute_uv=False)
        np.testing.assert_allclose(s.cpu().numpy(), s_ref, rtol=1e-10)
    cg.close()


def test_rank_deficient_block_falls_back_to_householder(gpu):
    """More columns in the block than the standardized matrix has rank: Cholesky fails and the
    PCA is redone with Householder QR; the top eigenvalues still match the oracle."""
    a = grgl_amd.grg_from_coalescent(60, 12, 1, seed=9).arrays()
    cg = capi.Graph(a)
    op = DeviceStdOperator(cg)
    ref = OracleStdOperator(a)
    e, U = randomized_pca(op, k=4, oversample=16)
    assert op.qr_mode == "householder"
    re, rU = oracle_pca(ref, k=4, oversample=16, omega=op.randn(op.shape[1], 20, 20261203).cpu().numpy())
    rel, ang = eigen_errors(e, U, re, rU)
    assert rel.max() < 1e-4, rel
    assert ang.max() < 1e-2, ang
    cg.close()
