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
    re, rU = oracle_pca(ref, k=10)
    rel, ang = eigen_errors(e, U, re, rU)
    assert rel.max() < 1e-4, rel
    assert ang.max() < 1e-2, ang
    cg.close()
