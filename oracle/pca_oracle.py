"""ORACLE — TEST INFRASTRUCTURE ONLY.

FP64 CPU restatement of the standardized-genotype PCA: the standardized operator of
doc/tutorials/LinearOperators.rst:195-207 (f = allele frequency over haplotypes,
mu = 2f, sigma = sqrt(2f(1-f)), columns with sigma = 0 dropped to 0) applied through the
oracle's matmul (by_individual), and the randomized SVD with power iterations of Halko,
Martinsson & Tropp (2011, Alg. 4.4) in numpy.  Used by tests and the benchmark to measure
the device PCA's eigenvalue and eigenvector errors; pinned in tests against a dense
eigendecomposition of Xstd Xstd^T built from the genotype matrix.
"""

import numpy as np

from . import DOWN, UP, OracleGraph


class OracleStdOperator:
    def __init__(self, arrays):
        self.g = OracleGraph(arrays)
        n_ind = self.g.num_samples // self.g.ploidy
        self.shape = (n_ind, self.g.num_mutations)
        counts = self.g.matmul(np.ones((1, self.g.num_samples), np.int64), UP)[0]
        self.freq = counts.astype(np.float64) / self.g.num_samples
        self.mu = 2.0 * self.freq
        self.sigma = np.sqrt(2.0 * self.freq * (1.0 - self.freq))
        self.inv_sigma = np.divide(1.0, self.sigma, out=np.zeros_like(self.sigma), where=self.sigma > 0)

    def matmat(self, B):
        R = self.g.matmul(np.ascontiguousarray((B * self.inv_sigma[:, None]).T), DOWN, by_individual=True)
        return R.T - ((self.mu * self.inv_sigma) @ B)[None, :]

    def rmatmat(self, C):
        S = self.g.matmul(np.ascontiguousarray(C.T), UP, by_individual=True)
        return (S.T - self.mu[:, None] * C.sum(axis=0)[None, :]) * self.inv_sigma[:, None]

    def dense_std(self, X_ind):
        """Xstd from a dense dosage matrix (individuals x sites)."""
        return (X_ind - self.mu[None, :]) * self.inv_sigma[None, :]


def randomized_pca(op, k=10, oversample=10, power_iters=5, seed=20261203, omega=None):
    """`omega` (M x (k + oversample)) is the random block; numpy's generator draws one from
    `seed` when it is not given (pass the device's block to run the same algorithm)."""
    n, m = op.shape
    ell = k + oversample
    if omega is None:
        omega = np.random.default_rng(seed).standard_normal((m, ell))
    Q, _ = np.linalg.qr(op.matmat(omega))
    for _ in range(power_iters):
        Z, _ = np.linalg.qr(op.rmatmat(Q))
        Q, _ = np.linalg.qr(op.matmat(Z))
    B = op.rmatmat(Q).T
    Ub, s, _ = np.linalg.svd(B, full_matrices=False)
    U = Q @ Ub
    return s[:k] ** 2 / m, U[:, :k]
