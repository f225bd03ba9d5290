"""Standardized genotype operator and randomized-SVD PCA on the device.

The standardized (individual x site) genotype matrix follows the reference's documented
definition (doc/tutorials/LinearOperators.rst:195-207): with allele frequency f_m,
mu_m = 2 f_m and sigma_m = sqrt(2 f_m (1 - f_m)),

    Xstd[:, m] = (X[:, m] - mu_m) / sigma_m      (0 where sigma_m = 0),

X being the diploid dosage matrix (by_individual).  Products never form Xstd:

    Xstd  @ B   = DOWN_by_individual((B / sigma)^T)^T - 1 (mu/sigma)^T B
    Xstd^T @ C  = (UP_by_individual(C^T)^T - mu (1^T C)) / sigma

Precision split: the graph traversals run in FP32 on the GPU (hand-written level kernels
through the C ABI); frequencies, the rank-1 corrections, QR and the small SVD stay in FP64
on the GPU.  Frequencies come from an exact int32 UP traversal of all-ones.

`randomized_pca` is the randomized range finder with power iterations (Halko, Martinsson
and Tropp 2011, Alg. 4.4) followed by an SVD of the small projected matrix: block
k + oversample, `power_iters` rounds of Q <- qr(A qr(A^T Q)).  Eigenvalues reported are
those of Xstd Xstd^T / M (the genetic relatedness matrix, GRM).
"""

import numpy as np

from . import capi


class DeviceStdOperator:
    """Xstd of a device-resident graph; inputs/outputs are float64 torch tensors on the GPU."""

    def __init__(self, graph, stream=None):
        import torch

        self.torch = torch
        self.g = graph  # capi.Graph
        self.stream = torch.cuda.current_stream().cuda_stream if stream is None else stream
        n_ind = graph.num_samples // graph.ploidy
        self.shape = (n_ind, graph.num_mutations)
        ones = torch.ones((1, graph.num_samples), dtype=torch.int32, device="cuda")
        counts = torch.zeros((1, graph.num_mutations), dtype=torch.int32, device="cuda")
        graph.matmul_ptr(capi.I32, ones.data_ptr(), graph.num_samples, 1, capi.UP, counts.data_ptr(),
                         counts.numel(), stream=self.stream)
        self.freq = counts[0].to(torch.float64) / graph.num_samples
        self.mu = 2.0 * self.freq
        self.sigma = torch.sqrt(2.0 * self.freq * (1.0 - self.freq))
        self.inv_sigma = torch.where(self.sigma > 0, 1.0 / torch.where(self.sigma > 0, self.sigma, 1.0),
                                     torch.zeros_like(self.sigma))

    def matmat(self, B):
        """Xstd @ B for B (M x k) float64 -> (N_ind x k) float64."""
        torch = self.torch
        k = B.shape[1]
        W = (B * self.inv_sigma[:, None]).T.contiguous().to(torch.float32)  # k x M
        R = torch.zeros((k, self.shape[0]), dtype=torch.float32, device="cuda")
        self.g.matmul_ptr(capi.F32, W.data_ptr(), self.shape[1], k, capi.DOWN, R.data_ptr(), R.numel(),
                          by_individual=True, stream=self.stream)
        corr = (self.mu * self.inv_sigma) @ B  # k
        return R.T.to(torch.float64) - corr[None, :]

    def rmatmat(self, C):
        """Xstd^T @ C for C (N_ind x k) float64 -> (M x k) float64."""
        torch = self.torch
        k = C.shape[1]
        V = C.T.contiguous().to(torch.float32)  # k x N_ind
        S = torch.zeros((k, self.shape[1]), dtype=torch.float32, device="cuda")
        self.g.matmul_ptr(capi.F32, V.data_ptr(), self.shape[0], k, capi.UP, S.data_ptr(), S.numel(),
                          by_individual=True, stream=self.stream)
        colsum = C.sum(dim=0)  # k
        return (S.T.to(torch.float64) - self.mu[:, None] * colsum[None, :]) * self.inv_sigma[:, None]

    # linear-algebra backend (torch, float64, on the device)
    def randn(self, rows, cols, seed):
        rng = np.random.default_rng(seed)
        return self.torch.from_numpy(rng.standard_normal((rows, cols))).to("cuda")

    def qr(self, Y):
        return self.torch.linalg.qr(Y, mode="reduced")[0]

    def svd_small(self, B):
        U, s, _ = self.torch.linalg.svd(B, full_matrices=False)
        return U, s

    def to_numpy(self, x):
        return x.detach().cpu().numpy()


def randomized_pca(op, k=10, oversample=10, power_iters=5, seed=20261203):
    """Top-k eigenpairs of the GRM Xstd Xstd^T / M.  Returns (eigenvalues[k], eigenvectors[N_ind, k])
    as numpy float64 arrays."""
    n, m = op.shape
    ell = k + oversample
    omega = op.randn(m, ell, seed)
    Q = op.qr(op.matmat(omega))
    for _ in range(power_iters):
        Z = op.qr(op.rmatmat(Q))
        Q = op.qr(op.matmat(Z))
    B = op.rmatmat(Q).T  # ell x M  (= Q^T Xstd)
    Ub, s = op.svd_small(B)
    U = Q @ Ub
    evals = op.to_numpy(s[:k]) ** 2 / m
    evecs = op.to_numpy(U[:, :k])
    return evals, evecs


def eigen_errors(evals, evecs, ref_evals, ref_evecs):
    """Relative eigenvalue errors and eigenvector angles (radians) against a reference."""
    rel = np.abs(evals - ref_evals) / np.abs(ref_evals)
    cos = np.abs(np.sum(evecs * ref_evecs, axis=0)) / (np.linalg.norm(evecs, axis=0) * np.linalg.norm(ref_evecs, axis=0))
    ang = np.arccos(np.clip(cos, -1.0, 1.0))
    return rel, ang
