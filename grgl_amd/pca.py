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

Every dense step is tall and skinny (M or N_ind rows, k + oversample columns).  Library
GEMMs whose reduction runs over the long dimension are not split on this hardware and run
as a handful of workgroups (a 1 x M by M x 20 product takes tens of milliseconds), so the
long reductions are cut into row chunks: Gram matrices are one batched product over
chunks plus a sum, the rank-1 corrections are column reductions.  Orthonormalisation is
CholeskyQR2 (Fukaya, Nakatsukasa et al. 2014: R = chol(Y^T Y), Y <- Y R^-1, twice), which
is backward stable once Y^T Y is safely positive definite; a failed factorisation is
detected once at the end and the whole PCA is then redone with Householder QR.  The final
small SVD of B = Q^T Xstd comes from the eigendecomposition of B B^T.

`randomized_pca` is the randomized range finder with power iterations (Halko, Martinsson
and Tropp 2011, Alg. 4.4) followed by an SVD of the small projected matrix: block
k + oversample, `power_iters` rounds of Q <- qr(A qr(A^T Q)).  Eigenvalues reported are
those of Xstd Xstd^T / M (the genetic relatedness matrix, GRM).
"""

import numpy as np

from . import capi


class TallSkinnyOps:
    """FP64 dense steps of the randomized SVD on tall-skinny torch tensors (any device)."""

    qr_mode = "cholesky2"

    def __init__(self, device="cuda"):
        import torch

        self.torch = torch
        self.device = device
        self._checks = []

    def randn(self, rows, cols, seed):
        """Seeded N(0,1) float64 block drawn on the operator's device (Philox on the GPU):
        the same seed gives the same block on every call."""
        torch = self.torch
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        return torch.randn((rows, cols), generator=gen, dtype=torch.float64, device=self.device)

    def gram(self, Y):
        """Y^T Y for a tall (rows x k) Y: per-chunk products in one batched GEMM, then a sum."""
        torch = self.torch
        rows, k = Y.shape
        chunks = max(1, min(256, rows // 1024))
        per = -(-rows // chunks)
        if per * chunks != rows:
            Y = torch.cat([Y, Y.new_zeros(per * chunks - rows, k)])
        Yc = Y.view(chunks, per, k)
        return torch.bmm(Yc.transpose(1, 2).contiguous(), Yc).sum(dim=0)

    def qr(self, Y):
        """Orthonormal basis of range(Y) (Y tall, full column rank)."""
        torch = self.torch
        if self.qr_mode == "householder":
            return torch.linalg.qr(Y, mode="reduced")[0]
        eye = torch.eye(Y.shape[1], dtype=Y.dtype, device=Y.device)
        R1, info1 = torch.linalg.cholesky_ex(self.gram(Y), upper=True)
        Y = Y @ torch.linalg.solve_triangular(R1, eye, upper=True)
        R2, info2 = torch.linalg.cholesky_ex(self.gram(Y), upper=True)
        Y = Y @ torch.linalg.solve_triangular(R2, eye, upper=True)
        # The second pass is accurate when the first left Y nearly orthonormal, i.e. R2 ~ I.
        # A singular or ill-conditioned block either fails a factorisation or leaves R2 far from I.
        dev = torch.nan_to_num((R2 - eye).abs().max(), nan=1.0e300)
        self._checks.append((info1 == 0) & (info2 == 0) & (dev < 0.25))
        return Y

    def begin(self):
        self._checks = []

    def factorizations_ok(self):
        """False when an orthonormalisation of the last run may have lost orthogonality."""
        return bool(self.torch.stack(self._checks).all()) if self._checks else True

    def svd_small(self, Bt):
        """Left singular vectors and singular values (descending) of B = Bt^T, Bt tall."""
        torch = self.torch
        w, V = torch.linalg.eigh(self.gram(Bt))
        w, V = w.flip(0), V.flip(1)
        return V, torch.sqrt(torch.clamp(w, min=0.0))

    def to_numpy(self, x):
        return x.detach().cpu().numpy()


class DeviceStdOperator(TallSkinnyOps):
    """Xstd of a device-resident graph; inputs/outputs are float64 torch tensors on the GPU."""

    def __init__(self, graph, stream=None):
        super().__init__("cuda")
        torch = self.torch
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
        corr = (B * (self.mu * self.inv_sigma)[:, None]).sum(dim=0)  # k
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


def randomized_pca(op, k=10, oversample=10, power_iters=5, seed=20261203, omega=None):
    """Top-k eigenpairs of the GRM Xstd Xstd^T / M.  Returns (eigenvalues[k], eigenvectors[N_ind, k])
    as numpy float64 arrays.  The random block is op.randn(M, k + oversample, seed) unless
    `omega` (a tensor on the operator's device) is given."""
    n, m = op.shape
    ell = k + oversample
    op.qr_mode = "cholesky2"
    while True:
        op.begin()
        if omega is None:
            omega = op.randn(m, ell, seed)
        Q = op.qr(op.matmat(omega))
        for _ in range(power_iters):
            Z = op.qr(op.rmatmat(Q))
            Q = op.qr(op.matmat(Z))
        if op.qr_mode != "householder" and not op.factorizations_ok():
            op.qr_mode = "householder"
            continue
        Ub, s = op.svd_small(op.rmatmat(Q))  # rmatmat(Q) = (Q^T Xstd)^T
        U = Q @ Ub
        return op.to_numpy(s[:k]) ** 2 / m, op.to_numpy(U[:, :k])


def eigen_errors(evals, evecs, ref_evals, ref_evecs):
    """Relative eigenvalue errors and eigenvector angles (radians) against a reference."""
    rel = np.abs(evals - ref_evals) / np.abs(ref_evals)
    cos = np.abs(np.sum(evecs * ref_evecs, axis=0)) / (np.linalg.norm(evecs, axis=0) * np.linalg.norm(ref_evecs, axis=0))
    ang = np.arccos(np.clip(cos, -1.0, 1.0))
    return rel, ang
