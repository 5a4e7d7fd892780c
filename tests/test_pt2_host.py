"""CPU tests of the PT2 host logic: GMRES (gmres.rs restatement) against a
dense solve, and the Woodbury preconditioner against its exact inverse."""
import numpy as np
import torch

from nwick.pt2 import Preconditioner, general_evp, gmres


def test_gmres_matches_dense_solve():
    rng = np.random.default_rng(0)
    n = 120
    A = np.diag(np.linspace(1.0, 5.0, n)) + 0.05 * rng.normal(size=(n, n))
    b = rng.normal(size=n)
    At, bt = torch.tensor(A), torch.tensor(b)
    d = torch.tensor(1.0 / np.diag(A))
    x, res, it, conv = gmres(torch, lambda v: At @ v, lambda v: v * d, bt, res_tol=1e-12, restart=15, max_iter=400)
    assert conv and res <= 1e-12
    assert np.max(np.abs(x.numpy() - np.linalg.solve(A, b))) < 1e-9


def test_woodbury_preconditioner_is_exact_inverse_of_projected_diagonal():
    """With M = diag(d), the rank-2 Woodbury form inverts diag(d) + u0 v0^T + u1 v1^T exactly."""
    rng = np.random.default_rng(1)
    n = 30
    d = rng.uniform(1.0, 3.0, n)
    s = rng.normal(size=n) * 0.2
    f = rng.normal(size=n) * 0.2
    e0 = -1.3
    proj = {"e0": e0, "s_a0": torch.tensor(s), "s_0a": torch.tensor(s), "f_a0": torch.tensor(f),
            "f_0a": torch.tensor(f)}
    P = Preconditioner(torch, torch.tensor(d), proj, kind="woodbury")
    u0 = -f + 2 * e0 * s
    u1 = -s
    A = np.diag(d) + np.outer(u0, s) + np.outer(u1, f)
    v = rng.normal(size=n)
    assert np.max(np.abs(P(torch.tensor(v)).numpy() - np.linalg.solve(A, v))) < 1e-12


def test_general_evp_lowest_root():
    rng = np.random.default_rng(2)
    B = rng.normal(size=(8, 8))
    S = B @ B.T + 8 * np.eye(8)
    H = rng.normal(size=(8, 8)); H = H + H.T
    e, C = general_evp(H, S)
    import scipy.linalg
    assert abs(e[0] - scipy.linalg.eigh(H, S, eigvals_only=True)[0]) < 1e-12
    assert abs(C[:, 0] @ S @ C[:, 0] - 1.0) < 1e-12
