"""CPU tests of the PT2 host logic: GMRES (gmres.rs restatement) against a
dense solve, real and complex (imaginary shift), with a lower-precision
Arnoldi operator refined at restarts; the preconditioners against the
inverses they stand for (projected diagonal; rank-2 Woodbury, with and
without an imaginary shift)."""
import numpy as np
import torch

from nwick.pt2 import Preconditioner, general_evp, gmres, omega_m_diag, omega_s_diag


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


def test_gmres_complex_matches_dense_solve():
    """A + i eps Q with A symmetric indefinite: the complex Givens path."""
    rng = np.random.default_rng(5)
    n = 90
    A = np.diag(np.linspace(-2.0, 4.0, n)) + 0.05 * rng.normal(size=(n, n))
    A = 0.5 * (A + A.T)
    Q = np.eye(n) + 0.02 * rng.normal(size=(n, n))
    Q = 0.5 * (Q + Q.T)
    C = A + 0.7j * Q
    b = rng.normal(size=n) + 0j
    Ct, bt = torch.tensor(C), torch.tensor(b)
    d = torch.tensor(1.0 / np.diag(C))
    x, res, it, conv = gmres(torch, lambda v: Ct @ v, lambda v: v * d, bt, res_tol=1e-12, restart=20, max_iter=2000)
    assert conv and res <= 1e-12
    assert np.max(np.abs(x.numpy() - np.linalg.solve(C, b))) < 1e-9


def test_gmres_low_precision_arnoldi_with_refined_restarts():
    """Arnoldi on a float32 copy of A, true residuals with A: converges to
    the float64 solution, beyond what float32 alone resolves."""
    rng = np.random.default_rng(6)
    n = 100
    A = np.diag(np.linspace(1.0, 3.0, n)) + 0.05 * rng.normal(size=(n, n))
    b = rng.normal(size=n)
    At, bt = torch.tensor(A), torch.tensor(b)
    A32 = At.float()
    d = torch.tensor(1.0 / np.diag(A))
    lo = lambda v: (A32 @ v.float()).double()
    x, res, it, conv = gmres(torch, lo, lambda v: v * d, bt, res_tol=1e-12, restart=15, max_iter=400,
                             apply_true=lambda v: At @ v)
    assert conv and res <= 1e-12
    assert np.max(np.abs(x.numpy() - np.linalg.solve(A, b))) < 1e-10


def _proj(rng, n, e0=-1.3):
    s = rng.normal(size=n) * 0.2
    f = rng.normal(size=n) * 0.2
    return s, f, {"e0": e0, "s_a0": torch.tensor(s), "s_0a": torch.tensor(s), "f_a0": torch.tensor(f),
                  "f_0a": torch.tensor(f)}


def test_diag_preconditioner_inverts_projected_diagonal():
    """kind "diag" divides by the diagonal of M^Omega (+ i eps Q), not of M."""
    rng = np.random.default_rng(3)
    n = 25
    d = rng.uniform(1.0, 3.0, n)
    sd = rng.uniform(0.5, 1.5, n)
    s, f, proj = _proj(rng, n)
    e0 = proj["e0"]
    om = d - 2 * f * s + 2 * e0 * s * s
    v = rng.normal(size=n)
    P = Preconditioner(torch, torch.tensor(d), proj, kind="diag")
    assert np.max(np.abs(P(torch.tensor(v)).numpy() - v / om)) < 1e-14
    eps = 0.4
    Pc = Preconditioner(torch, torch.tensor(d), proj, kind="diag", imag_shift=eps, s_diag=torch.tensor(sd))
    want = v / (om + 1j * eps * (sd - s * s))
    assert np.max(np.abs(Pc(torch.tensor(v + 0j)).numpy() - want)) < 1e-14
    assert np.allclose(omega_m_diag(torch.tensor(d), proj).numpy(), om, atol=1e-15)
    assert np.allclose(omega_s_diag(torch.tensor(sd), proj).numpy(), sd - s * s, atol=1e-15)


def test_woodbury_preconditioner_with_imaginary_shift():
    """Inverts diag(d + i eps sd) + u0 s^T + u1 f^T with u0 = -f + (2 E0 - i eps) s, u1 = -s."""
    rng = np.random.default_rng(4)
    n = 30
    d = rng.uniform(1.0, 3.0, n)
    sd = rng.uniform(0.5, 1.5, n)
    s, f, proj = _proj(rng, n)
    eps = 0.25
    P = Preconditioner(torch, torch.tensor(d), proj, kind="woodbury", imag_shift=eps, s_diag=torch.tensor(sd))
    u0 = -f + (2 * proj["e0"] - 1j * eps) * s
    A = np.diag(d + 1j * eps * sd) + np.outer(u0, s) + np.outer(-s, f)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.max(np.abs(P(torch.tensor(v)).numpy() - np.linalg.solve(A, v))) < 1e-12
