"""Dense CPU restatement of one NOCI-PT2 step (test infrastructure only).

Follows src/snoci/step.rs:198-429 with every matrix formed explicitly from the
oracle's H, S and generalised-Fock elements: current-space GEVP, NOCI density
(src/noci/naive.rs:231) and spin Fock matrices (src/scf/kernels.rs:46),
E0 = c^T F c, the projected operator of src/snoci/operators.rs:696-955, the
projected right-hand side (build_omega_v), and E_PT2 = a.M a + a.V + V.a.
The linear system is solved by least squares so linearly dependent candidate
spaces (whose E_PT2 is still unique) are handled.
"""
import numpy as np

import oracle as O  # oracle/oracle.py (the oracle directory is on sys.path in tests)


def candidate_left(slab, cands, current, mode):
    """F_ai, S_ai, H_ai with the candidate as bra: the reference evaluates the
    candidate-current blocks candidate-left (build_noci_hs / build_noci_fock
    with candidates first, src/snoci/operators.rs:128-159) and takes the
    current-candidate blocks as their adjoints (:113, :138).  The orientation
    matters for reference pairs with zero-overlap orbitals (m > 0) in the
    reference arithmetic (mode 0), whose elements are not symmetric."""
    na, ni = len(cands), len(current)
    F, S, H = np.empty((na, ni)), np.empty((na, ni)), np.empty((na, ni))
    for a, da in enumerate(cands):
        for i, di in enumerate(current):
            F[a, i], S[a, i] = slab.fs(da, di, mode)
            H[a, i] = slab.hs(da, di, mode)[0]
    return F, S, H


def dense_pt2(sysm, current, cands, mode=0, tol=1e-12, evp_tol=1e-8):
    from nwick.pt2 import fock, general_evp, noci_density
    slab = O.Slab(sysm, tol)
    Hc, Sc = O.hs_matrix(slab, current, mode=mode)
    e, C = general_evp(Hc, Sc, evp_tol)
    c = C[:, 0]
    fa, fb = fock(sysm.h, sysm.eri, *noci_density(sysm, current, c, tol))
    slab.update_fock(sysm, fa, fb)
    Fc, _ = O.fs_matrix(slab, current, mode=mode)
    e0 = float(c @ Fc @ c)
    F_ai, S_ai, H_ai = candidate_left(slab, cands, current, mode)
    Faa, Saa = O.fs_matrix(slab, list(cands), mode=mode)
    s0, f0 = S_ai @ c, F_ai @ c
    M = Faa - e0 * Saa - np.outer(f0, s0) - np.outer(s0, f0) + 2 * e0 * np.outer(s0, s0)
    v = H_ai @ c - e[0] * s0
    a = np.linalg.lstsq(M, -v, rcond=1e-12)[0]
    ept2 = float(a @ M @ a + 2 * a @ v)
    return {"e_noci": float(e[0]), "e0": e0, "ept2": ept2, "e_total": float(e[0]) + ept2, "a": a, "v": v,
            "residual": float(np.linalg.norm(M @ a + v))}


class DensePT2Backend:
    """PT2 step for the SNOCI driver from dense oracle matrices, solved with
    the same right-preconditioned GMRES the GPU path uses (on CPU tensors),
    so singular (linearly dependent) candidate spaces give the same
    amplitudes as the GPU solve."""

    def __init__(self, sysm, tol=1e-12, mode=0):
        self.sysm, self.tol, self.mode = sysm, tol, mode
        self.slab = O.Slab(sysm, tol)

    def __call__(self, current, cands, opts):
        import torch
        from nwick.pt2 import Preconditioner, fock, general_evp, gmres, noci_density
        slab, sysm = self.slab, self.sysm
        Hc, Sc = O.hs_matrix(slab, current, mode=self.mode)
        e, C = general_evp(Hc, Sc, self.tol)
        c = C[:, 0]
        if not cands:
            return {"ecurrent": float(e[0]), "coeffs": c}
        fa, fb = fock(sysm.h, sysm.eri, *noci_density(sysm, current, c, self.tol))
        slab.update_fock(sysm, fa, fb)
        Fc, _ = O.fs_matrix(slab, current, mode=self.mode)
        e0 = float(c @ Fc @ c)
        F_ai, S_ai, H_ai = candidate_left(slab, cands, current, self.mode)
        Faa, Saa = O.fs_matrix(slab, list(cands), mode=self.mode)
        s0, f0 = S_ai @ c, F_ai @ c
        Mraw = Faa - e0 * Saa
        M = Mraw - np.outer(f0, s0) - np.outer(s0, f0) + 2 * e0 * np.outer(s0, s0)
        v = H_ai @ c - e[0] * s0
        T = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64)
        proj = {"e0": e0, "s_a0": T(s0), "s_0a": T(s0), "f_a0": T(f0), "f_0a": T(f0)}
        prec = Preconditioner(torch, T(np.diag(Mraw)), proj, opts.preconditioner)
        Mt = T(M)
        a, res, iters, conv = gmres(torch, lambda x: Mt @ x, prec, -T(v), opts.res_tol, opts.restart,
                                    opts.gmres_max_iter)
        a = a.numpy()
        ept2 = float(a @ M @ a + 2 * a @ v)
        return {"ecurrent": float(e[0]), "coeffs": c, "ept2": ept2, "a": a, "v": v, "converged": conv,
                "residual": res}
