"""NOCI-PT2 on the GPU: projected shifted generalised-Fock operator and GMRES.

Mirrors the reference's PT2 step (src/snoci/step.rs:178-440):

  |Omega_a> = |Phi_a> - |Psi_0><Psi_0|Phi_a>
  M^Omega x = M x - f_a0 (s_0a . x) - s_a0 (f_0a . x) + 2 E0 s_a0 (s_0a . x)
              (apply_omega_m / apply_projection_correction, operators.rs:696-955)
  M_ab = F_ab - E0 S_ab, applied through spin-factorised per-spin tables
  (nw_factor_apply; apply_factorised_shifted_omega_m, operators.rs:762) or
  evaluated pairwise on the fly (nw_matvec_m; apply_candidate_m :375)
  solve M^Omega a = -V^Omega with right-preconditioned restarted GMRES
  (gmres.rs:406-544), E_PT2 = a.M^Omega a + a.V + V.a  (step.rs:414-429).

Everything that scales with the candidate count lives on the GPU in FP64:
the matvec (HIP kernels through the C ABI) and the Krylov vectors, Givens
rotations and the true residual (torch tensors on the device).  The NOCI
density and Fock matrices of the current space come from nw_noci_fock; the
host functions noci_density / fock below restate src/noci/naive.rs:231 and
src/scf/kernels.rs:46 for checking it.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .basis import to_records

SMALL = 1e-14


def general_evp(H, S, tol=1e-8):
    """Lowest-first generalised eigenpairs by canonical orthogonalisation,
    discarding overlap eigenvalues <= tol (src/maths/linalg.rs:239)."""
    w, U = np.linalg.eigh(S)
    keep = w > tol
    X = U[:, keep] / np.sqrt(w[keep])
    e, v = np.linalg.eigh(X.T @ H @ X)
    return e, X @ v


def _pair_density(Cl, Cg, S, tol):
    """Loewdin-paired transition density (naive.rs:73-229): returns
    (s_pair, rho) with rho = phase * s_red * (W or P) as pair_density."""
    Socc = Cl.T @ S @ Cg
    U, s, Vt = np.linalg.svd(Socc)
    V = Vt.T
    phase = np.linalg.det(U) * np.linalg.det(V)
    zeros = [k for k in range(len(s)) if abs(s[k]) <= tol]
    s_red = float(np.prod([x for x in s if abs(x) > tol])) if len(s) else 1.0
    lt, gt = Cl @ U, Cg @ V
    s_pair = phase * float(np.prod(s)) if not zeros else 0.0
    if len(zeros) == 0:
        w = np.array([1.0 / x if abs(x) > tol else 0.0 for x in s])
        rho = (gt * w) @ lt.T
    elif len(zeros) == 1:
        k = zeros[0]
        rho = np.outer(gt[:, k], lt[:, k])
    else:
        rho = np.zeros((Cl.shape[0], Cl.shape[0]))
    return s_pair, phase * s_red * rho


def noci_density(sysm, dets, c, tol=1e-8):
    """noci_density (src/noci/naive.rs:231-289) for a real NOCI state."""
    n = sysm.n_bas
    Da = np.zeros((n, n)); Db = np.zeros((n, n))

    def occ(d, sp):
        C = sysm.refs[d.parent][sp]
        o = d.oa if sp == 0 else d.ob
        return C[:, [i for i in range(sysm.nmo) if o >> i & 1]]

    for i, di in enumerate(dets):
        for j, dj in enumerate(dets):
            sa, ra = _pair_density(occ(di, 0), occ(dj, 0), sysm.S, tol)
            sb, rb = _pair_density(occ(di, 1), occ(dj, 1), sysm.S, tol)
            ph = (di.pha * dj.pha) * (di.phb * dj.phb)
            cij = c[i] * c[j] * ph
            Da += cij * sb * ra
            Db += cij * sa * rb
    return Da, Db


def fock(h, eri, Da, Db):
    """fock (src/scf/kernels.rs:46): F_sigma = h + J[Da + Db] - K[D_sigma]."""
    D = Da + Db
    J = np.einsum("pqrs,rs->pq", eri, D)
    Ka = np.einsum("prqs,rs->pq", eri, Da)
    Kb = np.einsum("prqs,rs->pq", eri, Db)
    Fa, Fb = h + J - Ka, h + J - Kb
    # the reference fills q <= p and mirrors
    return np.tril(Fa) + np.tril(Fa, -1).T, np.tril(Fb) + np.tril(Fb, -1).T


def omega_m_diag(m_diag, proj):
    """Diagonal of M^Omega from the unprojected diagonal of M
    (build_omega_m_diag, src/snoci/operators.rs:1032-1041)."""
    return (m_diag - proj["f_a0"] * proj["s_0a"] - proj["s_a0"] * proj["f_0a"]
            + 2.0 * proj["e0"] * proj["s_a0"] * proj["s_0a"])


def omega_s_diag(s_diag, proj):
    """Diagonal of Q = S - s_a0 s_0a^T (build_omega_s_diag, operators.rs:1049-1054)."""
    return s_diag - proj["s_a0"] * proj["s_0a"]


class Preconditioner:
    """build_preconditioner (src/snoci/operators.rs:954-985) with
    Preconditioner::new / apply (src/snoci/types.rs:161-279).

    m_diag, s_diag: unprojected diagonals of M and S.  kind "diag": the
    inverse of the diagonal of M^Omega + i eps Q.  Other kinds: the rank-2
    Woodbury inverse of diag(M + i eps S) plus the projection's two rank-1
    terms.  Diagonal entries below 1e-12 of the largest are regularised as
    conj(d) / floor^2.  Complex when imag_shift != 0."""

    def __init__(self, torch, m_diag, proj, kind="diag", imag_shift=0.0, s_diag=None):
        self.t = torch
        iq = 0.0
        if imag_shift != 0.0:
            if s_diag is None:
                raise ValueError("an imaginary shift needs the S diagonal")
            iq = 1j * imag_shift
            diag = (omega_m_diag(m_diag, proj) + iq * omega_s_diag(s_diag, proj) if kind == "diag"
                    else m_diag + iq * s_diag)
        else:
            diag = omega_m_diag(m_diag, proj) if kind == "diag" else m_diag
        dmax = float(diag.abs().max())
        dfloor = max(1e-12 * dmax, 1e-14)
        self.dinv = torch.where(diag.abs() > dfloor, 1.0 / diag, diag.conj() / (dfloor * dfloor))
        self.active = False
        if kind != "diag":
            dt = diag.dtype
            s_a0, f_a0 = proj["s_a0"].to(dt), proj["f_a0"].to(dt)
            u0 = -f_a0 + (2.0 * proj["e0"] - iq) * s_a0
            u1 = -s_a0
            self.v0, self.v1 = proj["s_0a"].to(dt), proj["f_0a"].to(dt)
            self.z0, self.z1 = self.dinv * u0, self.dinv * u1
            # unconjugated contractions (bilinear_dot)
            c00 = 1.0 + (self.v0 @ self.z0).item(); c01 = (self.v0 @ self.z1).item()
            c10 = (self.v1 @ self.z0).item(); c11 = 1.0 + (self.v1 @ self.z1).item()
            det = c00 * c11 - c01 * c10
            self.active = abs(det) > 1e-14
            if self.active:
                self.w = (c11 / det, -c01 / det, -c10 / det, c00 / det)

    def __call__(self, v):
        y = v * self.dinv
        if not self.active:
            return y
        # device-side contractions: no host synchronisation per application
        t0, t1 = self.v0 @ y, self.v1 @ y
        w00, w01, w10, w11 = self.w
        return y - self.z0 * (w00 * t0 + w01 * t1) - self.z1 * (w10 * t0 + w11 * t1)


def gmres(torch, apply, precondition, b, res_tol=1e-8, restart=30, max_iter=300, log=None, apply_true=None):
    """Right-preconditioned restarted GMRES (src/snoci/gmres.rs:406-544):
    Arnoldi with two orthogonalisation passes, Givens rotations, convergence
    on the RMS of the true residual ||b - A x|| / sqrt(n) recomputed at
    restarts.

    The Krylov basis stays in device memory as the rows of one matrix, and
    each orthogonalisation pass is classical Gram-Schmidt against all of it
    at once (two matrix-vector products; the second pass re-orthogonalises,
    as the reference's second modified pass does), so an Arnoldi step
    synchronises with the host once, to read its new Hessenberg column.

    apply_true: the operator of the true residual when it differs from the
    Arnoldi operator `apply` (a lower-precision copy of it): every restart is
    then an iterative-refinement step, b - A x in the higher precision and a
    Krylov correction built with the lower one.

    Real or complex b: inner products conjugate their first argument and the
    Givens rotations are the complex ones of the reference (cs real, sn =
    cs conj(y) / conj(x)); both reduce to the real forms for real b."""
    if apply_true is None:
        apply_true = apply
    n = b.shape[0]
    rms = math.sqrt(n)
    cplx = b.is_complex()
    dt = complex if cplx else float
    x = torch.zeros_like(b)
    rtrue = b.clone()  # b - A x for the zero start
    res = float(torch.linalg.vector_norm(rtrue)) / rms
    total = 0
    if res <= res_tol:
        return x, res, 0, True
    while total < max_iter:
        beta = float(torch.linalg.vector_norm(rtrue))
        if beta <= SMALL:
            return x, beta / rms, total, beta / rms <= res_tol
        inner = min(restart, max_iter - total)
        Q = torch.empty((inner + 1, n), dtype=b.dtype, device=b.device)
        Q[0] = rtrue / beta
        Hm = np.zeros((inner + 1, inner), dtype=dt)
        cs = np.zeros(inner); sn = np.zeros(inner, dtype=dt)
        g = np.zeros(inner + 1, dtype=dt); g[0] = beta
        kfinal = 0
        for k in range(inner):
            w = apply(precondition(Q[k]))
            Qk = Q[:k + 1]
            Qh = Qk.conj() if cplx else Qk
            h = torch.mv(Qh, w)
            w = w - torch.mv(Qk.T, h)
            h2 = torch.mv(Qh, w)
            w = w - torch.mv(Qk.T, h2)
            hnorm = torch.linalg.vector_norm(w)
            col = torch.cat([(h + h2), hnorm.reshape(1).to(h.dtype)]).cpu().numpy()  # the step's one sync
            hn = float(abs(col[k + 1]))
            Hm[:k + 1, k] = col[:k + 1]
            Hm[k + 1, k] = hn
            if hn > SMALL:
                Q[k + 1] = w / hn
            for j in range(k):
                tmp = cs[j] * Hm[j, k] + sn[j] * Hm[j + 1, k]
                Hm[j + 1, k] = -np.conj(sn[j]) * Hm[j, k] + cs[j] * Hm[j + 1, k]
                Hm[j, k] = tmp
            xk, yk = Hm[k, k], Hm[k + 1, k]
            ax, ay = abs(xk), abs(yk)
            den = math.sqrt(ax * ax + ay * ay)
            if den > SMALL:
                cs[k] = ax / den
                sn[k] = cs[k] * np.conj(yk) / np.conj(xk) if ax > SMALL else 1.0
            else:
                cs[k], sn[k] = 1.0, 0.0
            h0, h1 = Hm[k, k], Hm[k + 1, k]
            Hm[k, k] = cs[k] * h0 + sn[k] * h1
            Hm[k + 1, k] = -np.conj(sn[k]) * h0 + cs[k] * h1
            g0, g1 = g[k], g[k + 1]
            g[k] = cs[k] * g0 + sn[k] * g1
            g[k + 1] = -np.conj(sn[k]) * g0 + cs[k] * g1
            kfinal = k + 1
            est = abs(g[k + 1]) / rms
            if log:
                log(f"  gmres iter {total + k + 1}: residual estimate {est:.3e}")
            if est <= res_tol or hn <= SMALL:
                break
        y = np.zeros(kfinal, dtype=dt)
        for i in reversed(range(kfinal)):
            rhs = g[i] - sum(Hm[i, j] * y[j] for j in range(i + 1, kfinal))
            y[i] = rhs / Hm[i, i] if abs(Hm[i, i]) > SMALL else 0.0
        # x += M^-1 (Q y): the preconditioner is linear, so it is applied once
        yd = torch.as_tensor(y, dtype=b.dtype, device=b.device)
        x = x + precondition(torch.mv(Q[:kfinal].T, yd))
        total += kfinal
        rtrue = b - apply_true(x)
        res = float(torch.linalg.vector_norm(rtrue)) / rms
        if log:
            log(f"  gmres restart after {total} iterations: true residual {res:.3e}")
        if res <= res_tol:
            return x, res, total, True
    return x, res, total, res <= res_tol


class PT2:
    """NOCI-PT2 correction of a current NOCI space over a candidate space.

    eng must hold a build covering every ordered reference pair the
    candidate-current and candidate-candidate pairs touch."""

    def __init__(self, eng, sysm, current, candidates, tol=1e-8, mode=0, device="cuda:0", factor_tables="device",
                 precision="fp64"):
        """factor_tables: "device" (tables resident in device memory, the
        reference's default snoci.gmres.factor_tables = RAM with full_m = None,
        src/input/snoci.rs:25-33), "transient" (regenerated per matvec) or None
        (pairwise on-the-fly matvec, nw_matvec_m).  Candidate lists whose
        tables do not fit in device memory (NW_E_MEM), or whose tables would
        need per-spin contraction ranks above 4 (NW_E_RANK: triples against
        doubles), use the pairwise matvec, which covers ranks up to 8.

        precision: "fp64" (the reference's arithmetic) or "mixed": GMRES runs
        on the factorised operator with FP32 tables (FP64 sums, half the table
        bytes per matvec) inside FP64 iterative refinement, whose residuals
        and E_PT2 use the FP64 tables (see solve).  "mixed" needs the
        factorised operator."""
        import torch
        self.torch, self.eng, self.sysm = torch, eng, sysm
        self.current, self.candidates = list(current), list(candidates)
        self.tol, self.mode = tol, mode
        self.dev = torch.device(device)
        self.timings = {}
        self.factor_tables = factor_tables
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.fac = None
        self.nmatvec = self.nmatvec_lo = 0

    def setup(self):
        torch, eng = self.torch, self.eng
        t0 = time.perf_counter()
        cur_rec = to_records(self.current)
        Hc, Sc = eng.hs_matrix(cur_rec, mode=self.mode)
        e, C = general_evp(Hc, Sc, self.tol)
        self.ecurrent, c = float(e[0]), C[:, 0]
        self.coeffs = c
        # NOCI density and Fock on the GPU (noci_density + fock), then the slab's Fock intermediates
        fa, fb = eng.noci_fock(cur_rec, c, self.tol)
        eng.update_fock(fa, fb)
        Fc, _ = eng.fs_matrix(cur_rec, mode=self.mode)
        self.e0 = float(c @ Fc @ c)
        # candidate-current blocks through one combined determinant list
        allrec = to_records(self.current + self.candidates)
        ncur, nc = len(self.current), len(self.candidates)
        a_idx = np.repeat(np.arange(ncur, ncur + nc), ncur)
        i_idx = np.tile(np.arange(ncur), nc)
        req = np.stack([a_idx, i_idx], axis=1).astype(np.uint64)
        f_ai, s_ai = eng.fs_batched(allrec, req, mode=self.mode)
        h_ai, _ = eng.hs_batched(allrec, req, mode=self.mode)
        F_ai, S_ai, H_ai = (v.reshape(nc, ncur) for v in (f_ai, s_ai, h_ai))
        diag = np.stack([np.arange(ncur, ncur + nc)] * 2, axis=1).astype(np.uint64)
        f_aa, s_aa = eng.fs_batched(allrec, diag, mode=self.mode)
        T = lambda v: torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64, device=self.dev)
        s_a0, f_a0 = S_ai @ c, F_ai @ c
        self.proj = {"e0": self.e0, "s_a0": T(s_a0), "s_0a": T(s_a0), "f_a0": T(f_a0), "f_0a": T(f_a0)}
        self.v_omega = T(H_ai @ c - self.ecurrent * s_a0)
        self.cand_rec = torch.from_numpy(to_records(self.candidates).view(np.uint8).copy()).to(self.dev)
        self.y = torch.empty(nc, dtype=torch.float64, device=self.dev)
        self.fac = None
        if self.factor_tables is not None and nc > 0:
            from . import FACTOR_DEVICE, FACTOR_TRANSIENT, NwickError
            storage = FACTOR_DEVICE if self.factor_tables == "device" else FACTOR_TRANSIENT
            t1 = time.perf_counter()
            try:
                self.fac = eng.factor(self.cand_rec, mode=self.mode, storage=storage)
            except NwickError as e:
                if e.code not in (4, 5):  # NW_E_MEM, NW_E_RANK: pairwise matvec instead
                    raise
                self.fac = None
            self.timings["factor_s"] = time.perf_counter() - t1
        if self.fac is not None:
            # build_factorised_candidate_diags (src/snoci/operators.rs:354-365)
            md = torch.empty(nc, dtype=torch.float64, device=self.dev)
            sd = torch.empty_like(md)
            self.fac.diagonals(-self.e0, mdiag=md, sdiag=sd)
            self.m_diag, self.s_diag = md, sd
        else:
            self.m_diag, self.s_diag = T(f_aa - self.e0 * s_aa), T(s_aa)
        self.operator = "factorised" if self.fac is not None else "pairwise"
        if self.precision == "mixed":
            if self.fac is None:
                raise RuntimeError("mixed-precision PT2 needs the factorised operator")
            from . import PREC_FP64, PREC_MIXED
            self.fac.set_precision(PREC_MIXED)  # FP32 table copies
            self.fac.set_precision(PREC_FP64)
        self.timings["setup_s"] = time.perf_counter() - t0

    def _m(self, x, s_out=None):
        """(F - E0 S) x for a real vector, and S x into s_out if given."""
        y = self.torch.empty_like(self.y)
        if self.fac is not None:
            self.fac.apply(-self.e0, x, y=y, ys=s_out)
        else:
            self.eng.matvec_m(self.cand_rec, self.e0, x, y=y, ys=s_out, mode=self.mode)
        return y

    def apply(self, x, imag_shift=0.0):
        """(M^Omega + i eps Q) x: apply_factorised_shifted_omega_m
        (src/snoci/operators.rs:759-776) or, pairwise, apply_shifted_omega_m
        (:856-887), with the projection correction (:938-947).  A complex x
        takes both parts through one pass over the factor tables
        (nw_factor_apply_pair), or one pairwise matvec per part; eps != 0
        also uses S x."""
        torch = self.torch
        p = self.proj
        shift = imag_shift != 0.0
        if x.is_complex() and self.fac is not None:
            # both parts in one pass over the tables
            x = x.contiguous()
            y = torch.empty_like(x)
            svec = torch.empty_like(x) if shift else None
            self.fac.apply_pair(-self.e0, x, y=y, ys=svec)
        elif x.is_complex():
            parts = []
            for comp in (x.real, x.imag):
                sv = torch.empty_like(self.y) if shift else None
                parts.append((self._m(comp.contiguous(), sv), sv))
            y = torch.complex(parts[0][0], parts[1][0])
            svec = torch.complex(parts[0][1], parts[1][1]) if shift else None
        else:
            x = x.contiguous()
            svec = torch.empty_like(self.y) if shift else None
            y = self._m(x, svec)
        self.nmatvec += 1
        if self.fac is not None and self.fac.precision != 0:
            self.nmatvec_lo += 1
        if not x.is_complex() and not shift:
            sx, fx = float(p["s_0a"] @ x), float(p["f_0a"] @ x)
            return y - p["f_a0"] * sx - p["s_a0"] * fx + 2.0 * self.e0 * p["s_a0"] * sx
        dt = torch.complex128
        if getattr(self, "_proj_c", None) is None:
            self._proj_c = {k: p[k].to(dt) for k in ("s_a0", "f_a0", "s_0a", "f_0a")}
        pc = self._proj_c
        xc, y = x.to(dt), y.to(dt)
        sx, fx = (pc["s_0a"] @ xc).item(), (pc["f_0a"] @ xc).item()
        y = y - pc["f_a0"] * sx - pc["s_a0"] * fx + (2.0 * self.e0 - 1j * imag_shift) * pc["s_a0"] * sx
        if shift:
            y = y + (1j * imag_shift) * svec.to(dt)
        return y

    def solve(self, res_tol=1e-8, restart=30, max_iter=300, preconditioner="diag", log=None, imag_shift=0.0):
        """Solve (M^Omega + i eps Q) a = -V^Omega to an RMS residual of
        res_tol with restarted GMRES (gmres.rs:406-544) and return E_PT2 =
        Re(a^H (M^Omega + i eps Q) a + a^H V + V^H a) (step.rs:366-429); eps
        = imag_shift (snoci imag_shifts), complex a when eps != 0.
        precision "mixed": the Arnoldi matvecs read the FP32 tables, the true
        residual at each restart (the convergence test) and E_PT2 use the
        FP64 tables."""
        torch = self.torch
        self.nmatvec = self.nmatvec_lo = 0
        prec = Preconditioner(torch, self.m_diag, self.proj, preconditioner, imag_shift, self.s_diag)
        b = -self.v_omega
        if imag_shift != 0.0:
            b = b.to(torch.complex128)
        op = lambda v: self.apply(v, imag_shift)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if self.precision == "fp64":
            a, res, iters, conv = gmres(torch, op, prec, b, res_tol, restart, max_iter, log)
        else:
            from . import PREC_FP64, PREC_MIXED

            def apply_fp32(v):
                self.fac.set_precision(PREC_MIXED)
                try:
                    return self.apply(v, imag_shift)
                finally:
                    self.fac.set_precision(PREC_FP64)

            a, res, iters, conv = gmres(torch, apply_fp32, prec, b, res_tol, restart, max_iter, log,
                                        apply_true=op)
        ma = op(a)
        torch.cuda.synchronize()
        if a.is_complex():
            v = self.v_omega.to(a.dtype)
            ept2 = float((torch.vdot(a, ma) + torch.vdot(a, v) + torch.vdot(v, a)).real)
        else:
            ept2 = float(a @ ma) + 2.0 * float(a @ self.v_omega)
        self.timings["gmres_s"] = time.perf_counter() - t0
        return {"e_noci": self.ecurrent, "e0": self.e0, "ept2": ept2, "e_total": self.ecurrent + ept2,
                "residual_rms": res, "iterations": iters, "converged": conv, "matvecs": self.nmatvec,
                "matvecs_fp32": self.nmatvec_lo, "precision": self.precision, "imag_shift": imag_shift,
                "a": a}

