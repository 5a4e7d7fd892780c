"""Closed forms of what each literal reading of the reference's m > 0
evaluator changes in one Hamiltonian element (test infrastructure).

The reference's m > 0 path (src/nonorthogonalwicks/eval/
preparehamiltonianoverlap.rs:20550-20966 with eval/helpers.rs) differs from
the exact extended-Wick element in five places.  Each function below writes,
from the Rust text and with numpy determinants and cofactors taken from
explicit minors, the change in H that one reading makes when it is the only
one applied, reading the pair's intermediates from an exported slab block
(the layout of src/nonorthogonalwicks/layout.rs:21-189):

- v0:    the (m1, m2) enumeration adds v0[m1 + m2] on both one-assignment
         visits (:20637) while v0[1] already holds both (build.rs:229-232);
- vidx:  the one-column term reads v[m1][m2][mk] (:20648-20651) where the
         stored order is v[left][right][JK] (build.rs:242-259);
- jslot: the two-column term takes jslot(m1, m2, mk, mj) (:20690);
- dist:  distributions whose mixed determinant is at most tol in magnitude
         are skipped (adjugate_transpose_generic, helpers.rs:737-748);
- minor: first minors at most tol in magnitude are skipped (minor_adjt,
         helpers.rs:636-660);
- enuc:  E_nuc S is not added (the `_enuc` argument, :20555).

Every change is H(reading) - H(corrected); the corrected element is pinned
separately by brute force.  `element` restates the whole m > 0 element with
any subset of the readings, so that their interactions (a skipped first
minor carries the J slot of whichever index order is in force) are checked
too.
"""
from __future__ import annotations

import numpy as np

# slab offsets (per spin at 0 and 31) and scalar meta (per spin at 0 and 11)
O_X, O_Y, O_FH, O_FF, O_V, O_J = 0, 2, 4, 8, 12, 20
O_BB, O_VAB, O_VBA, O_II = 31, 62, 70, 78
M_TSP, M_PH, M_M, M_NOCC, M_F0H, M_V0, M_BB, M_VAB0 = 0, 1, 2, 3, 4, 8, 11, 22

RD_V0, RD_VIDX, RD_JSLOT, RD_DIST, RD_MINOR, RD_ENUC = 1, 2, 4, 8, 16, 32

# jslot, helpers.rs:165-193: (m_i, m_j, m_k, m_l) -> (slot, swap)
JSLOT = {(0, 0, 0, 0): (0, False), (0, 0, 0, 1): (1, False), (0, 1, 0, 0): (1, True),
         (0, 0, 1, 0): (2, False), (1, 0, 0, 0): (2, True), (0, 0, 1, 1): (3, False),
         (1, 1, 0, 0): (3, True), (0, 1, 0, 1): (4, False), (0, 1, 1, 0): (5, False),
         (1, 0, 0, 1): (5, True), (0, 1, 1, 1): (6, False), (1, 1, 0, 1): (6, True),
         (1, 0, 1, 0): (7, False), (1, 0, 1, 1): (8, False), (1, 1, 1, 0): (8, True),
         (1, 1, 1, 1): (9, False)}


def bits_of(v: int):
    return [i for i in range(v.bit_length()) if v >> i & 1]


def det(A):
    return 1.0 if A.shape[0] == 0 else float(np.linalg.det(A))


def cof(A):
    """Cofactor matrix from explicit minors: cof[r, c] = (-1)^(r+c) det A[r|c]."""
    L = A.shape[0]
    C = np.ones((L, L))
    if L > 1:
        for r in range(L):
            for c in range(L):
                C[r, c] = (-1) ** (r + c) * det(np.delete(np.delete(A, r, 0), c, 1))
    return C


class Spin:
    """One spin's intermediates of an ordered reference pair and one
    determinant pair's contraction labels and endpoint determinants."""

    def __init__(self, sub, off, meta, n, sp, bra_h, bra_p, ket_h, ket_p):
        o = off[O_BB:] if sp else off
        mt = meta[M_BB:] if sp else meta

        def m2(k):
            return sub[k:k + n * n].reshape(n, n)

        def m4(k):
            return None if k < 0 else sub[k:k + n ** 4].reshape(n, n, n, n)

        self.n = n
        self.m = int(mt[M_M])
        self.nocc = int(mt[M_NOCC])
        self.pref = mt[M_PH] * mt[M_TSP]
        self.f0h = mt[M_F0H:M_F0H + 2]
        self.v0 = mt[M_V0:M_V0 + 3]
        self.X = [m2(o[O_X]), m2(o[O_X + 1])]
        self.Y = [m2(o[O_Y]), m2(o[O_Y + 1])]
        self.FH = {(a, b): m2(o[O_FH + a * 2 + b]) for a in (0, 1) for b in (0, 1)}
        self.V = {(a, b, c): m2(o[O_V + (a * 2 + b) * 2 + c]) for a in (0, 1) for b in (0, 1) for c in (0, 1)}
        self.J = [m4(o[O_J + s]) for s in range(10)]
        nv = n - self.nocc
        # construct_determinant_indices (eval/prepare.rs:555-595)
        self.rows = [p - self.nocc for p in bra_p] + [nv + h for h in ket_h]
        self.cols = list(bra_h) + list(ket_p)
        self.L = len(self.rows)
        # build_d (src/maths/wick.rs:59): X on and below the diagonal, Y above
        self.D = []
        for mm in (0, 1):
            D = np.empty((self.L, self.L))
            for i, r in enumerate(self.rows):
                for j, c in enumerate(self.cols):
                    D[i, j] = (self.X[mm] if i >= j else self.Y[mm])[r, c]
            self.D.append(D)

    def mix(self, cbits):
        M = self.D[0].copy()
        for c in range(self.L):
            if cbits >> c & 1:
                M[:, c] = self.D[1][:, c]
        return M

    def dists(self, extra):
        """Assignments over `extra` operator bits then L column bits, m set bits."""
        return [b for b in range(1 << (self.L + extra)) if bin(b).count("1") == self.m]

    def s(self):
        return sum(det(self.mix(c)) for c in range(1 << self.L) if bin(c).count("1") == self.m)

    # --- changes of the same-spin two-body sum h2 ---------------------------
    def d_v0(self):
        return sum(0.5 * self.v0[1] * det(self.mix(b >> 2)) for b in self.dists(2)
                   if (b & 1) + (b >> 1 & 1) == 1)

    def d_vidx(self):
        tot = 0.0
        for b in self.dists(2):
            m1, m2, cb = b & 1, b >> 1 & 1, b >> 2
            C = cof(self.mix(cb))
            for k in range(self.L):
                mk = cb >> k & 1
                diff = self.V[(m1, m2, mk)] - self.V[(m1, mk, m2)]
                tot += -2.0 * sum(diff[self.cols[k], self.rows[r]] * C[r, k] for r in range(self.L))
        return tot

    def jval(self, key, i, j, rf, kf):
        slot, swap = JSLOT[key]
        Jt = self.J[slot]
        ri, cj, rr, cc = self.rows[i], self.cols[j], self.rows[rf], self.cols[kf]
        return Jt[rr, cc, ri, cj] if swap else Jt[ri, cj, rr, cc]

    def d_jslot(self):
        tot = 0.0
        L = self.L
        for b in self.dists(2):
            m1, m2, cb = b & 1, b >> 1 & 1, b >> 2
            M = self.mix(cb)
            for i in range(L):
                for j in range(L):
                    mj = cb >> j & 1
                    Cm = cof(np.delete(np.delete(M, i, 0), j, 1))
                    rfs = [r for r in range(L) if r != i]
                    kfs = [k for k in range(L) if k != j]
                    for k2, kf in enumerate(kfs):
                        mk = cb >> kf & 1
                        for r2, rf in enumerate(rfs):
                            lit = self.jval((m1, m2, mk, mj), i, j, rf, kf)
                            cor = self.jval((m1, mk, m2, mj), i, j, rf, kf)
                            tot += (-1) ** (i + j) * (lit - cor) * Cm[r2, k2]
        return tot


def spins(slab_view, n, bra, ket):
    sub, off, meta = slab_view
    return (Spin(sub, off, meta, n, 0, bits_of(bra.ha), bits_of(bra.pa), bits_of(ket.ha), bits_of(ket.pa)),
            Spin(sub, off, meta, n, 1, bits_of(bra.hb), bits_of(bra.pb), bits_of(ket.hb), bits_of(ket.pb)))


def prefactor(a: Spin, b: Spin, bra, ket):
    return (bra.pha * ket.pha * bra.phb * ket.phb) * a.pref * b.pref


def same_spin_change(slab_view, n, bra, ket, which):
    """H(reading) - H(corrected) for reading v0, vidx or jslot: the change
    of h2aa and h2bb, each entering H as pref * h2 * s_other / 2."""
    a, b = spins(slab_view, n, bra, ket)
    fn = {"v0": Spin.d_v0, "vidx": Spin.d_vidx, "jslot": Spin.d_jslot}[which]
    return prefactor(a, b, bra, ket) * 0.5 * (fn(a) * b.s() + fn(b) * a.s())


def minor_drop_change_l2(slab_view, n, bra, ket):
    """Reading `minor` on an alpha contraction of rank 2 whose D0[1, 0]
    (X0[rows[1], cols[0]]) is zero: every distribution that takes column 0
    from D0 has the 1x1 first minor D[0|1] = D0[1, 0] = 0, which the
    reference skips.  That minor's term is phase (-1)^(0+1) times the
    replacement determinant J(rows[0], cols[1]; rows[1], cols[0]) (cofactor 1
    of a 1x1 minor), so skipping it adds +J per such distribution."""
    a, b = spins(slab_view, n, bra, ket)
    assert a.L == 2 and a.D[0][1, 0] == 0.0
    tot = 0.0
    for bits in a.dists(2):
        m1, m2, cb = bits & 1, bits >> 1 & 1, bits >> 2
        if cb & 1:
            continue
        mk, mj = cb & 1, cb >> 1 & 1
        tot += a.jval((m1, mk, m2, mj), 0, 1, 1, 0)
    return prefactor(a, b, bra, ket) * 0.5 * tot * b.s()


def dist_drop_change_l1(slab_view, n, bra, ket):
    """Reading `dist` on an alpha contraction of rank 1 with m_alpha = 1,
    m_beta = 0 and D0[0, 0] = X0[rows[0], cols[0]] zero.  The same-spin
    distributions (m1, m2) = (1, 0) and (0, 1) have det D_mix = D0[0, 0] = 0
    and cofactor 1, so their terms are the pure replacements
      h1: -FH[1][0][c0, r0];   h2: -2 V[1][0][0][c0, r0] and -2 V[0][0][1][c0, r0];
    the different-spin distribution with m_alpha0 = 1 likewise contributes
      -VAB[1][0][0][c0, r0] det D_beta + sum_{r,k} II[1][0][0][0][r0, c0, rb_r, cb_k] cof_beta[r, k].
    The reference skips all of them; the overlap keeps its (nonzero) terms."""
    a, b = spins(slab_view, n, bra, ket)
    assert a.L == 1 and a.m == 1 and b.m == 0 and a.D[0][0, 0] == 0.0
    sub, off, meta = slab_view
    r0, c0 = a.rows[0], a.cols[0]
    h1 = -a.FH[(1, 0)][c0, r0]
    h2 = -2.0 * a.V[(1, 0, 0)][c0, r0] - 2.0 * a.V[(0, 0, 1)][c0, r0]
    nn = n * n
    vab = sub[off[O_VAB + 4]:off[O_VAB + 4] + nn].reshape(n, n)  # vab[ma0=1][mb0=0][mak=0]
    ii = sub[off[O_II + 8]:off[O_II + 8] + nn * nn].reshape(n, n, n, n)  # ii[1][0][0][0]
    Db = b.D[0]
    h2ab = -vab[c0, r0] * det(Db)
    if b.L:
        Cb = cof(Db)
        h2ab += sum(ii[r0, c0, b.rows[r], b.cols[k]] * Cb[r, k] for r in range(b.L) for k in range(b.L))
    sb = b.s()
    return -prefactor(a, b, bra, ket) * (h1 * sb + 0.5 * h2 * sb + h2ab)


def minor_skip_change(slab_view, n, bra, ket, tol):
    """Reading `minor` alone on any input: every (distribution, i, j) whose
    first minor D_mix[i|j] has |det| <= tol loses its two-column term
    (-1)^(i+j) sum_{k2,r} J(...) cof(D_mix[i|j])[r, k2] (stored index order)."""
    a, b = spins(slab_view, n, bra, ket)
    out = 0.0
    for sp, (x, other) in enumerate(((a, b), (b, a))):
        tot = 0.0
        L = x.L
        for bits in (x.dists(2) if L >= 2 else []):
            m1, m2, cb = bits & 1, bits >> 1 & 1, bits >> 2
            M = x.mix(cb)
            for i in range(L):
                for j in range(L):
                    Mn = np.delete(np.delete(M, i, 0), j, 1)
                    if abs(det(Mn)) > tol:
                        continue
                    Cm = cof(Mn)
                    mj = cb >> j & 1
                    kfs = [k for k in range(L) if k != j]
                    rfs = [r for r in range(L) if r != i]
                    for k2, kf in enumerate(kfs):
                        mk = cb >> kf & 1
                        for r2, rf in enumerate(rfs):
                            tot -= (-1) ** (i + j) * x.jval((m1, mk, m2, mj), i, j, rf, kf) * Cm[r2, k2]
        out += 0.5 * tot * other.s()
    return prefactor(a, b, bra, ket) * out


def _same_sum(x: Spin, rd, tol):
    """Same-spin sum over (m1, m2, column) assignments, pho.rs:20577-20741."""
    s = h1 = h2 = 0.0
    L = x.L
    for bits in x.dists(2):
        m1, m2, cb = bits & 1, bits >> 1 & 1, bits >> 2
        M = x.mix(cb)
        dd = det(M)
        if (rd & RD_DIST) and L > 0 and abs(dd) <= tol:
            if m1 == 0 and m2 == 0:
                s += dd
            continue
        C = cof(M)
        if m1 == 0 and m2 == 0:
            s += dd
        if m2 == 0:
            c = dd * x.f0h[m1]
            for k in range(L):
                F = x.FH[(m1, cb >> k & 1)]
                c -= dd + sum((F[x.cols[k], x.rows[r]] - M[r, k]) * C[r, k] for r in range(L))
            h1 += c
        v0 = x.v0[m1 + m2] * (0.5 if (m1 + m2 == 1 and not rd & RD_V0) else 1.0)
        c = v0 * dd
        for k in range(L):
            mk = cb >> k & 1
            V = x.V[(m1, m2, mk)] if rd & RD_VIDX else x.V[(m1, mk, m2)]
            c -= 2.0 * (dd + sum((V[x.cols[k], x.rows[r]] - M[r, k]) * C[r, k] for r in range(L)))
        if L >= 2:
            for i in range(L):
                for j in range(L):
                    Mn = np.delete(np.delete(M, i, 0), j, 1)
                    if (rd & RD_MINOR) and abs(det(Mn)) <= tol:
                        continue
                    Cm = cof(Mn)
                    mj = cb >> j & 1
                    kfs = [k for k in range(L) if k != j]
                    rfs = [r for r in range(L) if r != i]
                    for k2, kf in enumerate(kfs):
                        mk = cb >> kf & 1
                        key = (m1, m2, mk, mj) if rd & RD_JSLOT else (m1, mk, m2, mj)
                        for r2, rf in enumerate(rfs):
                            c += (-1) ** (i + j) * x.jval(key, i, j, rf, kf) * Cm[r2, k2]
        h2 += c
    return s, h1, h2


def _diff_sum(a: Spin, b: Spin, sub, off, meta, n, rd, tol):
    """Different-spin sum (get_det_adjt_diff, helpers.rs:564; pho.rs:20868-20955)."""
    nn = n * n

    def m2(k):
        return sub[off[k]:off[k] + nn].reshape(n, n)

    def m4(k):
        return sub[off[k]:off[k] + nn * nn].reshape(n, n, n, n)

    vab0 = meta[M_VAB0:M_VAB0 + 4].reshape(2, 2)
    tot = 0.0
    for ba in range(1 << (a.L + 1)):
        if bin(ba).count("1") != a.m:
            continue
        Ma = a.mix(ba >> 1)
        dA = det(Ma)
        if (rd & RD_DIST) and abs(dA) <= tol:
            continue
        Ca = cof(Ma)
        for bb in range(1 << (b.L + 1)):
            if bin(bb).count("1") != b.m:
                continue
            Mb = b.mix(bb >> 1)
            dB = det(Mb)
            if (rd & RD_DIST) and abs(dB) <= tol:
                continue
            Cb = cof(Mb)
            ma0, mb0 = ba & 1, bb & 1
            c = vab0[ma0, mb0] * dA * dB
            for k in range(a.L):
                V = m2(O_VAB + (ma0 * 2 + mb0) * 2 + (ba >> (k + 1) & 1))
                c -= sum(V[a.cols[k], a.rows[r]] * Ca[r, k] for r in range(a.L)) * dB
            for k in range(b.L):
                V = m2(O_VBA + (mb0 * 2 + ma0) * 2 + (bb >> (k + 1) & 1))
                c -= sum(V[b.cols[k], b.rows[r]] * Cb[r, k] for r in range(b.L)) * dA
            for i in range(a.L):
                for j in range(a.L):
                    ma1 = ba >> (j + 1) & 1
                    for k in range(b.L):
                        II = m4(O_II + ((ma0 * 2 + ma1) * 2 + mb0) * 2 + (bb >> (k + 1) & 1))
                        c += Ca[i, j] * sum(II[a.rows[i], a.cols[j], b.rows[r], b.cols[k]] * Cb[r, k]
                                            for r in range(b.L))
            tot += c
    return tot


def element(slab_view, n, bra, ket, rd, enuc, tol):
    """(H, S) of one determinant pair on an m > 0 reference pair with the
    readings `rd` (xw_hamiltonian_overlap_gen_prepared, pho.rs:20550-20966)."""
    sub, off, meta = slab_view
    a, b = spins(slab_view, n, bra, ket)
    sa, h1a, h2a = _same_sum(a, rd, tol)
    sb, h1b, h2b = _same_sum(b, rd, tol)
    h2ab = _diff_sum(a, b, sub, off, meta, n, rd, tol)
    p = prefactor(a, b, bra, ket)
    S = p * sa * sb
    H = p * (h1a * sb + h1b * sa + 0.5 * h2a * sb + 0.5 * h2b * sa + h2ab)
    if not rd & RD_ENUC:
        H += enuc * S
    return H, S
