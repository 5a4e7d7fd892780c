"""Learned screening of selected-NOCI candidates.

Selection in noci-rs (src/snoci/step.rs:178-440, select_candidates
operators.rs:1064) scores every candidate exactly.  Here the exact score is
the Epstein-Nesbet-like contribution

    score(D) = |<D|H - E S|Psi>|^2 / |E - H_DD|,   <D|H - E S|Psi> = v_D - E s_D

with v, s from the engine's projection of the candidates on the current state
(nw_project_hs).  A small network (nwick.mlp, trained on the GPU by the
project's own kernels) learns log10 score from 24 cheap pair descriptors on
the first selection iterations; later iterations evaluate exactly only the
candidates in the network's order until a chunk no longer holds one close to
the current k-th best exact score, plus a random audit, and fall back to
exhaustive exact screening when the audit shows that the network missed a
candidate that would have been selected.

Candidate determinants are generated vectorised (occupied-first references,
nmo <= 64): the excitation phase of holes h (removed highest first) and
particles p (added lowest first) on the occupied-first string is
(-1)^(sum h + n_h (n_occ - n_h) + n_h (n_h - 1) / 2), the closed form of
excitation_phase_bits for that case.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .basis import DET_DTYPE, to_records

N_FEATURES = 24
_TYPES = (("a", 1, 0), ("b", 0, 1), ("aa", 2, 0), ("bb", 0, 2), ("ab", 1, 1))


@dataclass
class Candidates:
    """Vectorised candidate set: parent, per spin holes/particles (-1 padded, 2 slots)."""
    parent: np.ndarray
    ah: np.ndarray
    ap: np.ndarray
    bh: np.ndarray
    bp: np.ndarray
    nocc_a: int
    nocc_b: int

    def __len__(self):
        return len(self.parent)

    def subset(self, idx):
        return Candidates(self.parent[idx], self.ah[idx], self.ap[idx], self.bh[idx], self.bp[idx], self.nocc_a,
                          self.nocc_b)

    def ranks(self):
        return (self.ah >= 0).sum(1), (self.bh >= 0).sum(1)


def _mask(idx):
    m = np.zeros(len(idx), dtype=np.uint64)
    for c in range(idx.shape[1]):
        v = idx[:, c]
        m |= np.where(v >= 0, np.left_shift(np.uint64(1), np.maximum(v, 0).astype(np.uint64)), np.uint64(0))
    return m


def _phase(h, nocc):
    nh = (h >= 0).sum(1)
    odd = (np.where(h >= 0, h, 0).sum(1) + nh * (nocc - nh) + nh * (nh - 1) // 2) & 1
    return np.where(odd == 1, -1.0, 1.0)


def records(c: Candidates) -> np.ndarray:
    """nw_det records of the candidates (low 64-bit words; nmo <= 64)."""
    n = len(c)
    out = np.zeros(n, dtype=DET_DTYPE)
    ref_a = np.uint64((1 << c.nocc_a) - 1)
    ref_b = np.uint64((1 << c.nocc_b) - 1)
    ha, pa, hb, pb = _mask(c.ah), _mask(c.ap), _mask(c.bh), _mask(c.bp)
    out["parent"] = c.parent
    out["pha"] = _phase(c.ah, c.nocc_a)
    out["phb"] = _phase(c.bh, c.nocc_b)
    out["oa"][:, 0] = (ref_a & ~ha) | pa
    out["ob"][:, 0] = (ref_b & ~hb) | pb
    out["ha"][:, 0], out["pa"][:, 0], out["hb"][:, 0], out["pb"][:, 0] = ha, pa, hb, pb
    return out


def random_candidates(nref, nmo, nocc_a, nocc_b, per_ref, seed, mix=(0.1, 0.1, 0.8 / 3, 0.8 / 3, 0.8 / 3)):
    """Seeded unique singles and doubles per reference (default 10% alpha
    singles, 10% beta singles, doubles split over aa/bb/ab), each reference's
    block ordered by rank then indices (generation-like order)."""
    assert nmo <= 64
    rng = np.random.default_rng(seed)
    parts = []
    for r in range(nref):
        blocks = []
        from math import comb
        va, vb = nmo - nocc_a, nmo - nocc_b
        pool = [nocc_a * va, nocc_b * vb, comb(nocc_a, 2) * comb(va, 2), comb(nocc_b, 2) * comb(vb, 2),
                nocc_a * va * nocc_b * vb]
        # each type capped at half of what exists (rejection sampling stays
        # cheap); the remainder goes to alpha-beta doubles
        wants = [min(int(per_ref * f), p // 2) for f, p in zip(mix[:-1], pool[:-1])]
        wants.append(per_ref - sum(wants))
        if wants[-1] > pool[-1] // 2:
            raise ValueError("per_ref too large for this orbital space")
        for (name, ka, kb), want in zip(_TYPES, wants):
            if want == 0:
                continue
            got = np.zeros((0, 8), dtype=np.int64)
            tries = 0
            while len(got) < want:
                tries += 1
                if tries > 50:
                    raise ValueError(f"cannot draw {want} unique {name} excitations")
                m = 2 * (want - len(got)) + 64
                cols = []
                for k, no, lo, hi in ((ka, nocc_a, 0, nocc_a), (ka, nocc_a, nocc_a, nmo), (kb, nocc_b, 0, nocc_b),
                                      (kb, nocc_b, nocc_b, nmo)):
                    if k == 0:
                        cols.append(np.full((m, 2), -1))
                        continue
                    # k distinct values of [lo, hi) per row: the k smallest of i.i.d. keys
                    x = np.sort(np.argpartition(rng.random((m, hi - lo)), k - 1, axis=1)[:, :k], 1) + lo
                    if k == 1:
                        x = np.concatenate([x, np.full((m, 1), -1)], 1)
                    cols.append(x)
                new = np.concatenate(cols, 1)
                got = np.unique(np.concatenate([got, new]), axis=0)
            sel = rng.choice(len(got), want, replace=False)
            blocks.append(got[np.sort(sel)])
        e = np.concatenate(blocks)
        ra = (e[:, 0:2] >= 0).sum(1)
        rb = (e[:, 4:6] >= 0).sum(1)
        order = np.lexsort(tuple(e[:, k] for k in range(7, -1, -1)) + (ra, ra + rb))
        e = e[order]
        parts.append(np.concatenate([np.full((len(e), 1), r), e], 1))
    a = np.concatenate(parts)
    return Candidates(a[:, 0].astype(np.uint32), a[:, 1:3], a[:, 3:5], a[:, 5:7], a[:, 7:9], nocc_a, nocc_b)


# ---------------------------------------------------------------------------
# Descriptors
# ---------------------------------------------------------------------------
def reference_fock(sysm):
    """Per reference and spin: the reference determinant's own spin Fock matrix
    in its MO basis (h + J - K of its occupied densities, src/scf/kernels.rs:46)."""
    from .pt2 import fock
    out = []
    for Ca, Cb in sysm.refs:
        Da = Ca[:, :sysm.nocc_a] @ Ca[:, :sysm.nocc_a].T
        Db = Cb[:, :sysm.nocc_b] @ Cb[:, :sysm.nocc_b].T
        fa, fb = fock(sysm.h, sysm.eri, Da, Db)
        out.append((Ca.T @ fa @ Ca, Cb.T @ fb @ Cb))
    return out


def _pair_sv(sysm, p, d, sp):
    no = sysm.nocc_a if sp == 0 else sysm.nocc_b
    A = sysm.refs[p][sp][:, :no].T @ sysm.S @ sysm.refs[d][sp][:, :no]
    return np.sort(np.linalg.svd(A, compute_uv=False))


def parent_table(sysm, state):
    """Per parent reference: weight in the current state, diagonal energy
    relative to E, log10 reduced overlap, zero count, the two smallest
    occupied-overlap singular values per spin and their mean, all with the
    dominant reference (largest weight)."""
    nref = sysm.nref
    wpar = state["parent_weight"]
    dom = int(np.argmax(wpar))
    pf = np.zeros((nref, 10))
    for p in range(nref):
        sa, sb = _pair_sv(sysm, p, dom, 0), _pair_sv(sysm, p, dom, 1)
        nz = np.concatenate([sa, sb])
        zc = (nz <= 1e-8).sum()
        red = np.prod(np.where(nz > 1e-8, nz, 1.0))
        pf[p] = [wpar[p], state["parent_energy"][p] - state["energy"], np.log10(max(red, 1e-30)), zc,
                 sa[0], sa[1] if len(sa) > 1 else 1.0, sb[0], sb[1] if len(sb) > 1 else 1.0, nz.mean(),
                 1.0 if p == dom else 0.0]
    return pf


def static_features(sysm, cand: Candidates, fmo, hpar):
    """The state-independent descriptor columns: excitation ranks (0-1), hole
    and particle orbital energies of the parent's Fock per spin (2-9), the
    orbital-energy gap (10), the exact coupling |H_Dp| with the parent
    reference (11) and the one-body Fock coupling of the excitation (22)."""
    nref = sysm.nref
    n = len(cand)
    F = np.zeros((n, N_FEATURES), dtype=np.float32)
    ra, rb = cand.ranks()
    F[:, 0], F[:, 1] = ra, rb
    par = cand.parent.astype(np.int64)
    eps = [np.stack([np.diag(fmo[p][sp]) for p in range(nref)]) for sp in (0, 1)]
    de = np.zeros(n)
    coup = np.zeros(n)
    for sp, (h, q, col) in enumerate(((cand.ah, cand.ap, 2), (cand.bh, cand.bp, 6))):
        e = eps[sp]
        for k in range(2):
            hk, qk = h[:, k], q[:, k]
            eh = np.where(hk >= 0, e[par, np.maximum(hk, 0)], 0.0)
            ep = np.where(qk >= 0, e[par, np.maximum(qk, 0)], 0.0)
            F[:, col + k], F[:, col + 2 + k] = eh, ep
            de += ep - eh
        fm = np.stack([fmo[p][sp] for p in range(nref)])
        h0, q0, h1, q1 = h[:, 0], q[:, 0], h[:, 1], q[:, 1]
        single = (h0 >= 0) & (h1 < 0)
        double = h1 >= 0
        f00 = fm[par, np.maximum(h0, 0), np.maximum(q0, 0)]
        f11 = fm[par, np.maximum(h1, 0), np.maximum(q1, 0)]
        f01 = fm[par, np.maximum(h0, 0), np.maximum(q1, 0)]
        f10 = fm[par, np.maximum(h1, 0), np.maximum(q0, 0)]
        coup += np.where(single, np.abs(f00), 0.0) + np.where(double, np.abs(f00 * f11 - f01 * f10), 0.0)
    F[:, 10] = de
    F[:, 11] = np.log10(np.abs(hpar) + 1e-14)
    F[:, 22] = np.log10(coup + 1e-12)
    return F


def dynamic_features(xp, Fs, par, pf, hpar, hdom, hdiag, energy):
    """Complete the descriptor rows with the state-dependent columns: the
    parent table (12-19, 23), the exact coupling |H_Dd| with the dominant
    reference (20) and the first-order amplitude |H_Dp| / |H_DD - E| (21).
    xp is numpy or torch (then every array is a device tensor)."""
    F = Fs.clone() if hasattr(Fs, "clone") else Fs.copy()
    t = pf[par]
    F[:, 12:20] = t[:, :8]
    F[:, 23] = t[:, 8]
    F[:, 20] = xp.log10(xp.abs(hdom) + 1e-14)
    den = xp.clip(xp.abs(hdiag - energy), 1e-3, None) if xp is np else (hdiag - energy).abs().clamp(min=1e-3)
    F[:, 21] = xp.log10(xp.abs(hpar) / den + 1e-14)
    return F


def dynamic_features_into(X, mu, sd, par, pf, hpar, hdom, hdiag, energy):
    """dynamic_features written normalised, (F - mu) / sd, into the
    state-dependent columns of the device matrix X in place (the static
    columns of X already hold their normalised values)."""
    t = pf[par]
    X[:, 12:20] = (t[:, :8] - mu[12:20]) / sd[12:20]
    X[:, 23] = (t[:, 8] - mu[23]) / sd[23]
    X[:, 20] = (hdom.abs().add(1e-14).log10() - mu[20]) / sd[20]
    den = (hdiag - energy).abs().clamp(min=1e-3)
    X[:, 21] = ((hpar.abs() / den + 1e-14).log10() - mu[21]) / sd[21]


def features(sysm, cand: Candidates, fmo, state, couplings):
    """All 24 pair descriptors on the host (numpy): static_features then
    dynamic_features.  couplings = (H_Dp, H_Dd, H_DD)."""
    hp, hd, hdiag = couplings
    Fs = static_features(sysm, cand, fmo, hp)
    pf = parent_table(sysm, state).astype(np.float32)
    return dynamic_features(np, Fs, cand.parent.astype(np.int64), pf, hp, hd, hdiag, state["energy"])


# ---------------------------------------------------------------------------
# Exact scores through the engine
# ---------------------------------------------------------------------------
class ExactScorer:
    """Candidate projections v, s on the current state (nw_project_hs) and the
    candidates' diagonal H_DD (nw_hs_batched), all on the GPU."""

    def __init__(self, eng, cand: Candidates, device="cuda:0", mode=0):
        import torch
        self.torch, self.eng, self.mode = torch, eng, mode
        self.dev = torch.device(device)
        rec = records(cand)
        self.rec = rec
        self.rec_d = torch.from_numpy(rec.view(np.uint8).copy()).to(self.dev).view(-1, DET_DTYPE.itemsize)
        n = len(cand)
        req = np.stack([np.arange(n), np.arange(n)], 1).astype(np.uint64)
        self.hdiag, self.sdiag = eng.hs_batched(rec, req, mode=mode)
        self.hdiag_d = torch.from_numpy(np.ascontiguousarray(self.hdiag)).to(self.dev)
        self.n_exact = 0

    def coupling(self, ket_rec, ket_of):
        """H_{D, k(D)} for every candidate D (bra) with ket ket_rec[ket_of[D]]."""
        n = len(self.rec)
        allrec = np.concatenate([self.rec, ket_rec])
        req = np.stack([np.arange(n), n + np.asarray(ket_of, dtype=np.int64)], 1).astype(np.uint64)
        h, _ = self.eng.hs_batched(allrec, req, mode=self.mode)
        return h

    def project_single(self, ket_rec):
        """H_{D,k} for every candidate D against one determinant k, as a device
        tensor (one projection launch with a one-determinant current space)."""
        torch = self.torch
        k = torch.from_numpy(ket_rec.view(np.uint8).copy()).to(self.dev).view(-1, DET_DTYPE.itemsize)
        allrec = torch.cat([k, self.rec_d]).contiguous()
        n = self.rec_d.shape[0]
        v = torch.empty(n, dtype=torch.float64, device=self.dev)
        s = torch.empty(n, dtype=torch.float64, device=self.dev)
        one = torch.ones(1, dtype=torch.float64, device=self.dev)
        self.eng.project_hs(allrec, 1, one, v=v, s=s, mode=self.mode)
        return v

    def current(self, cur_rec, coef):
        """The current space and its coefficients on the device, for repeated
        scores_d calls within one selection iteration."""
        torch = self.torch
        cur = torch.from_numpy(cur_rec.view(np.uint8).copy()).to(self.dev).view(-1, DET_DTYPE.itemsize)
        c = torch.as_tensor(np.ascontiguousarray(coef), dtype=torch.float64, device=self.dev)
        return cur, c

    def scores_d(self, cur_rec, coef, energy, idx=None, cur=None):
        """Exact scores |<D|H - E S|Psi>|^2 / max(|E - H_DD|, 1e-3) of the
        candidates idx (a device index tensor or array; all when None), in idx
        order, as a device tensor.  cur: current() of (cur_rec, coef)."""
        torch = self.torch
        cur, c = cur if cur is not None else self.current(cur_rec, coef)
        if idx is None:
            sub = self.rec_d
            hd = self.hdiag_d
        else:
            it = idx if torch.is_tensor(idx) else torch.as_tensor(np.asarray(idx, dtype=np.int64), device=self.dev)
            sub = self.rec_d.index_select(0, it)
            hd = self.hdiag_d.index_select(0, it)
        n = sub.shape[0]
        allrec = torch.cat([cur, sub]).contiguous()
        v = torch.empty(n, dtype=torch.float64, device=self.dev)
        s = torch.empty(n, dtype=torch.float64, device=self.dev)
        self.eng.project_hs(allrec, cur.shape[0], c, v=v, s=s, mode=self.mode)
        num = v - energy * s
        self.n_exact += n
        return num * num / (energy - hd).abs().clamp(min=1e-3)

    def scores(self, cur_rec, coef, energy, idx=None):
        """scores_d as a host array."""
        return self.scores_d(cur_rec, coef, energy, idx).cpu().numpy()


def pick_best(sc_d, taken_d, k):
    """select_candidates (src/snoci/operators.rs:1064-1085) with threshold 0:
    the k best scores above zero among the candidates not yet taken, best
    first, equal scores in index order (a descending stable sort), on the
    scores' device; returns host indices."""
    import torch
    vals, idx = torch.sort(sc_d.masked_fill(taken_d, -1.0), descending=True, stable=True)
    return idx[:k][vals[:k] > 0].cpu().numpy()


def kth_best_d(sc_d, k):
    """The k-th largest exact score of a device score vector whose entries
    without a score are -1; 0 when fewer than k are scored."""
    import torch
    if k > sc_d.shape[0]:
        return 0.0
    v = float(torch.topk(sc_d, k).values[-1])
    return v if v >= 0.0 else 0.0


def solve_current(eng, cur_rec, mode=0, tol=1e-8):
    """Lowest root of the current space (general_evp of src/maths/linalg.rs:239)."""
    from .pt2 import general_evp
    H, S = eng.hs_matrix(cur_rec, mode=mode)
    e, C = general_evp(H, S, tol)
    return float(e[0]), C[:, 0], H


@dataclass
class SelectionResult:
    energies: list = field(default_factory=list)
    selected: list = field(default_factory=list)
    screen_s: list = field(default_factory=list)
    exact_evals: list = field(default_factory=list)
    audit_misses: list = field(default_factory=list)
    fallbacks: int = 0
    recall: list = field(default_factory=list)
    captured: list = field(default_factory=list)  # exact score of the picks / of the true best picks
    train: dict = field(default_factory=dict)


def select(eng, sysm, cand: Candidates, iters=4, k_add=64, learned=False, learn_iters=2, top_frac=0.05,
           audit_frac=0.01, chunk_frac=0.01, chunk_ratio=0.1, chunk_patience=2, max_frac=0.5, epochs=8, batch=4096,
           lr=2e-3, seed=0, mode=0, measure_recall=False, log=None):
    """Selected NOCI from the references: each iteration adds the k_add
    best-scoring candidates.  learned=True trains the network on the exact
    scores of the first learn_iters iterations and screens the later ones.
    Timings in screen_s cover scoring and picking (descriptors, inference,
    exact evaluation, audit and fallback); measure_recall adds an untimed
    exhaustive scoring of each screened iteration to report recall."""
    from .basis import Det
    from .mlp import MLP
    refs = [Det(i, (1 << sysm.nocc_a) - 1, (1 << sysm.nocc_b) - 1) for i in range(sysm.nref)]
    cur_rec = to_records(refs)
    ref_rec = to_records(refs)
    scorer = ExactScorer(eng, cand, mode=mode)
    torch, dev = scorer.torch, scorer.dev
    fmo = reference_fock(sysm)
    rng = np.random.default_rng(seed)
    taken = np.zeros(len(cand), dtype=bool)
    taken_d = torch.zeros(len(cand), dtype=torch.bool, device=dev)
    hdom_of = {}
    Xcur = None

    def screened(cur_rec, c, e, pf, hdom):
        """One screened iteration: network ranking, exact scores of the top
        fraction and of further chunks until chunk_patience chunks in a row
        hold no candidate within chunk_ratio of the k_add-th best exact score,
        a random audit of the rest, and the exhaustive fallback on an audit
        miss.  Returns (scores, exact evaluations, misses, evaluated top, 0/1)."""
        cur = scorer.current(cur_rec, c)
        # the normalised descriptors: static columns kept from training,
        # the state-dependent ones rewritten
        dynamic_features_into(Xcur, mu, sd, par_d, pf, hpar_d, hdom, hdiag_d, e)
        pred = torch.empty(N, dtype=torch.float32, device=dev)
        net.forward(Xcur, out=pred)
        pred.masked_fill_(taken_d, -float("inf"))
        nfree = N - int(taken.sum())
        order = torch.argsort(pred, descending=True)[:nfree]
        sc_d = torch.full((N,), -1.0, dtype=torch.float64, device=dev)
        chunk = max(k_add, int(chunk_frac * N))
        done = min(nfree, max(k_add, int(top_frac * N)))
        first = torch.sort(order[:done]).values
        s_f = scorer.scores_d(cur_rec, c, e, first, cur=cur)
        sc_d[first] = s_f
        best = torch.topk(s_f, min(k_add, len(s_f))).values  # running k_add best exact scores
        quiet = 0
        while done < nfree and done < max_frac * N:
            kth = float(best[-1]) if len(best) >= k_add else 0.0
            nxt = torch.sort(order[done:done + chunk]).values
            s_n = scorer.scores_d(cur_rec, c, e, nxt, cur=cur)
            sc_d[nxt] = s_n
            best = torch.topk(torch.cat([best, s_n]), min(k_add, len(best) + len(s_n))).values
            done += len(nxt)
            quiet = quiet + 1 if float(s_n.max()) < chunk_ratio * kth else 0
            if quiet >= chunk_patience:
                break
        top = order[:done]
        rest = order[done:]
        m = min(len(rest), max(1, int(audit_frac * N)))
        audit = torch.unique(rest[torch.randint(len(rest), (m,), generator=gen, device=dev)]) \
            if len(rest) else rest
        n_eval = done
        misses = 0
        if len(audit):
            s_a = scorer.scores_d(cur_rec, c, e, audit, cur=cur)
            sc_d[audit] = s_a
            n_eval += len(audit)
            # the exact evaluation judges the network: an audited candidate it
            # ranked out that beats the k_add-th best is a miss
            allbest = torch.topk(torch.cat([best, s_a]), min(k_add, len(best) + len(s_a))).values
            kth = float(allbest[-1]) if len(allbest) >= k_add else 0.0
            misses = int((s_a > kth).sum())
        fell = 0
        if misses > 0 or done >= max_frac * N:
            fell = 1
            sc_d = scorer.scores_d(cur_rec, c, e, cur=cur)
            n_eval += N
        return sc_d, n_eval, misses, top, fell
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    res = SelectionResult()
    Xs, ys = [], []
    net, mu, sd = None, None, None
    e_ref0 = None
    N = len(cand)
    if learned:
        t0 = time.perf_counter()
        hpar = scorer.coupling(ref_rec, cand.parent.astype(np.int64))
        Fs = torch.from_numpy(static_features(sysm, cand, fmo, hpar)).to(dev)
        par_d = torch.from_numpy(cand.parent.astype(np.int64)).to(dev)
        hpar_d = torch.from_numpy(hpar).to(dev)
        hdiag_d = torch.from_numpy(np.ascontiguousarray(scorer.hdiag)).to(dev)
        res.train["static_descriptor_s"] = time.perf_counter() - t0
    for it in range(iters):
        e, c, H = solve_current(eng, cur_rec, mode)
        if e_ref0 is None:
            e_ref0 = e
        res.energies.append(e)
        par = cur_rec["parent"].astype(int)
        wpar = np.bincount(par, weights=c * c, minlength=sysm.nref)
        isref = (cur_rec["ha"][:, 0] == 0) & (cur_rec["hb"][:, 0] == 0)
        penergy = np.zeros(sysm.nref)
        penergy[par[isref]] = np.diag(H)[isref]
        state = {"parent_weight": wpar, "parent_energy": penergy, "energy": e, "reference_energy": e_ref0}
        t0 = time.perf_counter()
        use_net = learned and it >= learn_iters
        F = None
        if learned:
            dom = int(np.argmax(wpar))
            pf = torch.from_numpy(parent_table(sysm, state).astype(np.float32)).to(dev)
            if dom not in hdom_of:  # couplings with the dominant reference, once per reference
                hdom_of[dom] = scorer.project_single(ref_rec[dom:dom + 1])
            hdom = hdom_of[dom]
            if not use_net:
                F = dynamic_features(torch, Fs, par_d, pf, hpar_d, hdom, hdiag_d, e)
        if not use_net:
            sc_d = scorer.scores_d(cur_rec, c, e)
            n_eval = N
            if learned:
                Xs.append(F)
                ys.append(torch.log10(sc_d + 1e-20).float())
        else:
            sc_d, n_eval, misses, top, fell = screened(cur_rec, c, e, pf, hdom)
            res.audit_misses.append(misses)
            res.fallbacks += fell
        pick = pick_best(sc_d, taken_d, k_add)
        res.screen_s.append(time.perf_counter() - t0)
        res.exact_evals.append(int(n_eval))
        if use_net and measure_recall:
            full = scorer.scores(cur_rec, c, e)
            scorer.n_exact -= N
            best = np.argsort(-np.where(taken, -1.0, full))[:k_add]
            res.recall.append(float(np.isin(best, top.cpu().numpy()).mean()))
            res.captured.append(float(full[pick].sum() / full[best].sum()))
        cur_it, c_it = cur_rec, c  # this iteration's current space, for the warm-up below
        taken[pick] = True
        taken_d[torch.from_numpy(pick).to(dev)] = True
        res.selected.append(pick)
        cur_rec = np.concatenate([cur_rec, scorer.rec[np.sort(pick)]])
        cur_rec = cur_rec[np.argsort(cur_rec["parent"], kind="stable")]  # runs of one parent together
        if log:
            log(f"selection iteration {it}: E={e:.10f} evaluated={n_eval} picked={len(pick)} "
                f"screen {res.screen_s[-1]:.3f} s" + (f" misses={res.audit_misses[-1]}" if use_net else ""))
        if learned and it == learn_iters - 1:
            X = torch.cat(Xs)
            y = torch.cat(ys).cpu().numpy()
            mu, sd = X.mean(0), X.std(0) + 1e-6
            Xn = ((X - mu) / sd).contiguous()
            # emphasise the top of the distribution: the selection only needs its order
            wt = np.where(y >= np.quantile(y, 0.9), 5.0, 1.0).astype(np.float32)
            yd = torch.from_numpy(y - y.mean()).to(dev)
            wd = torch.from_numpy(wt).to(dev)
            net = MLP(N_FEATURES, seed=seed)
            t1 = time.perf_counter()
            el = net.train(Xn, yd, wd, epochs=epochs, batch=batch, lr=lr, seed=seed)
            # first use of the inference path (network forward, full-size sort,
            # masked assignment) once here, with the training time, so the
            # screened iterations carry only their own work
            Xcur = ((Fs - mu) / sd).contiguous()
            # first use of the screened path (allocations and kernels at their
            # sizes) once here, with the training time, on the state of this
            # iteration; its result is discarded
            n_before = scorer.n_exact
            screened(cur_it, c_it, e, pf, hdom)
            pick_best(torch.zeros(N, dtype=torch.float64, device=dev), taken_d, k_add)
            scorer.n_exact = n_before
            torch.cuda.synchronize()
            res.train.update({"samples": int(len(y)), "epochs": epochs, "loss": [float(v) for v in el],
                              "seconds": time.perf_counter() - t1})
            del Xs, X, Xn
    e, c, H = solve_current(eng, cur_rec, mode)
    res.energies.append(e)
    return res


# ---------------------------------------------------------------------------
# PT2 candidates ranked by the screener
# ---------------------------------------------------------------------------
def candidate_dets(cand: Candidates, idx):
    """nwick.basis.Det objects of the candidates idx (for nwick.pt2.PT2)."""
    from .basis import Det
    ra, rb = (1 << cand.nocc_a) - 1, (1 << cand.nocc_b) - 1
    out = []
    for i in np.asarray(idx, dtype=np.int64):
        f = lambda a: [int(v) for v in a[i] if v >= 0]
        out.append(Det(int(cand.parent[i]), ra, rb, f(cand.ah), f(cand.ap), f(cand.bh), f(cand.bp)))
    return out


@dataclass
class ScreenedPT2:
    ept2: float
    e_current: float
    kept: np.ndarray            # candidate indices in the PT2 space
    exact_evals: int            # candidates scored exactly (training, ranking, audit)
    tail_estimate: float        # estimated first-order weight of the dropped candidates
    audit_tail: float           # the audit's own estimate of it
    fallback: bool
    solve: dict = field(default_factory=dict)
    seconds: dict = field(default_factory=dict)


def screened_pt2(eng, sysm, cand: Candidates, tau=1e-6, train_frac=0.1, top_frac=0.05, chunk_frac=0.02,
                 chunk_patience=2, audit_frac=0.02, epochs=30, batch=1024, lr=2e-3, seed=0, mode=0, solve_kw=None,
                 factor_tables="device", log=None):
    """NOCI-PT2 of the reference space over the candidates the learned
    screener ranks as contributing.  The reference solves the PT2 equations
    over every candidate (src/snoci/step.rs:198-429); here the network,
    trained on the exact first-order weights |<D|H - E S|Psi>|^2 / |E - H_DD|
    of a random train_frac of the candidates, orders the rest; they are
    scored exactly in that order, a chunk at a time, until chunk_patience
    chunks in a row each hold less than tau / 4 of weight and the estimated
    weight of everything unscored (the last chunk's mean weight times the
    unscored count) is below tau / 2.  A random audit of the unscored rest
    judges the network: when the audit's own estimate of the dropped weight
    exceeds tau, every candidate is scored exactly and the PT2 space is cut
    by weight instead (the exact fallback).  The PT2 equations are then
    solved over the kept candidates with nwick.pt2.PT2."""
    import torch

    from .basis import Det
    from .mlp import MLP
    from .pt2 import PT2
    solve_kw = dict(solve_kw or {})
    t0 = time.perf_counter()
    refs = [Det(i, (1 << sysm.nocc_a) - 1, (1 << sysm.nocc_b) - 1) for i in range(sysm.nref)]
    ref_rec = to_records(refs)
    scorer = ExactScorer(eng, cand, mode=mode)
    dev = scorer.dev
    N = len(cand)
    e, c, H = solve_current(eng, ref_rec, mode)
    wpar = np.bincount(np.arange(sysm.nref), weights=c * c, minlength=sysm.nref)
    state = {"parent_weight": wpar, "parent_energy": np.diag(H).copy(), "energy": e, "reference_energy": e}
    fmo = reference_fock(sysm)
    hpar = scorer.coupling(ref_rec, cand.parent.astype(np.int64))
    dom = int(np.argmax(wpar))
    hdom = scorer.project_single(ref_rec[dom:dom + 1])
    Fs = torch.from_numpy(static_features(sysm, cand, fmo, hpar)).to(dev)
    par_d = torch.from_numpy(cand.parent.astype(np.int64)).to(dev)
    pf = torch.from_numpy(parent_table(sysm, state).astype(np.float32)).to(dev)
    F = dynamic_features(torch, Fs, par_d, pf, torch.from_numpy(hpar).to(dev), hdom, scorer.hdiag_d, e)
    cur = scorer.current(ref_rec, c)
    sc_d = torch.full((N,), -1.0, dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    # training on exact weights of a random subset
    tr = torch.randperm(N, generator=gen, device=dev)[:max(64, int(train_frac * N))]
    tr = torch.sort(tr).values
    sc_d[tr] = scorer.scores_d(ref_rec, c, e, tr, cur=cur)
    mu, sd = F.mean(0), F.std(0) + 1e-6
    X = ((F - mu) / sd).contiguous()
    y = torch.log10(sc_d[tr] + 1e-20).float()
    net = MLP(N_FEATURES, seed=seed)
    net.train(X[tr].contiguous(), (y - y.mean()).contiguous(), None, epochs=epochs, batch=batch, lr=lr, seed=seed)
    t1 = time.perf_counter()
    pred = torch.empty(N, dtype=torch.float32, device=dev)
    net.forward(X, out=pred)
    net.close()
    pred[tr] = -float("inf")  # already scored
    order = torch.argsort(pred, descending=True)[:N - len(tr)]
    done, quiet = 0, 0
    chunk = max(64, int(chunk_frac * N))
    first = max(chunk, int(top_frac * N))
    tail = float("inf")
    while done < len(order):
        step = first if done == 0 else chunk
        nxt = torch.sort(order[done:done + step]).values
        s_n = scorer.scores_d(ref_rec, c, e, nxt, cur=cur)
        sc_d[nxt] = s_n
        done += len(nxt)
        tail = float(s_n.mean()) * (len(order) - done)
        quiet = quiet + 1 if float(s_n.sum()) < 0.25 * tau else 0
        if quiet >= chunk_patience and tail < 0.5 * tau:
            break
    rest = order[done:]
    audit_tail, fallback = 0.0, False
    if len(rest):
        m = min(len(rest), max(16, int(audit_frac * N)))
        aud = torch.unique(rest[torch.randint(len(rest), (m,), generator=gen, device=dev)])
        s_a = scorer.scores_d(ref_rec, c, e, aud, cur=cur)
        sc_d[aud] = s_a
        audit_tail = float(s_a.mean()) * len(rest)
        if audit_tail > tau:
            fallback = True
            sc_d = scorer.scores_d(ref_rec, c, e, cur=cur)
    scored = (sc_d >= 0).cpu().numpy()
    w = sc_d.cpu().numpy()
    if fallback:
        # keep the largest weights until what is left sums below tau / 2
        srt = np.argsort(-w)
        cum = np.cumsum(w[srt][::-1])[::-1]
        keep_n = int(np.searchsorted(-cum, -0.5 * tau))
        kept = np.sort(srt[:max(keep_n, 1)])
        tail = float(w[srt[keep_n:]].sum())
    else:
        kept = np.nonzero(scored)[0]
    t2 = time.perf_counter()
    if log:
        log(f"screened PT2: kept {len(kept)} of {N} candidates, exact evaluations {scorer.n_exact}, "
            f"tail {tail:.2e} (audit {audit_tail:.2e}), fallback {fallback}")
    pt = PT2(eng, sysm, refs, candidate_dets(cand, kept), tol=1e-8, mode=mode, factor_tables=factor_tables)
    pt.setup()
    r = pt.solve(**solve_kw)
    if pt.fac is not None:
        pt.fac.close()
    t3 = time.perf_counter()
    return ScreenedPT2(ept2=r["ept2"], e_current=pt.ecurrent, kept=kept, exact_evals=int(scorer.n_exact),
                       tail_estimate=tail, audit_tail=audit_tail, fallback=fallback, solve=r,
                       seconds={"train": t1 - t0, "rank_and_score": t2 - t1, "pt2": t3 - t2})


def candidates_from_dets(dets, nocc_a, nocc_b) -> Candidates:
    """The vectorised candidate arrays of a list of singles and doubles
    (nwick.basis.Det, occupied-first references)."""
    n = len(dets)
    arr = {k: np.full((n, 2), -1, dtype=np.int64) for k in ("ah", "ap", "bh", "bp")}
    for i, d in enumerate(dets):
        for k in arr:
            bits = [b for b in range(128) if getattr(d, {"ah": "ha", "ap": "pa", "bh": "hb", "bp": "pb"}[k]) >> b & 1]
            if len(bits) > 2:
                raise ValueError("candidate arrays hold singles and doubles")
            arr[k][i, :len(bits)] = bits
    return Candidates(np.array([d.parent for d in dets], dtype=np.uint32), arr["ah"], arr["ap"], arr["bh"],
                      arr["bp"], nocc_a, nocc_b)
