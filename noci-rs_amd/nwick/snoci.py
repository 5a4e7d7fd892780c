"""Selected NOCI (SNOCI) driver over the GPU PT2 step.

Host-side mirror of src/snoci/step.rs:178-550 (snoci_step),
src/snoci/candidate.rs (CandidatePool), src/basis/excitation.rs:156-370
(excitation labels and generate_excited_basis applied to possibly excited
determinants) and src/snoci/operators.rs:1064-1083 (select_candidates).

Each iteration solves the current space, runs NOCI-PT2 over the candidate
pool (the PT2 step's matrix elements and GMRES run on the GPU through
nwick.pt2.PT2), scores candidates by |a_i V_i| and moves the best `max_add`
of them into the current space.  Candidates are identified by their label
string exactly as the reference builds it, which fixes the pool's order and
its de-duplication and therefore which of several tied candidates is chosen.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from itertools import combinations

from .basis import Det


def _bits(mask):
    out = []
    while mask:
        low = mask & -mask
        out.append(low.bit_length() - 1)
        mask ^= low
    return out


def _suffix(ah, ap, bh, bp):
    """excitation_label (excitation.rs:156-190)."""
    parts = []
    if ah:
        parts.append("alpha {} -> {}".format(" ".join(map(str, _bits(ah))), " ".join(map(str, _bits(ap)))))
    if bh:
        parts.append("beta {} -> {}".format(" ".join(map(str, _bits(bh))), " ".join(map(str, _bits(bp)))))
    return "(" + "; ".join(parts) + ")"


@dataclass
class State:
    """A determinant with its reference-style label (DetState's label/parent/occupations)."""
    det: Det
    label: str
    ref_oa: int
    ref_ob: int


def reference_states(labels, nocc_a, nocc_b):
    oa, ob = (1 << nocc_a) - 1, (1 << nocc_b) - 1
    return [State(Det(i, oa, ob), lab, oa, ob) for i, lab in enumerate(labels)]


def excite(r: State, nmo, orders):
    """generate_excited_basis (excitation.rs:276-370) for one state `r`,
    which may itself be excited relative to its parent reference."""
    d = r.det
    occ_a = [p for p in range(nmo) if d.oa >> p & 1]
    occ_b = [p for p in range(nmo) if d.ob >> p & 1]
    vir_a = [p for p in range(nmo) if not d.oa >> p & 1]
    vir_b = [p for p in range(nmo) if not d.ob >> p & 1]
    out = []
    for k in sorted(set(orders)):
        for ka in range(k + 1):
            kb = k - ka
            for ah in combinations(occ_a, ka):
                for ap in combinations(vir_a, ka):
                    for bh in combinations(occ_b, kb):
                        for bp in combinations(vir_b, kb):
                            oa = d.oa
                            for i in ah:
                                oa &= ~(1 << i)
                            for a in ap:
                                oa |= 1 << a
                            ob = d.ob
                            for i in bh:
                                ob &= ~(1 << i)
                            for a in bp:
                                ob |= 1 << a
                            th_a, tp_a = r.ref_oa & ~oa, oa & ~r.ref_oa
                            th_b, tp_b = r.ref_ob & ~ob, ob & ~r.ref_ob
                            nd = Det(d.parent, r.ref_oa, r.ref_ob, _bits(th_a), _bits(tp_a), _bits(th_b), _bits(tp_b))
                            out.append(State(nd, f"{r.label} {_suffix(th_a, tp_a, th_b, tp_b)}", r.ref_oa,
                                             r.ref_ob))
    return out


class CandidatePool:
    """CandidatePool (candidate.rs): label-keyed pool of candidate states."""

    def __init__(self, selected, nmo, orders):
        self.nmo, self.orders = nmo, orders
        self.candidates = [c for s in selected for c in excite(s, nmo, orders)]

    def update(self, selected_space, newly_selected):
        if not newly_selected:
            return
        sel = {s.label for s in newly_selected}
        self.candidates = [c for c in self.candidates if c.label not in sel]
        new = [c for s in newly_selected for c in excite(s, self.nmo, self.orders)]
        existing = {s.label for s in selected_space} | {c.label for c in self.candidates}
        self.candidates += [c for c in new if c.label not in existing]


def select_candidates(candidates, scores, sigma, max_add):
    """select_candidates (operators.rs:1064-1083): scores above sigma,
    descending, stable for ties, at most max_add."""
    ranked = [(c, s) for c, s in zip(candidates, scores) if s > sigma]
    ranked.sort(key=lambda t: -t[1])
    return [c for c, _ in ranked[:max_add]]


@dataclass
class SNOCIOptions:
    """SNOCIOptions defaults (src/input/snoci.rs:140-160)."""
    orders: tuple = (1, 2)
    max_iter: int = 5
    max_add: int = 1
    max_dim: int = 100
    sigma: float = 1e-6
    tol: float = 1e-8
    res_tol: float = 1e-8
    restart: int = 30
    gmres_max_iter: int = 300
    preconditioner: str = "woodbury"


@dataclass
class SNOCIResult:
    ecurrent: float
    coeffs: object
    current: list
    candidates: list
    selected: list
    ept2: float | None = None
    history: list = field(default_factory=list)


class GpuPT2Backend:
    """One NOCI-PT2 solve per iteration through the engine (nwick.pt2.PT2)."""

    def __init__(self, eng, sysm, tol=1e-12, mode=0, device="cuda:0"):
        self.eng, self.sysm, self.tol, self.mode, self.device = eng, sysm, tol, mode, device

    def __call__(self, current, candidates, opts: SNOCIOptions):
        from .pt2 import PT2
        pt = PT2(self.eng, self.sysm, current, candidates, tol=self.tol, mode=self.mode, device=self.device)
        pt.setup()
        r = pt.solve(res_tol=opts.res_tol, restart=opts.restart, max_iter=opts.gmres_max_iter,
                     preconditioner=opts.preconditioner)
        return {"ecurrent": r["e_noci"], "coeffs": pt.coeffs, "ept2": r["ept2"],
                "a": r["a"].cpu().numpy(), "v": pt.v_omega.cpu().numpy(), "converged": r["converged"],
                "residual": r["residual_rms"]}


def snoci(backend, refs, nmo, opts: SNOCIOptions, log=None):
    """snoci_step (step.rs:178-550) from the reference states `refs`."""
    selected = list(refs)
    pool = None
    final = None
    for it in range(opts.max_iter):
        if pool is None:
            pool = CandidatePool(selected, nmo, opts.orders)
        if not pool.candidates:
            r = backend(selected_dets(selected), [], opts)
            return SNOCIResult(r["ecurrent"], r["coeffs"], selected, [], [])
        r = backend(selected_dets(selected), selected_dets(pool.candidates), opts)
        if not r["converged"]:
            raise RuntimeError(f"SNOCI iteration {it}: GMRES did not converge (residual {r['residual']:.3e})")
        scores = abs(r["a"] * r["v"])
        remaining = max(opts.max_dim - len(selected), 0)
        state = SNOCIResult(r["ecurrent"], r["coeffs"], list(selected), list(pool.candidates), [], r["ept2"])
        if log:
            log(f"SNOCI iteration {it}: ncur={len(selected)} ncand={len(pool.candidates)} "
                f"E={r['ecurrent']:.12f} EPT2={r['ept2']:.3e}")
        if remaining == 0:
            return state
        state.selected = select_candidates(pool.candidates, scores, opts.sigma, min(opts.max_add, remaining))
        if not state.selected or abs(r["ept2"]) < opts.tol:
            return state
        selected = selected + state.selected
        pool.update(selected, state.selected)
        final = state
    return final


def selected_dets(states):
    return [s.det for s in states]
