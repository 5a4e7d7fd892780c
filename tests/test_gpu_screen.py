"""Candidate projection (nw_project_hs) against the oracle, and the learned
screener against exhaustive exact screening on a small selected-NOCI run."""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, screen, synth

pytestmark = pytest.mark.gpu


def test_project_hs_matches_oracle():
    sysm = synth.make_system(14, 3, 2, 3, seed=61, orthonormal_ao=False, n_aux=28, orth_refs=(2,))
    cur = basis.excited_basis(3, 14, 3, 2, per_ref=4, seed=61, include_refs=True)
    cand = basis.excited_basis(3, 14, 3, 2, per_ref=30, seed=62)
    coef = np.random.default_rng(0).normal(size=len(cur))
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    for mode in (0, 1):
        v, s = eng.project_hs(basis.to_records(cur + cand), len(cur), coef, mode=mode)
        slab = O.Slab(sysm)
        # bra = candidate, ket = current (build_noci_hs(data, candidates, selected_space), the
        # orientation that matters for the reference arithmetic on m > 0 pairs)
        Hai = np.array([[slab.hs(a, i, mode=mode)[0] for i in cur] for a in cand])
        Sai = np.array([[slab.hs(a, i, mode=mode)[1] for i in cur] for a in cand])
        vo, so = Hai @ coef, Sai @ coef
        assert np.max(np.abs(v - vo)) < 1e-10 and np.max(np.abs(s - so)) < 1e-10


@pytest.fixture(scope="module")
def small_selection():
    sysm = synth.make_system(24, 3, 3, 6, seed=81)
    cand = screen.random_candidates(6, 24, 3, 3, 2500, seed=82)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    print(f"screen case: n_bas=24 nocc=3 refs=6 candidates={len(cand)}", flush=True)
    return sysm, cand, eng


def test_learned_screening_matches_exhaustive(small_selection):
    """The trained network keeps the exact best candidates inside its top 10%
    (random screening of the same size would keep about 10%), the audit finds
    no miss, and the selected energy is within 1 uHa of exhaustive screening."""
    sysm, cand, eng = small_selection
    kw = dict(iters=5, k_add=16, seed=3)
    ex = screen.select(eng, sysm, cand, learned=False, **kw)
    ln = screen.select(eng, sysm, cand, learned=True, learn_iters=2, top_frac=0.05, audit_frac=0.02, epochs=30,
                       batch=1024, lr=2e-3, measure_recall=True, **kw)
    print(f"screen: exhaustive E={ex.energies[-1]:.10f} learned E={ln.energies[-1]:.10f} recall={ln.recall} "
          f"captured={ln.captured} "
          f"misses={ln.audit_misses} fallbacks={ln.fallbacks} train loss {ln.train['loss'][0]:.3f}->"
          f"{ln.train['loss'][-1]:.3f} evals {ex.exact_evals} vs {ln.exact_evals}", flush=True)
    assert min(ln.recall) >= 0.6
    assert abs(ln.energies[-1] - ex.energies[-1]) < 1e-6  # the screened energy within 1 uHa
    assert sum(ln.exact_evals[2:]) < 0.5 * sum(ex.exact_evals[2:])
