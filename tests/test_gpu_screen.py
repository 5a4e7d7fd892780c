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


def test_project_hs_is_order_independent():
    """Candidate projections are summed without atomics (per block and row,
    then over the row's blocks in column-run order): repeated calls, and a
    call on the current space listed in another run order, give the same
    bits for every candidate; the values match the dense H, S of the same
    pairs (bra = candidate, as the projection orients them: the references of
    this system overlap weakly, so the two orientations differ by rounding
    amplified to ~1e-8) times the coefficients."""
    sysm = synth.make_system(20, 3, 3, 4, seed=63, n_aux=40)
    cur = basis.excited_basis(4, 20, 3, 3, per_ref=40, seed=63, include_refs=True)
    cand = basis.excited_basis(4, 20, 3, 3, per_ref=400, seed=64)
    coef = np.random.default_rng(1).normal(size=len(cur))
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    rec = basis.to_records(cur + cand)
    v1, s1 = eng.project_hs(rec, len(cur), coef)
    for _ in range(3):
        v2, s2 = eng.project_hs(rec, len(cur), coef)
        assert np.array_equal(v1, v2) and np.array_equal(s1, s2)
    n, nc = len(cur), len(cand)
    req = np.array([(n + a, i) for a in range(nc) for i in range(n)], dtype=np.uint64)
    h, sv = eng.hs_batched(rec, req)
    print(f"projection: {nc} candidates x {n} current", flush=True)
    assert np.max(np.abs(v1 - h.reshape(nc, n) @ coef)) < 1e-10
    assert np.max(np.abs(s1 - sv.reshape(nc, n) @ coef)) < 1e-10
    eng.close()


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


def test_screened_pt2_matches_full():
    """PT2 candidates ranked by the screener, on a real molecule (H6 chain,
    6-31G, 1.5 Angstrom spacing; RHF and two spin-biased UHF references; all
    2,997 singles and doubles): trained on the exact first-order weights
    |<D|H - E S|Psi>|^2 / |E - H_DD| of 10% of the candidates, the network
    orders the rest, which are scored exactly in that order until the
    unscored weight is below tau (judged by a random audit, with an
    exhaustive exact fallback), and the PT2 equations (imaginary shift 1)
    are solved over the kept candidates only.  With tau = 3 mHa of weight a
    third of the candidates is dropped and E_PT2 moves by far less than tau;
    the PT2 over every candidate is the exact judge."""
    from nwick.molecule import Molecule
    from nwick.pt2 import PT2
    from nwick.scf import ao_data, reference_states, to_system
    recipes = [{"label": "RHF (0, 0)"},
               {"label": "UHF (1, -1)", "spin_bias": {"pattern": [1, -1, 1, -1, 1, -1], "pol": 0.75}},
               {"label": "UHF (-1, 1)", "spin_bias": {"pattern": [-1, 1, -1, 1, -1, 1], "pol": 0.75}}]
    ao = ao_data(Molecule([("H", (0, 0, 1.5 * k)) for k in range(6)], "6-31g"))
    sysm = to_system(ao, reference_states(ao, recipes))
    dets = basis.excited_basis(sysm.nref, sysm.nmo, sysm.nocc_a, sysm.nocc_b, per_ref=None, orders=(1, 2))
    cand = screen.candidates_from_dets(dets, sysm.nocc_a, sysm.nocc_b)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    kw = dict(res_tol=1e-7, restart=40, max_iter=800, imag_shift=1.0)
    refs = [basis.Det(i, (1 << sysm.nocc_a) - 1, (1 << sysm.nocc_b) - 1) for i in range(sysm.nref)]
    full = PT2(eng, sysm, refs, dets, tol=1e-8)
    full.setup()
    rf = full.solve(**kw)
    full.fac.close()
    tau = 3e-3
    r = screen.screened_pt2(eng, sysm, cand, tau=tau, seed=5, solve_kw=kw)
    print(f"screened PT2: full E_PT2 {rf['ept2']:.10f} ({len(cand)} candidates), screened {r.ept2:.10f} "
          f"({len(r.kept)} kept, {r.exact_evals} exact scores, tail {r.tail_estimate:.2e}, audit "
          f"{r.audit_tail:.2e}, fallback {r.fallback}) {r.seconds}", flush=True)
    assert rf["converged"] and r.solve["converged"]
    assert abs(r.ept2 - rf["ept2"]) < tau
    assert abs(r.ept2 - rf["ept2"]) < 1e-4
    assert len(r.kept) < 0.75 * len(cand)
    eng.close()
