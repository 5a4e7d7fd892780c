"""GPU parity of the generalised-Fock / overlap path (NOCI-PT2 operator M = F - E0 S):
Fock-intermediate refresh, F,S elements and the matrix-free matvec, against the oracle."""
import functools

import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def pt2_case():
    sysm = synth.make_system(12, 3, 2, 3, seed=61, orth_refs=(1,), orth_count=1, orthonormal_ao=False, n_aux=24)
    dets = basis.excited_basis(3, 12, 3, 2, per_ref=30, seed=6, include_refs=True)
    rng = np.random.default_rng(7)
    fa = rng.normal(size=(12, 12)); fa = 0.5 * (fa + fa.T)
    fb = rng.normal(size=(12, 12)); fb = 0.5 * (fb + fb.T)
    print(f"pt2_case: n=12 ndet={len(dets)}", flush=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    eng.update_fock(fa, fb)
    slab = O.Slab(sysm)
    slab.update_fock(sysm, fa, fb)
    return sysm, dets, eng, slab, fa, fb


@pytest.mark.parametrize("mode", [0, 1])
def test_fs_matrix_vs_oracle(pt2_case, mode):
    sysm, dets, eng, slab, fa, fb = pt2_case
    rec = basis.to_records(dets)
    F, S = eng.fs_matrix(rec, mode=mode)
    Fo, So = O.fs_matrix(slab, dets, mode=mode)
    assert np.max(np.abs(S - So)) < TOL
    assert np.max(np.abs(F - Fo)) < TOL


def test_fs_batched_vs_oracle(pt2_case):
    sysm, dets, eng, slab, fa, fb = pt2_case
    rec = basis.to_records(dets)
    rng = np.random.default_rng(8)
    req = rng.integers(0, len(dets), size=(2000, 2)).astype(np.uint64)
    f, s = eng.fs_batched(rec, req, mode=1)
    for q in range(0, len(req), 5):
        fo, so = slab.fs(dets[int(req[q, 0])], dets[int(req[q, 1])], 1)
        assert abs(f[q] - fo) < TOL and abs(s[q] - so) < TOL


def test_s_same_and_m_batched_vs_oracle(pt2_case):
    """Per-spin overlap factors (excitation phases applied) against the
    oracle's xw_overlap factor of each spin, their product against the pair
    overlap, and M = F - e0 S per request against the oracle's F and S, in
    both modes (m = 1 reference pairs included)."""
    sysm, dets, eng, slab, fa, fb = pt2_case
    rec = basis.to_records(dets)
    rng = np.random.default_rng(11)
    req = rng.integers(0, len(dets), size=(1500, 2)).astype(np.uint64)
    sa = eng.s_same_batched(rec, req, 0)
    sb = eng.s_same_batched(rec, req, 1)
    e0 = -2.75
    for mode in (0, 1):
        m = eng.m_batched(rec, req, e0, mode=mode)
        for q in range(0, len(req), 3):
            a, b = dets[int(req[q, 0])], dets[int(req[q, 1])]
            soa, _ = slab.fs_spin(a, b, 0, mode)
            sob, _ = slab.fs_spin(a, b, 1, mode)
            fo, so = slab.fs(a, b, mode)
            assert abs(sa[q] - soa) < TOL and abs(sb[q] - sob) < TOL
            assert abs(sa[q] * sb[q] - so) < TOL
            assert abs(m[q] - (fo - e0 * so)) < TOL
    with pytest.raises(nwick.NwickError):
        eng.s_same_batched(rec, np.array([[0, len(dets)]], dtype=np.uint64), 0)


@pytest.mark.parametrize("mode", [0, 1])
def test_matvec_vs_dense(pt2_case, mode):
    sysm, dets, eng, slab, fa, fb = pt2_case
    rec = basis.to_records(dets)
    Fo, So = O.fs_matrix(slab, dets, mode=mode)
    e0 = -3.25
    rng = np.random.default_rng(9)
    x = rng.normal(size=len(dets))
    x[::7] = 0.0
    y, ys = eng.matvec_m(rec, e0, x, ys=np.empty(len(dets)), mode=mode)
    assert np.max(np.abs(y - (Fo - e0 * So) @ x)) < 1e-9
    assert np.max(np.abs(ys - So @ x)) < 1e-9


def test_fock_refresh_replaces_previous(pt2_case):
    sysm, dets, eng, slab, fa, fb = pt2_case
    rec = basis.to_records(dets)
    eng.update_fock(2.0 * fa, fb)
    F2, _ = eng.fs_matrix(rec, mode=1)
    eng.update_fock(fa, fb)
    F1, _ = eng.fs_matrix(rec, mode=1)
    Fo, _ = O.fs_matrix(slab, dets, mode=1)
    assert np.max(np.abs(F1 - Fo)) < TOL
    assert np.max(np.abs(F2 - F1)) > 1e-3


@functools.lru_cache(maxsize=1)
def _pt2_dense_case():
    """A small PT2 problem on the GPU and its dense operators from the
    oracle's F, S, H blocks with the same Fock matrices."""
    from nwick.pt2 import fock, general_evp, noci_density
    sysm = synth.make_system(10, 2, 2, 3, seed=71, orthonormal_ao=False, n_aux=20)
    refs = basis.excited_basis(3, 10, 2, 2, per_ref=0, include_refs=True, counts={})
    allx = basis.excited_basis(3, 10, 2, 2, per_ref=None, include_refs=False)
    # ref-0 singles and doubles plus ref-1 singles: a linearly independent candidate space
    cands = [d for d in allx if d.parent == 0 or (d.parent == 1 and d.rank_a + d.rank_b == 1)]
    print(f"pt2 dense check: ncur={len(refs)} ncand={len(cands)}", flush=True)
    slab = O.Slab(sysm)
    Hc, Sc = O.hs_matrix(slab, refs, mode=1)
    e, C = general_evp(Hc, Sc)
    c = C[:, 0]
    fa, fb = fock(sysm.h, sysm.eri, *noci_density(sysm, refs, c))
    slab.update_fock(sysm, fa, fb)
    Fc, _ = O.fs_matrix(slab, refs, mode=1)
    e0 = c @ Fc @ c
    allm = refs + cands
    ncur = len(refs)
    Fall, Sall = O.fs_matrix(slab, allm, mode=1)
    Hall, _ = O.hs_matrix(slab, allm, mode=1)
    F_ai, S_ai, H_ai = Fall[ncur:, :ncur], Sall[ncur:, :ncur], Hall[ncur:, :ncur]
    Faa, Saa = Fall[ncur:, ncur:], Sall[ncur:, ncur:]
    s0, f0 = S_ai @ c, F_ai @ c
    M = Faa - e0 * Saa - np.outer(f0, s0) - np.outer(s0, f0) + 2 * e0 * np.outer(s0, s0)
    Q = Saa - np.outer(s0, s0)
    v = H_ai @ c - e[0] * s0
    return sysm, refs, cands, {"e_noci": e[0], "e0": e0, "M": M, "Q": Q, "v": v}


def test_pt2_gmres_vs_dense_solve():
    """GPU PT2 (matrix-free matvec + GPU GMRES) against a dense solve."""
    from nwick.pt2 import PT2
    sysm, refs, cands, dn = _pt2_dense_case()
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    pt = PT2(eng, sysm, refs, cands, mode=1)
    pt.setup()
    # the synthetic Fock operator makes M^Omega indefinite (cond ~2e4): restarted
    # GMRES(30) stagnates, so the test runs unrestarted GMRES (restart >= n)
    r = pt.solve(res_tol=1e-10, restart=400, max_iter=800)
    print(f"pt2 gmres: converged={r['converged']} residual={r['residual_rms']:.3e} iters={r['iterations']}", flush=True)
    assert r["converged"]
    M, v = dn["M"], dn["v"]
    a = np.linalg.solve(M, -v)
    ept2 = a @ M @ a + 2 * a @ v
    print(f"pt2: gpu {r['ept2']:.12f} dense {ept2:.12f} iters {r['iterations']}", flush=True)
    assert abs(r["e0"] - dn["e0"]) < 1e-10 and abs(r["e_noci"] - dn["e_noci"]) < 1e-10
    assert abs(r["ept2"] - ept2) < 1e-8


@pytest.mark.parametrize("operator", ["factorised", "pairwise"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["diag", "woodbury"])
def test_pt2_imag_shift_vs_dense_solve(operator, precision, kind):
    """(M^Omega + i eps Q) a = -V with complex GMRES (imag_shifts) against a
    dense complex solve: E_PT2 = Re(a^H (M + i eps Q) a + 2 a^H V)."""
    from nwick.pt2 import PT2
    if operator == "pairwise" and precision == "mixed":
        pytest.skip("FP32 tables belong to the factorised operator")
    sysm, refs, cands, dn = _pt2_dense_case()
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    pt = PT2(eng, sysm, refs, cands, mode=1, factor_tables="device" if operator == "factorised" else None,
             precision=precision)
    pt.setup()
    eps = 0.5
    r = pt.solve(res_tol=1e-10, restart=60, max_iter=3000, preconditioner=kind, imag_shift=eps)
    Cm = dn["M"] + 1j * eps * dn["Q"]
    v = dn["v"]
    a = np.linalg.solve(Cm, -v)
    ept2 = float(np.real(np.conj(a) @ Cm @ a + 2 * np.conj(a) @ v))
    print(f"{operator} {precision} {kind} eps={eps}: gpu {r['ept2']:.12f} dense {ept2:.12f} residual "
          f"{r['residual_rms']:.2e} iters {r['iterations']} matvecs {r['matvecs']} (FP32 {r['matvecs_fp32']})",
          flush=True)
    assert r["converged"] and r["a"].is_complex()
    assert np.max(np.abs(r["a"].cpu().numpy() - a)) < 1e-7 * max(1.0, float(np.max(np.abs(a))))
    assert abs(r["ept2"] - ept2) < 1e-9


def test_noci_fock_matches_host_restatement():
    """nw_noci_fock (GPU pair densities, D = CC T CC^T, J/K by GEMV) against
    noci_density + fock on the host, with m = 1 reference pairs present."""
    from nwick.pt2 import fock, noci_density
    sysm = synth.make_system(12, 3, 2, 3, seed=75, orthonormal_ao=False, n_aux=24, orth_refs=(2,))
    dets = basis.excited_basis(3, 12, 3, 2, per_ref=6, seed=75, include_refs=True)
    c = np.random.default_rng(1).normal(size=len(dets))
    eng = nwick.Engine.from_system(sysm)
    fa, fb, da, db = eng.noci_fock(basis.to_records(dets), c, tol=1e-8, densities=True)
    Da, Db = noci_density(sysm, dets, c, 1e-8)
    Fa, Fb = fock(sysm.h, sysm.eri, Da, Db)
    print(f"noci_fock: max|dDa|={np.max(np.abs(da - Da)):.3e} max|dFa|={np.max(np.abs(fa - Fa)):.3e}", flush=True)
    for got, want in ((da, Da), (db, Db), (fa, Fa), (fb, Fb)):
        assert np.max(np.abs(got - want)) < 1e-10 * max(1.0, np.max(np.abs(want)))
