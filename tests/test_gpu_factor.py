"""GPU parity of the spin-factorised one-body / overlap operator (nw_factor_*)
against the oracle restatement (oracle/spin_factor.py) and the pairwise
matvec (nw_matvec_m): component ids, per-spin tables for m = 0 and m > 0
reference pairs in both arithmetic modes, y = (F + lambda S) x and S x with
resident and regenerated tables, FP32 tables and their error bound,
diagonals, the stale-table guard, split source groups, and NOCI-PT2
energies through the factorised operator, in FP64 and with FP32 tables
inside FP64 iterative refinement."""
import numpy as np
import pytest

import nwick
import oracle as O
import spin_factor as OF
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def fcase():
    sysm = synth.make_system(12, 3, 2, 3, seed=61, orth_refs=(1,), orth_count=1, orthonormal_ao=False, n_aux=24)
    dets = basis.excited_basis(3, 12, 3, 2, per_ref=30, seed=6, include_refs=True)
    rng = np.random.default_rng(7)
    fa = rng.normal(size=(12, 12)); fa = 0.5 * (fa + fa.T)
    fb = rng.normal(size=(12, 12)); fb = 0.5 * (fb + fb.T)
    print(f"factor case: n=12 ndet={len(dets)}", flush=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    eng.update_fock(fa, fb)
    slab = O.Slab(sysm)
    slab.update_fock(sysm, fa, fb)
    fac = OF.SpinFactorisation(dets, sysm.nref)
    return sysm, dets, eng, slab, fac, fa, fb


def test_ids_match_oracle(fcase):
    sysm, dets, eng, slab, fac, fa, fb = fcase
    f = eng.factor(basis.to_records(dets))
    a, b = f.ids()
    assert np.array_equal(a, fac.aid) and np.array_equal(b, fac.bid)
    for p in range(sysm.nref):
        info = f.info(p)
        assert (info["ncomp_a"], info["ncomp_b"]) == (fac.ncomp(p, 0), fac.ncomp(p, 1))
    f.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_tables_vs_oracle(fcase, mode):
    sysm, dets, eng, slab, fac, fa, fb = fcase
    f = eng.factor(basis.to_records(dets), mode=mode)
    assert any(m != (0, 0) for m in slab.m.values())
    tabs = OF.tables(slab, dets, fac, mode)
    worst = 0.0
    for (q, p, sp), ref in tabs.items():
        got = f.table(q, p, sp)
        worst = max(worst, float(np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref)))))
    print(f"mode {mode}: max table deviation {worst:.2e}", flush=True)
    assert worst < TOL
    f.close()


@pytest.mark.parametrize("storage", [nwick.FACTOR_DEVICE, nwick.FACTOR_TRANSIENT])
@pytest.mark.parametrize("mode", [0, 1])
def test_apply_vs_oracle_and_pairwise(fcase, mode, storage):
    sysm, dets, eng, slab, fac, fa, fb = fcase
    rec = basis.to_records(dets)
    f = eng.factor(rec, mode=mode, storage=storage)
    tabs = OF.tables(slab, dets, fac, mode)
    rng = np.random.default_rng(11)
    x = rng.normal(size=len(dets))
    x[::5] = 0.0
    lam = 2.75
    y, ys = f.apply(lam, x, ys=np.empty(len(dets)))
    yo, yso = OF.apply(fac, tabs, lam, x)
    scale = max(1.0, float(np.max(np.abs(yo))))
    assert np.max(np.abs(y - yo)) < 1e-9 * scale
    assert np.max(np.abs(ys - yso)) < 1e-9 * scale
    # the pairwise operator (bra = lower index) agrees: same-parent pairs have m = 0
    yp, ysp = eng.matvec_m(rec, -lam, x, ys=np.empty(len(dets)), mode=mode)
    assert np.max(np.abs(y - yp)) < 1e-9 * scale
    assert np.max(np.abs(ys - ysp)) < 1e-9 * scale
    f.close()


@pytest.mark.parametrize("storage", [nwick.FACTOR_DEVICE, nwick.FACTOR_TRANSIENT])
def test_fp32_tables_and_apply(fcase, storage):
    """FP32 tables are the FP64 tables rounded to nearest (bit-exact copies
    for resident tables; regenerated ones are evaluated in FP64 and rounded,
    so within one FP32 ulp).  The FP32 apply differs from the FP64 apply by
    at most 8 * 2^-24 of the absolute sum of the terms, |Fa Sb| + |Sa Fb| +
    |lam Sa Sb| over |x| (table rounding on both factors plus the rounding
    of materialised rows), computed through the oracle on absolute tables."""
    sysm, dets, eng, slab, fac, fa, fb = fcase
    rec = basis.to_records(dets)
    f = eng.factor(rec, mode=1, storage=storage)
    keys = [(q, p, sp) for q in range(sysm.nref) for p in range(sysm.nref) for sp in (0, 1)]
    t64 = {k: f.table(*k) for k in keys}
    rng = np.random.default_rng(12)
    x = rng.normal(size=len(dets))
    lam = -1.9
    y64, ys64 = f.apply(lam, x, ys=np.empty(len(dets)))
    bytes64, abytes64 = f.info()["table_bytes"], f.info()["apply_bytes"]
    assert 0 < abytes64 <= bytes64
    f.set_precision(nwick.PREC_MIXED)
    assert 2 * f.info()["table_bytes"] == bytes64 and 2 * f.info()["apply_bytes"] == abytes64
    for k in keys:
        got, want = f.table(*k), t64[k].astype(np.float32).astype(np.float64)
        if storage == nwick.FACTOR_DEVICE:
            assert np.array_equal(got, want), k
        else:
            assert np.all(np.abs(got - want) <= 2.0 ** -23 * np.abs(want)), k
    y32, ys32 = f.apply(lam, x, ys=np.empty(len(dets)))
    tabs = OF.tables(slab, dets, fac, 1)
    yabs, ysabs = OF.apply(fac, {k: np.abs(v) for k, v in tabs.items()}, abs(lam), np.abs(x))
    ey, es = np.abs(y32 - y64) / yabs, np.abs(ys32 - ys64) / ysabs
    print(f"storage {storage}: FP32 apply error / absolute term sum: y {ey.max():.2e}, ys {es.max():.2e} "
          f"(2^-24 = {2.0 ** -24:.2e})", flush=True)
    assert np.all(np.abs(y32 - y64) <= 8 * 2.0 ** -24 * yabs + 1e-300)
    assert np.all(np.abs(ys32 - ys64) <= 8 * 2.0 ** -24 * ysabs + 1e-300)
    assert ey.max() > 0  # the FP32 tables were read
    # back to FP64: the FP64 result again, bit for bit
    f.set_precision(nwick.PREC_FP64)
    y2 = f.apply(lam, x)
    assert np.array_equal(y2, y64)
    with pytest.raises(nwick.NwickError):
        f.set_precision(nwick.PREC_BF16X3)
    f.close()


@pytest.mark.parametrize("storage", [nwick.FACTOR_DEVICE, nwick.FACTOR_TRANSIENT])
@pytest.mark.parametrize("prec", [nwick.PREC_FP64, nwick.PREC_MIXED])
def test_apply_pair_equals_two_applies(fcase, storage, prec):
    """Two interleaved vectors in one pass give, bit for bit, what two
    separate applies give (same summation order per vector), for FP64 and
    FP32 tables; before and after a single apply shares the buffers."""
    import torch
    sysm, dets, eng, slab, fac, fa, fb = fcase
    rec = basis.to_records(dets)
    f = eng.factor(rec, mode=1, storage=storage)
    f.set_precision(prec)
    n = len(dets)
    rng = np.random.default_rng(21)
    xr, xi = rng.normal(size=n), rng.normal(size=n)
    lam = 0.7
    yr, ysr = f.apply(lam, xr, ys=np.empty(n))
    yi, ysi = f.apply(lam, xi, ys=np.empty(n))
    y2, ys2 = f.apply_pair(lam, xr + 1j * xi, y=np.empty(n, dtype=complex), ys=np.empty(n, dtype=complex))
    assert np.array_equal(y2.real, yr) and np.array_equal(y2.imag, yi)
    assert np.array_equal(ys2.real, ysr) and np.array_equal(ys2.imag, ysi)
    # device tensors, then a single apply again (plans of both vector counts coexist)
    xt = torch.tensor(xr + 1j * xi, device="cuda:0")
    yt = torch.empty_like(xt)
    f.apply_pair(lam, xt, y=yt)
    assert np.array_equal(yt.cpu().numpy(), y2)
    assert np.array_equal(f.apply(lam, xr), yr)
    f.close()


def test_fp32_tables_follow_refresh(fcase):
    sysm, dets, eng, slab, fac, fa, fb = fcase
    rec = basis.to_records(dets)
    f = eng.factor(rec)
    f.set_precision(nwick.PREC_MIXED)
    x = np.random.default_rng(3).normal(size=len(dets))
    eng.update_fock(fa, 3.0 * fb)
    with pytest.raises(nwick.NwickError):
        f.apply(0.5, x)
    f.refresh()
    y32 = f.apply(0.5, x)
    f.set_precision(nwick.PREC_FP64)
    y64 = f.apply(0.5, x)
    scale = float(np.max(np.abs(y64)))
    assert np.max(np.abs(y32 - y64)) < 1e-5 * scale
    eng.update_fock(fa, fb)
    f.close()


def test_diagonals(fcase):
    sysm, dets, eng, slab, fac, fa, fb = fcase
    f = eng.factor(basis.to_records(dets))
    lam = -1.5
    md, sd = f.diagonals(lam)
    for i in range(len(dets)):
        fo, so = slab.fs(dets[i], dets[i], 0)
        assert abs(sd[i] - so) < TOL and abs(md[i] - (fo + lam * so)) < TOL * max(1.0, abs(fo))
    f.close()


def test_stale_tables_and_refresh(fcase):
    sysm, dets, eng, slab, fac, fa, fb = fcase
    rec = basis.to_records(dets)
    f = eng.factor(rec)
    x = np.random.default_rng(2).normal(size=len(dets))
    eng.update_fock(2.0 * fa, fb)
    with pytest.raises(nwick.NwickError) as ei:
        f.apply(0.0, x)
    assert ei.value.code == 6
    f.refresh()
    y = f.apply(0.0, x)
    yp = eng.matvec_m(rec, 0.0, x)
    assert np.max(np.abs(y - yp)) < 1e-9 * max(1.0, float(np.max(np.abs(yp))))
    eng.update_fock(fa, fb)
    f.refresh()
    f.close()


@pytest.mark.parametrize("storage", [nwick.FACTOR_DEVICE, nwick.FACTOR_TRANSIENT])
def test_split_groups_vs_pairwise(storage):
    """All singles and doubles of two references: the beta-reference group of
    every parent holds several hundred entries, so stage 1 splits it into
    pieces and reduces them; checked against the pairwise matvec."""
    sysm = synth.make_system(16, 4, 4, 2, seed=5, kappa_sigma=0.3)
    dets = basis.excited_basis(2, 16, 4, 4, per_ref=None, include_refs=True)
    print(f"split case: ndet={len(dets)}", flush=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    f = eng.factor(rec, storage=storage)
    counts = [f.info(p) for p in range(2)]
    assert all(c["ncomp_a"] > 64 for c in counts)
    # the reference row and column of each parent are summed by dual passes
    # (each table once for both directions, same-parent tables over their
    # upper triangle): an apply reads well under the resident table bytes
    info = f.info()
    if storage == nwick.FACTOR_DEVICE:
        assert info["apply_bytes"] < 0.8 * info["table_bytes"], info
    rng = np.random.default_rng(4)
    x = rng.normal(size=len(dets))
    lam = -0.8
    y, ys = f.apply(lam, x, ys=np.empty(len(dets)))
    yp, ysp = eng.matvec_m(rec, -lam, x, ys=np.empty(len(dets)))
    scale = max(1.0, float(np.max(np.abs(yp))))
    assert np.max(np.abs(y - yp)) < 1e-9 * scale
    assert np.max(np.abs(ys - ysp)) < 1e-9 * scale
    f.close()


def test_pt2_factorised_equals_pairwise():
    """E_PT2 through resident tables, regenerated tables and the pairwise
    matvec.  The synthetic Fock operator makes M^Omega indefinite, so the
    solve is unrestarted GMRES (restart >= n) as in the dense PT2 check."""
    from nwick.pt2 import PT2
    sysm = synth.make_system(10, 2, 2, 3, seed=71, orthonormal_ao=False, n_aux=20)
    refs = basis.excited_basis(3, 10, 2, 2, per_ref=0, include_refs=True, counts={})
    allx = basis.excited_basis(3, 10, 2, 2, per_ref=None, include_refs=False)
    cands = [d for d in allx if d.parent == 0 or (d.parent == 1 and d.rank_a + d.rank_b == 1)]
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    out = {}
    for ft in ("device", "transient", None):
        pt = PT2(eng, sysm, refs, cands, mode=1, factor_tables=ft)
        pt.setup()
        assert pt.operator == ("pairwise" if ft is None else "factorised")
        r = pt.solve(res_tol=1e-10, restart=400, max_iter=800)
        print(f"factor_tables={ft}: E_PT2 {r['ept2']:.12f} iters {r['iterations']} converged {r['converged']}",
              flush=True)
        assert r["converged"]
        out[ft] = r["ept2"]
    assert abs(out["device"] - out[None]) < 1e-9
    assert abs(out["transient"] - out[None]) < 1e-9


@pytest.mark.parametrize("storage", ["device", "transient"])
def test_pt2_mixed_refinement_equals_fp64(storage):
    """E_PT2 from GMRES whose Arnoldi matvecs read the FP32 tables and whose
    restart residuals use the FP64 tables equals the FP64 E_PT2, and its
    FP64 residual meets the tolerance.  This system is indefinite and
    ill-conditioned, so each FP64 restart residual gains only a few digits
    over the previous one; the iteration limit allows for that."""
    from nwick.pt2 import PT2
    sysm = synth.make_system(10, 2, 2, 3, seed=71, orthonormal_ao=False, n_aux=20)
    refs = basis.excited_basis(3, 10, 2, 2, per_ref=0, include_refs=True, counts={})
    allx = basis.excited_basis(3, 10, 2, 2, per_ref=None, include_refs=False)
    cands = [d for d in allx if d.parent == 0 or (d.parent == 1 and d.rank_a + d.rank_b == 1)]
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    res = {}
    for prec in ("fp64", "mixed"):
        pt = PT2(eng, sysm, refs, cands, mode=1, factor_tables=storage, precision=prec)
        pt.setup()
        restarts = []
        r = pt.solve(res_tol=1e-10, restart=400, max_iter=4000,
                     log=lambda m: restarts.append(m.strip()) if "restart" in m else None)
        print(f"{storage} {prec}: E_PT2 {r['ept2']:.12f} residual {r['residual_rms']:.2e} iters {r['iterations']} "
              f"FP32 matvecs {r['matvecs_fp32']} of {r['matvecs']}; " + "; ".join(restarts), flush=True)
        assert r["converged"] and r["residual_rms"] <= 1e-10
        res[prec] = r
    assert res["mixed"]["matvecs_fp32"] > 0 and res["fp64"]["matvecs_fp32"] == 0
    assert abs(res["mixed"]["ept2"] - res["fp64"]["ept2"]) < 1e-9


@pytest.mark.parametrize("storage", [nwick.FACTOR_DEVICE, nwick.FACTOR_TRANSIENT])
def test_block_products_equal_per_group_sums(monkeypatch, storage):
    """The single x single groups of each parent as one dense product on FP64
    MFMA (k_fac_block, one-vector applies) against the per-group sums
    (NW_FAC_BLOCK=0): FP64 and FP32 tables, and against the pairwise matvec;
    two-vector applies are the per-group sums either way."""
    sysm = synth.make_system(16, 4, 4, 2, seed=5, kappa_sigma=0.3)
    dets = basis.excited_basis(2, 16, 4, 4, per_ref=None, include_refs=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    rng = np.random.default_rng(21)
    x = rng.normal(size=len(dets))
    x2 = rng.normal(size=(len(dets), 2))
    lam = -0.6
    monkeypatch.setenv("NW_FAC_BLOCK", "0")
    f0 = eng.factor(rec, storage=storage)
    monkeypatch.delenv("NW_FAC_BLOCK")
    f1 = eng.factor(rec, storage=storage)
    i0, i1 = f0.info(), f1.info()
    print(f"storage {storage}: {i1['block_groups']} groups in {i1['blocks']} dense blocks", flush=True)
    assert i0["block_groups"] == 0 and i1["block_groups"] >= 32 and i1["blocks"] >= 2
    yp, ysp = eng.matvec_m(rec, -lam, x, ys=np.empty(len(dets)))
    scale = max(1.0, float(np.max(np.abs(yp))))
    for prec in (nwick.PREC_FP64, nwick.PREC_MIXED):
        f0.set_precision(prec)
        f1.set_precision(prec)
        y0, ys0 = f0.apply(lam, x, ys=np.empty(len(dets)))
        y1, ys1 = f1.apply(lam, x, ys=np.empty(len(dets)))
        d = max(float(np.max(np.abs(y1 - y0))), float(np.max(np.abs(ys1 - ys0))))
        print(f"prec {prec}: max |block - per-group| {d:.2e} (scale {scale:.1f})", flush=True)
        # FP64 sums in another order; with FP32 tables the materialised rows
        # are rounded to FP32, so the two orders may round them differently
        rel = 1e-12 if prec == nwick.PREC_FP64 else 2.0 ** -20
        assert d < rel * scale
        if prec == nwick.PREC_FP64:
            assert np.max(np.abs(y1 - yp)) < 1e-9 * scale and np.max(np.abs(ys1 - ysp)) < 1e-9 * scale
        # two-vector applies keep the per-group sums in both
        z0 = f0.apply_pair(lam, x2)
        z1 = f1.apply_pair(lam, x2)
        assert np.array_equal(z1, z0)
        assert np.max(np.abs(z1[:, 0] - f1.apply(lam, np.ascontiguousarray(x2[:, 0])))) < rel * scale
    f0.close()
    f1.close()


def test_engine_closed_before_its_factor():
    """Closing the engine frees its factor objects; closing such a factor
    afterwards is a no-op (either order of destruction is safe)."""
    sysm = synth.make_system(8, 2, 2, 2, seed=3, n_aux=16)
    dets = basis.excited_basis(2, 8, 2, 2, per_ref=6, seed=3, include_refs=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    f = eng.factor(basis.to_records(dets))
    g = eng.factor(basis.to_records(dets))
    g.close()
    eng.close()
    f.close()
    assert f.h is None
