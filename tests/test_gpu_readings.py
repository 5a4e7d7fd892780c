"""The GPU's reference-arithmetic m > 0 path against the pinned oracle.

- Skip readings on the device: a slab block with one contraction entry set to
  zero (written into a saved wicks.bin and loaded back through the disk
  cache) makes a first minor or a whole distribution vanish; the GPU's
  reference mode skips exactly what the oracle's readings skip, checked
  against the oracle on the same block and the numpy restatement.
- The n_bas = 100 shape of the headline workload in FP64: 10 + 10 electrons,
  two references (one with an occupied alpha orbital exactly orthogonal to
  the other reference, so the cross pair has m_alpha = 1), 40 determinants
  each, both modes, against the oracle at 1e-10 per element.
- A determinant whose parent index does not fit 16 bits is an argument
  error, and the grid-stride tile loop of the evaluation kernels gives the
  same H, S with a grid capped far below the tile count.
"""
import os

import numpy as np
import pytest

import m_readings as R
import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _zero_in_cache(eng, directory, x, w, r, c):
    """Set X0_alpha[r, c] of pair (x, w) to zero in a saved wicks.bin."""
    path = os.path.join(directory, "wicks.bin")
    slab = np.fromfile(path, dtype="<f8")
    base = 0
    for p in range(x * eng.nref + w):
        base += int(eng.pair_slab(*divmod(p, eng.nref))[1][94])
    off = eng.pair_slab(x, w)[1]
    slab[base + off[R.O_X] + r * eng.nmo + c] = 0.0
    slab.tofile(path)


def _zero_in_oracle(slab, x, w, r, c):
    import copy
    out = copy.deepcopy(slab)
    out.slab[out.base[(x, w)] + out.off[(x, w)][R.O_X] + r * slab.nmo + c] = 0.0
    return out


@pytest.mark.parametrize("kind", ["minor", "dist"])
def test_skip_readings_on_device(tmp_path, kind):
    sysm = synth.make_system(8, 3, 3, 2, seed=71, orth_refs=(1,), orth_count=1, orthonormal_ao=False,
                             n_aux=16, enuc=1.7)
    dets = basis.excited_basis(2, 8, 3, 3, per_ref=None, include_refs=True)
    eng = nwick.Engine.from_system(sysm)
    meta = eng.build(tol=1e-8)
    assert (meta[(0, 1)].ma, meta[(0, 1)].mb) == (1, 0)
    oslab = O.Slab(sysm, tol=1e-8)
    rank = 2 if kind == "minor" else 1
    bras = [k for k, d in enumerate(dets) if d.parent == 0 and d.rank_a == rank and d.rank_b <= 1][:4]
    kets = [k for k, d in enumerate(dets) if d.parent == 1 and d.rank_a == 0 and d.rank_b <= 1][:3]
    rec = basis.to_records(dets)
    for ia in bras:
        a = dets[ia]
        rows = [p - 3 for p in R.bits_of(a.pa)]
        cols = R.bits_of(a.ha)
        r, c = (rows[1], cols[0]) if kind == "minor" else (rows[0], cols[0])
        d = tmp_path / f"{kind}{ia}"
        d.mkdir()
        eng.save_wicks(str(d))
        _zero_in_cache(eng, str(d), 0, 1, r, c)
        eng2 = nwick.Engine.from_system(sysm)
        eng2.load_wicks(str(d), tol=1e-8)
        mod = _zero_in_oracle(oslab, 0, 1, r, c)
        req = np.array([[ia, ib] for ib in kets], dtype=np.uint64)
        for mode in (0, 1):
            h, s = eng2.hs_batched(rec, req, mode=mode)
            for q, ib in enumerate(kets):
                ho, so = mod.hs(a, dets[ib], mode)
                hn, sn = R.element(mod.pair_view(0, 1), 8, a, dets[ib], O.RD_ALL if mode == 0 else 0,
                                   sysm.enuc, 1e-8)
                assert abs(h[q] - ho) < TOL and abs(s[q] - so) < TOL
                assert abs(hn - ho) < TOL
        # the planted zero is what makes the two modes' skip differ here
        skip = O.RD_MINOR if kind == "minor" else O.RD_DIST
        assert any(abs(mod.hs_readings(a, dets[ib], skip)[0] - mod.hs_readings(a, dets[ib], 0)[0]) > 1e-7
                   for ib in kets)
        eng2.close()


@pytest.fixture(scope="module")
def n100():
    sysm = synth.make_system(100, 10, 10, 2, seed=2024, n_aux=300, orth_refs=(1,), orth_count=1)
    counts = {"a": 4, "b": 4, "aa": 10, "bb": 10, "ab": 11}
    dets = basis.excited_basis(2, 100, 10, 10, per_ref=39, seed=2024, include_refs=True, counts=counts)
    print(f"n_bas=100: 10+10 e, 2 refs, ndet={len(dets)}; oracle build of pairs (0,0) (0,1) (1,1)", flush=True)
    pairs = [(0, 0), (0, 1), (1, 1)]
    slab = O.Slab(sysm, tol=1e-8, pairs=pairs)
    return sysm, dets, slab


@pytest.mark.parametrize("mode", [0, 1])
def test_n100_fp64_vs_oracle(n100, mode):
    sysm, dets, slab = n100
    assert slab.m[(0, 1)] == (1, 0) and slab.m[(0, 0)] == (0, 0)
    eng = nwick.Engine.from_system(sysm)
    mask = np.array([[1, 1], [0, 1]], dtype=np.uint8)
    meta = eng.build(tol=1e-8, pair_mask=mask)
    assert (meta[(0, 1)].ma, meta[(0, 1)].mb) == (1, 0)
    rec = basis.to_records(dets)
    H, S = eng.hs_matrix(rec, mode=mode)
    Ho, So = O.hs_matrix(slab, dets, mode=mode)
    dh, ds = float(np.max(np.abs(H - Ho))), float(np.max(np.abs(S - So)))
    print(f"n_bas=100 mode {mode}: max|dH|={dh:.3e} max|dS|={ds:.3e} max|H|={np.max(np.abs(Ho)):.3f}", flush=True)
    assert ds < TOL and dh < TOL
    eng.close()


def test_parent_beyond_16_bits_is_an_argument_error():
    sysm = synth.make_system(8, 2, 2, 2, seed=3, n_aux=16)
    dets = basis.excited_basis(2, 8, 2, 2, per_ref=4, seed=3, include_refs=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    rec = basis.to_records(dets)
    rec["parent"][1] = 65536  # 65536 mod 2^16 = 0 would alias a valid parent
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_matrix(rec)
    assert e.value.code == 1  # NW_E_ARG
    with pytest.raises(nwick.NwickError):
        eng.hs_batched(rec, np.array([[0, 1]], dtype=np.uint64))
    eng.close()


def test_grid_stride_tiles(monkeypatch):
    sysm = synth.make_system(12, 3, 2, 3, seed=8, orth_refs=(2,), n_aux=24)
    dets = basis.excited_basis(3, 12, 3, 2, per_ref=60, seed=8, include_refs=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    monkeypatch.delenv("NW_EVAL_GRID_CAP", raising=False)
    H, S = eng.hs_matrix(rec, mode=0)
    F, S2 = eng.fs_matrix(rec, mode=0)
    x = np.random.default_rng(1).normal(size=len(dets))
    y, ys = eng.matvec_m(rec, 0.3, x, ys=np.empty(len(dets)))
    monkeypatch.setenv("NW_EVAL_GRID_CAP", "3")
    Hc, Sc = eng.hs_matrix(rec, mode=0)
    Fc, S2c = eng.fs_matrix(rec, mode=0)
    yc, ysc = eng.matvec_m(rec, 0.3, x, ys=np.empty(len(dets)))
    assert np.array_equal(H, Hc) and np.array_equal(S, Sc) and np.array_equal(F, Fc) and np.array_equal(S2, S2c)
    assert np.max(np.abs(y - yc)) < 1e-12 * max(1.0, np.max(np.abs(y)))
    assert np.max(np.abs(ys - ysc)) < 1e-12 * max(1.0, np.max(np.abs(ys)))
    eng.close()
