"""Streamed build -> evaluate -> reuse over groups of reference pairs
(nw_hs_matrix_streamed, the engine's counterpart of a disk-backed slab):
the same H, S as one full build whatever the grouping, against the oracle;
a budget below one pair's block is refused; and a 16-reference n_bas = 100,
10+10-electron system whose FP64 slab (about 340 GB) exceeds one device's
memory, checked element by element against the oracle on four reference
pairs."""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import basis, synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _greedy_groups(sizes, budget):
    groups, cur = [], 0
    for b in sizes:
        if not groups or cur + b > budget:
            groups.append(0)
            cur = 0
        groups[-1] += 1
        cur += b
    return len(groups)


@pytest.fixture(scope="module")
def small():
    n, na, nb, nref = 16, 3, 3, 5
    sysm = synth.make_system(n, na, nb, nref, seed=17, orth_refs=(3,), orth_count=1, orthonormal_ao=False, n_aux=40)
    dets = basis.excited_basis(nref, n, na, nb, per_ref=24, seed=17, include_refs=True)
    print(f"streamed: n={n} nocc={na}+{nb} nref={nref} ndet={len(dets)}", flush=True)
    return sysm, dets


@pytest.mark.parametrize("mode", [0, 1])
def test_streamed_equals_full_build(small, mode):
    sysm, dets = small
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    meta = eng.build(tol=1e-8)
    assert any(m.ma == 1 for m in meta.values())
    H0, S0 = eng.hs_matrix(rec, mode=mode)
    nref = sysm.nref
    # block bytes of the upper-triangle pairs in (x, w) order
    sizes = [int(eng.pair_slab(x, w)[1][94]) * 8 for x in range(nref) for w in range(x, nref)]
    Ho, So = O.hs_matrix(O.Slab(sysm, tol=1e-8), dets, mode=mode)
    assert np.max(np.abs(H0 - Ho)) < TOL and np.max(np.abs(S0 - So)) < TOL
    for budget in (max(sizes), 3 * max(sizes), sum(sizes)):
        H, S, groups = eng.hs_matrix_streamed(rec, tol=1e-8, mode=mode, slab_bytes=budget)
        dh, ds = float(np.max(np.abs(H - H0))), float(np.max(np.abs(S - S0)))
        print(f"mode {mode} budget {budget}: {groups} groups, max|dH| {dh:.2e} max|dS| {ds:.2e}", flush=True)
        assert groups == _greedy_groups(sizes, budget)
        assert dh < 1e-12 * max(1.0, np.max(np.abs(H0))) and ds < 1e-12
    # the context keeps the last group's build: pairs of earlier groups are gone
    H, S, groups = eng.hs_matrix_streamed(rec, tol=1e-8, mode=mode, slab_bytes=max(sizes))
    assert groups > 1
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_matrix(rec, mode=mode)
    assert e.value.code == 6  # NW_E_STATE: reference pair not built
    eng.close()


def test_streamed_mixed_precision_equals_full_build(small):
    sysm, dets = small
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.set_precision(nwick.PREC_BF16X3)
    eng.set_mixed_floor(0.0)
    meta = eng.build(tol=1e-8)
    assert all(m.fp32 for m in meta.values())
    H0, S0 = eng.hs_matrix(rec)
    nref = sysm.nref
    blk = [int(eng.pair_slab(x, w)[1][94]) * 8 for x in range(nref) for w in range(x, nref)]
    # FP32 pairs keep half-length rank-4 blocks: a budget of the largest FP64
    # block holds at least one pair
    smeta = {}
    H, S, groups = eng.hs_matrix_streamed(rec, tol=1e-8, slab_bytes=max(blk), meta=smeta)
    print(f"bf16x3 streamed: {groups} groups, max|dH| {np.max(np.abs(H - H0)):.2e}", flush=True)
    assert groups > 1
    assert sorted(smeta) == [(x, w) for x in range(nref) for w in range(x, nref)]
    for k, m in smeta.items():
        assert (m.ma, m.mb, m.fp32, m.sigma_min) == (meta[k].ma, meta[k].mb, meta[k].fp32, meta[k].sigma_min)
    assert np.max(np.abs(H - H0)) < 1e-12 * max(1.0, np.max(np.abs(H0))) and np.max(np.abs(S - S0)) < 1e-12
    eng.close()


def test_budget_below_one_pair_is_refused(small):
    sysm, dets = small
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_matrix_streamed(rec, tol=1e-8, slab_bytes=1 << 20)
    assert e.value.code == 4  # NW_E_MEM
    eng.close()


def test_sixteen_references_n100_beyond_device_memory():
    nref = 16
    # reference 1's first alpha orbital is orthogonal to the occupied spaces
    # of references 0, 2, 3, 4 (16 x 10 occupied orbitals span the whole basis)
    sysm = synth.make_system(100, 10, 10, nref, seed=2025, n_aux=300, orth_refs=(1,), orth_count=1,
                             orth_against=(0, 2, 3, 4))
    counts = {"a": 2, "b": 2, "aa": 3, "bb": 3, "ab": 4}
    dets = basis.excited_basis(nref, 100, 10, 10, per_ref=14, seed=2025, include_refs=True, counts=counts)
    print(f"n_bas=100: 10+10 e, {nref} refs, ndet={len(dets)}, 136 upper-triangle pairs, FP64", flush=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    H, S, groups = eng.hs_matrix_streamed(rec, tol=1e-8, mode=0)
    t = eng.timings()
    print(f"streamed: {groups} groups, total {t['total_ms']:.0f} ms (builds {t['transform_ms']:.0f} ms, "
          f"evaluations {t['eval_ms']:.1f} ms)", flush=True)
    eng.close()
    assert groups >= 2
    assert np.allclose(H, H.T, rtol=0, atol=0) and np.allclose(S, S.T, rtol=0, atol=0)
    # two m = 1 pairs, a same-reference pair (elements of order 100 Ha) and a
    # pair of different references, each in a different group or position
    pairs = [(0, 1), (1, 4), (9, 9), (6, 11)]
    slab = O.Slab(sysm, tol=1e-8, pairs=pairs)
    assert slab.m[(0, 1)] == (1, 0) and slab.m[(1, 4)] == (1, 0) and slab.m[(6, 11)] == (0, 0)
    par = np.array([d.parent for d in dets])
    for x, w in pairs:
        dh = ds = hmax = 0.0
        for i in np.flatnonzero(par == x):
            for j in np.flatnonzero(par == w):
                if i > j:
                    continue
                ho, so = slab.hs(dets[i], dets[j], 0)
                dh, ds, hmax = max(dh, abs(H[i, j] - ho)), max(ds, abs(S[i, j] - so)), max(hmax, abs(ho))
        print(f"oracle pair ({x},{w}): max|H| {hmax:.3e} max|dH| {dh:.2e} max|dS| {ds:.2e}", flush=True)
        assert dh < TOL * max(1.0, hmax) and ds < TOL
