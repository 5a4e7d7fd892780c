"""Host logic of the learned screener: vectorised candidate records against
the per-determinant generator (masks, occupations, excitation phases), and
the candidate pick and k-th best score against numpy's stable sort."""
import numpy as np

from nwick import screen
from nwick.basis import Det, excitation_phase_bits, to_records


def test_vectorised_records_match_det_records():
    c = screen.random_candidates(3, 30, 5, 4, 400, seed=9)
    assert len(c) == 3 * 400
    rec = screen.records(c)
    ref_a, ref_b = (1 << 5) - 1, (1 << 4) - 1
    for k in range(0, len(c), 37):
        tup = [tuple(int(v) for v in x[k] if v >= 0) for x in (c.ah, c.ap, c.bh, c.bp)]
        d = Det(int(c.parent[k]), ref_a, ref_b, *tup)
        r = to_records([d])[0]
        for f in ("parent", "pha", "phb", "oa", "ob", "ha", "pa", "hb", "pb"):
            assert np.array_equal(np.asarray(rec[f][k]), np.asarray(r[f])), f


def test_phase_closed_form_matches_bitwise_rule():
    rng = np.random.default_rng(4)
    for _ in range(300):
        nocc, nmo = int(rng.integers(1, 9)), 20
        k = int(rng.integers(1, min(3, nocc + 1)))
        h = np.sort(rng.choice(nocc, k, replace=False))
        p = np.sort(rng.choice(np.arange(nocc, nmo), k, replace=False))
        hm = sum(1 << int(i) for i in h)
        pm = sum(1 << int(i) for i in p)
        want = excitation_phase_bits((1 << nocc) - 1, hm, pm)
        hh = np.full((1, 2), -1)
        hh[0, :k] = h
        assert screen._phase(hh, nocc)[0] == want


def test_candidates_unique_and_grouped():
    c = screen.random_candidates(2, 24, 4, 4, 300, seed=1)
    keys = np.concatenate([c.parent[:, None], c.ah, c.ap, c.bh, c.bp], 1)
    assert len(np.unique(keys, axis=0)) == len(c)
    assert np.all(np.diff(c.parent.astype(int)) >= 0)


def test_pick_best_and_kth_match_stable_sort():
    """The device pick equals numpy's descending stable sort of the untaken
    scores (ties in index order), cut at k and at zero; kth_best_d is the
    k-th largest scored entry (unscored entries are -1), or 0."""
    import torch
    from nwick.screen import kth_best_d, pick_best
    rng = np.random.default_rng(9)
    for n, k in [(1000, 64), (5000, 1), (300, 299), (300, 300), (300, 500), (10, 3)]:
        for kind in ("ties", "unique", "sparse"):
            if kind == "ties":
                sc = rng.integers(-1, 5, size=n).astype(float)
            elif kind == "unique":
                sc = rng.random(n)
            else:
                sc = np.full(n, -1.0)
                m = max(1, n // 10)
                sc[rng.choice(n, m, replace=False)] = rng.integers(0, 3, size=m)
            taken = rng.random(n) < 0.2
            masked = np.where(taken, -1.0, sc)
            want = np.argsort(-masked, kind="stable")[:k]
            want = want[masked[want] > 0]
            got = pick_best(torch.tensor(sc), torch.tensor(taken), k)
            assert np.array_equal(got, want), (n, k, kind)
            valid = np.sort(sc[sc >= 0])
            kth = valid[-k] if len(valid) >= k else 0.0
            assert kth_best_d(torch.tensor(sc), k) == kth, (n, k, kind)
