"""Multi-system batch (nw_batch_*): every kernel grouping and schedule, the
tuner and graph replay, against the single-system nw_hs_matrix path and
the CPU oracle.  The fused kernels call the same element arithmetic as the
single-system kernels, so the results are compared at 1e-13."""
import numpy as np
import pytest

import nwick
import oracle as O
from nwick import synth
from nwick.basis import to_records

pytestmark = pytest.mark.gpu
SPEC = [(4, 4, 40), (6, 6, 30), (8, 3, 24)]   # (n, nref, excitations per reference)


@pytest.fixture(scope="module")
def systems():
    out = []
    for sysm, dets in synth.batch_systems(seed=77, systems=SPEC):
        eng = nwick.Engine.from_system(sysm)
        eng.build(tol=1e-8)
        rec = to_records(dets)
        H, S = eng.hs_matrix(rec, mode=0)
        out.append((sysm, dets, eng, rec, H, S))
        print(f"batch system n_bas={sysm.n_bas} nocc={sysm.nocc_a} nref={sysm.nref} ndet={len(dets)}", flush=True)
    return out


def _batch(systems, mode=0):
    b = nwick.Batch()
    for sysm, dets, eng, rec, H, S in systems:
        b.add(eng, rec, mode=mode)
    return b


def _check(b, systems, tol=1e-13):
    for k, (sysm, dets, eng, rec, H, S) in enumerate(systems):
        Hb, Sb = b.result(k)
        assert np.max(np.abs(Hb - H)) <= tol and np.max(np.abs(Sb - S)) <= tol


def test_single_system_path_matches_oracle(systems):
    sysm, dets, eng, rec, H, S = systems[2]
    sub = dets[:60] + dets[-60:]
    Hs, Ss = eng.hs_matrix(to_records(sub), mode=0)
    Ho, So = O.hs_matrix(O.Slab(sysm, 1e-8), sub, mode=0)
    assert np.max(np.abs(Hs - Ho)) < 1e-10 and np.max(np.abs(Ss - So)) < 1e-10


@pytest.mark.parametrize("grouping,dyn", [(nwick.GROUP_CODE, 1), (nwick.GROUP_CLASS, 1), (nwick.GROUP_CLASS, 0),
                                          (nwick.GROUP_ALL, 1), (nwick.GROUP_ALL, 0)])
def test_every_schedule_matches_single_system(systems, grouping, dyn):
    print(f"schedule grouping={grouping} dynamic={dyn}", flush=True)
    b = _batch(systems)
    shapes = [{"nocc": 2, "code": c, "shape": sh} for c, sh in ((12, 2), (13, 6), (22, 5), (23, 3))]
    b.set_schedule({"grouping": grouping, "dynamic": dyn, "grid_per_cu": 2, "buckets": shapes})
    b.plan(tune=False, graph=False)
    b.run()
    _check(b, systems)


def test_tuned_graph_replays(systems):
    """Tuner, HIP graph capture, and three replays with eager single-system
    work in between; every replay equals the eager result."""
    b = _batch(systems)
    b.plan(tune=True, graph=True)
    sched = b.schedule()
    print(f"tuned: grouping={sched['grouping']} dynamic={sched['dynamic']} grid/CU={sched['grid_per_cu']} "
          f"buckets={len(sched['buckets'])}", flush=True)
    assert len(sched["buckets"]) > 0
    for rep in range(3):
        sysm, dets, eng, rec, H, S = systems[rep % len(systems)]
        eng.hs_matrix(rec, mode=0)
        b.run()
        _check(b, systems)
    ms, npairs = b.last()
    assert npairs == sum(len(s[1]) * (len(s[1]) + 1) // 2 for s in systems) and ms > 0
    # the exported schedule reproduces the plan in a fresh batch
    b2 = _batch(systems)
    b2.set_schedule(sched)
    b2.plan(tune=False, graph=True)
    b2.run()
    _check(b2, systems)


def test_corrected_mode_batch(systems):
    b = _batch(systems, mode=1)
    b.plan(tune=False, graph=True)
    b.run()
    for k, (sysm, dets, eng, rec, H, S) in enumerate(systems):
        H1, S1 = eng.hs_matrix(rec, mode=1)
        Hb, Sb = b.result(k)
        assert np.max(np.abs(Hb - H1)) <= 1e-13 and np.max(np.abs(Sb - S1)) <= 1e-13
