"""nw_hs_batched_noalloc (no device allocation; temporaries in a fixed
device array per GPU) against nw_hs_batched on the same device arrays: two
calls in a row, the second with fewer determinants and requests, outputs
bit-identical; sizes beyond its limits and host arrays are refused."""
import numpy as np
import pytest

import nwick
from nwick import basis, synth

pytestmark = pytest.mark.gpu


def _dev(sysm, dets, req):
    import torch
    rec = basis.to_records(dets)
    d = torch.from_numpy(rec.view(np.uint8).copy()).to("cuda:0")
    r = torch.from_numpy(np.ascontiguousarray(req, dtype=np.int64)).to("cuda:0")
    return d, r


def test_noalloc_matches_batched_twice():
    import torch
    sysm = synth.make_system(12, 3, 2, 3, seed=77, orth_refs=(2,), orth_count=1, n_aux=24, enuc=0.4)
    dets = basis.excited_basis(3, 12, 3, 2, per_ref=40, seed=77, include_refs=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    rng = np.random.default_rng(5)
    sizes = [(len(dets), 5000), (len(dets) // 2, 700)]
    for nd, nq in sizes:
        sub = dets[:nd]
        req = rng.integers(0, nd, size=(nq, 2))
        d, r = _dev(sysm, sub, req)
        for mode in (0, 1):
            h0 = torch.full((nq,), float("nan"), dtype=torch.float64, device="cuda:0")
            s0 = torch.full_like(h0, float("nan"))
            h1 = torch.full_like(h0, float("nan"))
            s1 = torch.full_like(h0, float("nan"))
            eng.hs_batched_device(d, r, h0, s0, mode=mode)
            eng.hs_batched_device(d, r, h1, s1, mode=mode, noalloc=True)
            torch.cuda.synchronize()
            a0, b0, a1, b1 = (t.cpu().numpy() for t in (h0, s0, h1, s1))
            print(f"noalloc: ndet={nd} nreq={nq} mode={mode} max|H|={np.max(np.abs(a0)):.3f}", flush=True)
            assert np.all(np.isfinite(a1)) and np.all(np.isfinite(b1))
            assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
    eng.close()


def test_noalloc_refuses_oversize_and_host_arrays():
    import torch
    sysm = synth.make_system(10, 2, 2, 2, seed=78, n_aux=20)
    dets = basis.excited_basis(2, 10, 2, 2, per_ref=10, seed=78, include_refs=True)
    eng = nwick.Engine.from_system(sysm)
    eng.build()
    d, r = _dev(sysm, dets, np.zeros((4, 2), dtype=np.int64))
    h = torch.empty(4, dtype=torch.float64, device="cuda:0")
    s = torch.empty_like(h)
    big = torch.zeros(((1 << 20) + 1, 2), dtype=torch.int64, device="cuda:0")
    hb = torch.empty((1 << 20) + 1, dtype=torch.float64, device="cuda:0")
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_batched_device(d, big, hb, hb.clone(), noalloc=True)
    assert e.value.code == 2  # NW_E_SHAPE
    many = torch.zeros(((1 << 18) + 1) * nwick.DET_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(nwick.NwickError) as e:
        eng.hs_batched_device(many, r, h, s, noalloc=True)
    assert e.value.code == 2
    rec = basis.to_records(dets)
    req = np.zeros((4, 2), dtype=np.uint64)
    hh, ss = np.empty(4), np.empty(4)
    rc = eng.L.nw_hs_batched_noalloc(eng.ctx, nwick._ptr(rec), len(rec), nwick._ptr(req), 4, nwick._ptr(hh),
                                     nwick._ptr(ss), 0)
    assert rc == 1  # NW_E_ARG: host arrays
    eng.close()
