"""The reference's golden fixture energies through the GPU engine.

Front end (integrals, SCF states) on the host, then every Wick H, S, F element,
the PT2 matvec and the GMRES solve on the GPU through the C ABI; energies are
compared with the reference's tests/fixtures/*/expected.json at its 1e-8 Ha
tolerance, and the H, S matrices element-wise with the CPU oracle at 1e-10.
The H2 cc-pVDZ case is the reference's README NOCI example (RHF + two
spin-biased UHF states, r = 1.5 Angstrom), checked against the oracle."""
import numpy as np
import pytest

import nwick
import oracle as O
from fixture_runs import SCF_TOL, candidates, close, noci_energy, ref_dets, snoci_options, snoci_refs, system
from nwick import snoci as SN
from nwick.basis import to_records
from nwick.molecule import Molecule
from nwick.pt2 import PT2
from nwick.scf import ao_data, reference_states, to_system

pytestmark = pytest.mark.gpu


def _engine(sysm):
    eng = nwick.Engine.from_system(sysm)
    eng.build(tol=SCF_TOL)
    return eng


@pytest.mark.parametrize("name", ["REF_NOCI_H2_STO-3G_1_5", "PT2_H2_3-21G_1_5_WICKS", "SNOCI_H4_STO-3G_1_5_WICKS",
                                  "REF_NOCI_H4_STO-3G_1_75", "REF_NOCI_F2_STO-3G_1_75_WICKS",
                                  "REF_NOCI_LiH_6-31G_2_8_WICKS"])
def test_gpu_reference_noci_energy(name):
    fx, ao, st, sysm = system(name)
    print(f"fixture {name}: n_bas={sysm.n_bas} nref={sysm.nref}", flush=True)
    eng = _engine(sysm)
    dets = ref_dets(sysm)
    H, S = eng.hs_matrix(to_records(dets), mode=0)
    Ho, So = O.hs_matrix(O.Slab(sysm, SCF_TOL), dets, mode=0)
    assert np.max(np.abs(H - Ho)) < 1e-10 and np.max(np.abs(S - So)) < 1e-10
    e = noci_energy(H, S)
    print(f"{name}: GPU E_NOCI {e:.12f} (reference {fx['reference_noci_energy']:.12f})", flush=True)
    assert close(e, fx["reference_noci_energy"])


@pytest.mark.parametrize("name", ["PT2_H2_STO-3G_1_5_WICKS", "PT2_H2_3-21G_1_5_WICKS"])
def test_gpu_noci_pt2_energy(name):
    fx, ao, st, sysm = system(name)
    print(f"fixture {name}", flush=True)
    eng = _engine(sysm)
    g = fx["gmres"]
    pt = PT2(eng, sysm, ref_dets(sysm), candidates(fx, sysm), tol=SCF_TOL, mode=0)
    pt.setup()
    r = pt.solve(res_tol=g["res_tol"], restart=g["restart"], max_iter=g["max_iter"], preconditioner="woodbury")
    print(f"{name}: GPU E_NOCI+PT2 {r['e_total']:.12f} (reference {fx['noci_pt2_energy']:.12f}) "
          f"iters={r['iterations']}", flush=True)
    assert r["converged"]
    assert close(r["e_total"], fx["noci_pt2_energy"])


@pytest.mark.parametrize("name", ["SNOCI_H2_3-21G_1_5_WICKS", "SNOCI_H4_STO-3G_1_5_WICKS"])
def test_gpu_snoci_energy(name):
    fx, ao, st, sysm = system(name)
    print(f"fixture {name}", flush=True)
    eng = _engine(sysm)
    res = SN.snoci(SN.GpuPT2Backend(eng, sysm, SCF_TOL), snoci_refs(st, sysm), sysm.nmo, snoci_options(fx))
    print(f"{name}: GPU E_SNOCI {res.ecurrent:.12f} (reference {fx['snoci_energy']:.12f})", flush=True)
    assert close(res.ecurrent, fx["snoci_energy"])


def test_gpu_h2_ccpvdz_noci_matches_oracle():
    """H2 / cc-pVDZ, r = 1.5 Angstrom: RHF + UHF {1,-1} + UHF {-1,1} at pol 0.75
    (10 basis functions, 1 alpha + 1 beta electron): 3x3 H and S element-wise
    against the oracle, and the NOCI energy."""
    mol = Molecule([("H", (0, 0, -0.75)), ("H", (0, 0, 0.75))], "cc-pvdz")
    ao = ao_data(mol)
    st = reference_states(ao, [{"label": "RHF (0, 0)"},
                               {"label": "UHF (1, -1)", "spin_bias": {"pattern": [1, -1], "pol": 0.75}},
                               {"label": "UHF (-1, 1)", "spin_bias": {"pattern": [-1, 1], "pol": 0.75}}])
    sysm = to_system(ao, st)
    eng = _engine(sysm)
    dets = ref_dets(sysm)
    H, S = eng.hs_matrix(to_records(dets), mode=0)
    Ho, So = O.hs_matrix(O.Slab(sysm, SCF_TOL), dets, mode=0)
    print(f"h2 cc-pvdz: E_NOCI gpu {noci_energy(H, S):.12f} oracle {noci_energy(Ho, So):.12f}", flush=True)
    assert np.max(np.abs(H - Ho)) < 1e-10 and np.max(np.abs(S - So)) < 1e-10
    assert abs(noci_energy(H, S) - noci_energy(Ho, So)) < 1e-10


@pytest.mark.parametrize("name", ["DET_H2_STO-3G_1_5_WICKS", "DET_F2_sto-3g_1_75_WICKS"])
def test_gpu_deterministic_energy(name):
    """The deterministic NOCI-QMC fixtures through the GPU: H, S over the
    references and their singles and doubles (300 determinants for F2 /
    STO-3G, every element class with contraction ranks up to 4 per spin on
    9+9-electron references), element-wise against the oracle, and the
    lowest root of the pencil against the reference's energy."""
    fx, ao, st, sysm = system(name)
    dets = candidates(fx, sysm)
    dets = [d for r in range(sysm.nref) for d in ([ref_dets(sysm)[r]] + [x for x in dets if x.parent == r])]
    print(f"fixture {name}: {len(dets)} determinants", flush=True)
    eng = _engine(sysm)
    H, S = eng.hs_matrix(to_records(dets), mode=0)
    Ho, So = O.hs_matrix(O.Slab(sysm, SCF_TOL), dets, mode=0)
    assert np.max(np.abs(H - Ho)) < 1e-10 * max(1.0, np.max(np.abs(Ho))) and np.max(np.abs(S - So)) < 1e-10
    e = noci_energy(H, S)
    print(f"{name}: GPU E_det {e:.12f} (reference {fx['deterministic_noci_energy']:.12f})", flush=True)
    assert abs(e - fx["deterministic_noci_energy"]) < fx["tolerance"]
