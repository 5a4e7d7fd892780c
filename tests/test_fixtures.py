"""The reference's own golden energies, reproduced on the CPU.

Front end (integrals, SCF states) -> oracle Wick H, S, F -> NOCI, NOCI-PT2 and
SNOCI energies, compared with tests/fixtures/*/expected.json of the reference
at its own tolerance of 1e-8 Ha.  These pin the oracle (Wick build, H/S/F
evaluation, m classification, noci_density, Fock, PT2 projection and energy,
SNOCI selection) end to end against numbers the reference produced."""
import numpy as np
import pytest

import oracle as O
from fixture_runs import (FIXTURES, SCF_TOL, TOL_E, candidates, close, noci_energy, ref_dets, snoci_options,
                          snoci_refs, system)
from nwick import snoci as SN
from nwick.molecule import Molecule, integrals
from pt2_dense import DensePT2Backend, dense_pt2

NOCI_CASES = ["REF_NOCI_H2_STO-3G_1_5", "PT2_H2_3-21G_1_5_WICKS", "SNOCI_H4_STO-3G_1_5_WICKS",
              "REF_NOCI_H4_STO-3G_1_75", "REF_NOCI_F2_STO-3G_1_75_WICKS", "REF_NOCI_LiH_6-31G_2_8_WICKS"]


def test_h2_sto3g_integrals_textbook():
    """H2 / STO-3G at R = 1.4 bohr: the integrals tabulated by Szabo and Ostlund (Sec. 3.5.2)."""
    mol = Molecule([("H", (0, 0, 0)), ("H", (0, 0, 1.4))], "sto-3g", unit="bohr")
    S, T, V, eri = integrals(mol)
    assert abs(S[0, 1] - 0.6593) < 1e-4
    assert abs(T[0, 0] - 0.7600) < 1e-4 and abs(T[0, 1] - 0.2365) < 1e-4
    assert abs(V[0, 0] - (-1.8804)) < 1e-4 and abs(V[0, 1] - (-1.1948)) < 1e-4
    for idx, val in (((0, 0, 0, 0), 0.7746), ((0, 0, 1, 1), 0.5697), ((0, 1, 0, 1), 0.2970), ((0, 0, 0, 1), 0.4441)):
        assert abs(eri[idx] - val) < 1e-4


def test_p_integrals_are_centre_derivatives_of_s_integrals():
    """A normalised p primitive equals (1/sqrt(a)) d/dA of the normalised s
    primitive of the same exponent on the same centre A; central differences
    of s-only integrals therefore fix every p integral (overlap, kinetic,
    nuclear attraction, ERI) independently of the Hermite recurrences.  The
    moving function sits on a charge-free centre so no nucleus moves."""
    from nwick import molecule as M
    M.CHARGES["X"] = 0
    M.BASIS_SETS["_t_s"] = {"X": [(0, [(0.8, 1.0)])], "He": [(0, [(0.5, 1.0)])], "Li": [(0, [(1.3, 1.0)])]}
    M.BASIS_SETS["_t_p"] = {"X": [(1, [(0.8, 1.0)])], "He": [(0, [(0.5, 1.0)])], "Li": [(0, [(1.3, 1.0)])]}
    try:
        base = [("X", np.array([0.1, -0.2, 0.3])), ("He", np.array([0.0, 0.4, 1.5])),
                ("Li", np.array([0.9, 0.2, -0.4]))]
        Sp, Tp, Vp, Ep = integrals(Molecule([(a, tuple(x)) for a, x in base], "_t_p", unit="bohr"))
        a, h = 0.8, 1e-4

        def s_ints(shift, d):
            at = [(s, x.copy()) for s, x in base]
            at[0][1][d] += shift
            return integrals(Molecule([(s, tuple(x)) for s, x in at], "_t_s", unit="bohr"))

        for d in range(3):
            plus, minus = s_ints(h, d), s_ints(-h, d)
            Sd, Td, Vd, Ed = [(p_ - m_) / (2 * h) / np.sqrt(a) for p_, m_ in zip(plus, minus)]
            # p molecule rows: X p_x, p_y, p_z = 0..2, He s = 3, Li s = 4; s molecule: X s = 0, He = 1, Li = 2
            assert abs(Sp[d, 3] - Sd[0, 1]) < 1e-7
            assert abs(Tp[d, 4] - Td[0, 2]) < 1e-7
            assert abs(Vp[d, 3] - Vd[0, 1]) < 1e-7
            assert abs(Ep[d, 3, 4, 4] - Ed[0, 1, 2, 2]) < 1e-7
            assert abs(Ep[d, 4, 3, 3] - Ed[0, 2, 1, 1]) < 1e-7
    finally:
        del M.BASIS_SETS["_t_s"], M.BASIS_SETS["_t_p"], M.CHARGES["X"]


@pytest.mark.parametrize("name", NOCI_CASES)
def test_scf_and_reference_noci_energy(name):
    fx, ao, st, sysm = system(name)
    got = sorted(s.e for s in st)
    want = sorted(fx["scf_energies"])
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert close(g, w), (g, w)
    slab = O.Slab(sysm, SCF_TOL)
    H, S = O.hs_matrix(slab, ref_dets(sysm), mode=0)
    e = noci_energy(H, S)
    print(f"{name}: E_NOCI {e:.12f} (reference {fx['reference_noci_energy']:.12f})")
    assert close(e, fx["reference_noci_energy"])


@pytest.mark.parametrize("name", ["PT2_H2_STO-3G_1_5_WICKS", "PT2_H2_3-21G_1_5_WICKS"])
def test_noci_pt2_energy(name):
    fx, ao, st, sysm = system(name)
    r = dense_pt2(sysm, ref_dets(sysm), candidates(fx, sysm), mode=0, tol=SCF_TOL)
    print(f"{name}: E_NOCI+PT2 {r['e_total']:.12f} (reference {fx['noci_pt2_energy']:.12f})")
    assert close(r["e_noci"], fx["reference_noci_energy"])
    assert close(r["e_total"], fx["noci_pt2_energy"])


@pytest.mark.parametrize("name", ["SNOCI_H2_STO-3G_1_5_WICKS", "SNOCI_H2_3-21G_1_5_WICKS",
                                  "SNOCI_H4_STO-3G_1_5_WICKS"])
def test_snoci_energy(name):
    fx, ao, st, sysm = system(name)
    res = SN.snoci(DensePT2Backend(sysm, SCF_TOL), snoci_refs(st, sysm), sysm.nmo, snoci_options(fx))
    print(f"{name}: E_SNOCI {res.ecurrent:.12f} (reference {fx['snoci_energy']:.12f}), ncur={len(res.current)}")
    assert close(res.ecurrent, fx["snoci_energy"])


@pytest.mark.parametrize("name", ["DET_H2_STO-3G_1_5_WICKS", "DET_F2_sto-3g_1_75_WICKS"])
def test_deterministic_energy(name):
    """Deterministic NOCI-QMC (tests/deterministic.rs:91-160): propagation in
    the references' singles/doubles space converges to the lowest root of the
    (H, S) pencil over that space, which the general eigensolver gives
    directly.  F2 / STO-3G: 300 determinants (every reference's singles and
    doubles fill the 100-determinant space), at the reference's 1e-7."""
    fx, ao, st, sysm = system(name)
    for g, w in zip(sorted(s.e for s in st), sorted(fx["scf_energies"])):
        assert close(g, w), (g, w)
    dets = candidates(fx, sysm)
    dets = [d for r in range(sysm.nref) for d in ([ref_dets(sysm)[r]] + [x for x in dets if x.parent == r])]
    slab = O.Slab(sysm, SCF_TOL)
    H, S = O.hs_matrix(slab, dets, mode=0)
    e = noci_energy(H, S)
    print(f"{name}: {len(dets)} dets, E_det {e:.12f} (reference {fx['deterministic_noci_energy']:.12f})")
    assert abs(e - fx["deterministic_noci_energy"]) < fx["tolerance"]


def test_h2_321g_scan_last_geometry_scf():
    """REF_NOCI_H2_3-21G_1_0_WICKS at its last geometry (1.00 Angstrom): the
    RHF and both spin-biased UHF states collapse to one energy."""
    fx, ao, st, sysm = system("REF_NOCI_H2_3-21G_1_0_WICKS")
    for g, w in zip(sorted(s.e for s in st), sorted(fx["final_scf_energies"])):
        assert close(g, w), (g, w)
