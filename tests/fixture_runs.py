"""Drivers that run the reference's own test fixtures (tests/fixtures/*,
inputs and expected energies copied into tests/golden/reference_fixtures.json)
through the front end (integrals, SCF states) and either the CPU oracle or
the GPU engine."""
import json
import os

import numpy as np

from nwick import snoci as SN
from nwick.basis import Det, excited_basis
from nwick.molecule import Molecule
from nwick.pt2 import general_evp
from nwick.scf import ao_data, reference_states, to_system

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))
TOL_E = 1e-8   # the reference's own fixture tolerance (tests/common/mod.rs assert_close)
SCF_TOL = 1e-12  # post.tol used by the reference's fixture runs (tests/pt2.rs:104-125)


def system(name):
    fx = FIXTURES[name]
    mol = Molecule([(a, tuple(x)) for a, x in fx["atoms"]], fx["basis"])
    ao = ao_data(mol)
    st = reference_states(ao, fx["states"])
    return fx, ao, st, to_system(ao, st)


def ref_dets(sysm):
    oa, ob = (1 << sysm.nocc_a) - 1, (1 << sysm.nocc_b) - 1
    return [Det(i, oa, ob) for i in range(sysm.nref)]


def noci_energy(H, S):
    e, _ = general_evp(H, S, SCF_TOL)
    return float(e[0])


def candidates(fx, sysm):
    return excited_basis(sysm.nref, sysm.nmo, sysm.nocc_a, sysm.nocc_b, per_ref=None, orders=tuple(fx["orders"]))


def snoci_options(fx):
    g = fx["gmres"]
    s = fx["snoci"]
    return SN.SNOCIOptions(orders=tuple(fx["orders"]), max_iter=s["max_iter"], max_add=s["max_add"],
                           res_tol=g["res_tol"], restart=g["restart"], gmres_max_iter=g["max_iter"])


def snoci_refs(st, sysm):
    return SN.reference_states([s.label for s in st if s.noci], sysm.nocc_a, sysm.nocc_b)


def close(a, b, tol=TOL_E):
    return abs(a - b) <= tol * max(1.0, abs(b)) if np.isfinite(a) else False
