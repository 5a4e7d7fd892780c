"""Wick disk cache in the reference's format: wicks.meta parsed by an
independent reader written from the reference's type definitions
(src/nonorthogonalwicks/storage.rs:163-180 WicksDiskMeta, types.rs
SameSpinOffset / DiffSpinOffset / SameSpinMeta / DiffSpinMeta, bincode 1
fixed-width little-endian encoding), wicks.bin against the device slab, and a
load into a fresh context reproducing H and S exactly."""
import os
import struct

import numpy as np
import pytest

import nwick
from nwick import basis, synth

pytestmark = pytest.mark.gpu
NONE = 2 ** 64 - 1


class BincodeReader:
    def __init__(self, data):
        self.d, self.i = data, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.d, self.i)
        self.i += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]


def read_disk_meta(path):
    r = BincodeReader(open(path, "rb").read())
    out = {"version": r.take("I"), "nref": r.take("Q"), "slab_len": r.take("Q")}
    same_off = lambda: {"x": r.take("2Q"), "y": r.take("2Q"), "xrdm": r.take("2Q"), "yrdm": r.take("2Q"),
                        "fh": r.take("4Q"), "ff": r.take("4Q"), "v": r.take("8Q"), "hcol0": r.take("Q"),
                        "j": r.take("10Q")}
    out["off"] = []
    for _ in range(r.take("Q")):
        aa, bb = same_off(), same_off()
        out["off"].append({"aa": aa, "bb": bb, "vab": r.take("8Q"), "vba": r.take("8Q"), "iiab": r.take("16Q")})
    same_meta = lambda: {"tilde_s_prod": r.take("d"), "phase": r.take("d"), "m": r.take("Q"), "nmo": r.take("Q"),
                         "nocc": r.take("Q"), "f0f": r.take("2d"), "f0h": r.take("2d"), "v0": r.take("3d")}
    out["meta"] = []
    for _ in range(r.take("Q")):
        aa, bb = same_meta(), same_meta()
        out["meta"].append({"aa": aa, "bb": bb, "nmo_ab": r.take("Q"), "vab0": r.take("4d"), "vba0": r.take("4d")})
    assert r.i == len(r.d)
    return out


def test_cache_round_trip(tmp_path):
    sysm = synth.make_system(10, 3, 2, 3, seed=44, orthonormal_ao=False, n_aux=20, orth_refs=(2,))
    dets = basis.excited_basis(3, 10, 3, 2, per_ref=25, seed=44, include_refs=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.build(tol=1e-8)
    H, S = eng.hs_matrix(rec, mode=1)
    eng.save_wicks(str(tmp_path))
    meta = read_disk_meta(os.path.join(tmp_path, "wicks.meta"))
    slab = np.fromfile(os.path.join(tmp_path, "wicks.bin"), dtype="<f8")
    assert meta["version"] == 2 and meta["nref"] == 3 and meta["slab_len"] == len(slab)
    assert len(meta["off"]) == len(meta["meta"]) == 9
    base = 0
    m_seen = set()
    for p in range(9):
        x, w = divmod(p, 3)
        blk, off, mt = eng.pair_slab(x, w)
        o = meta["off"][p]
        assert o["aa"]["x"][0] == base                      # pairs back to back in x * nref + w order
        assert o["aa"]["xrdm"] == (0, 0) and o["aa"]["yrdm"] == (0, 0)
        rel = lambda a: -1 if a == NONE else a - base
        # our 95-entry relative offset table: aa block, bb block, vab, vba, iiab
        for sp, key, o0 in ((0, "aa", 0), (1, "bb", 31)):
            s = o[key]
            got = [*map(rel, s["x"]), *map(rel, s["y"]), *map(rel, s["fh"]), *map(rel, s["ff"]), *map(rel, s["v"]),
                   *map(rel, s["j"]), rel(s["hcol0"])]
            assert got == list(off[o0:o0 + 31])
        assert [rel(a) for a in o["vab"]] == list(off[62:70]) and [rel(a) for a in o["vba"]] == list(off[70:78])
        assert [rel(a) for a in o["iiab"]] == list(off[78:94])
        assert np.array_equal(slab[base:base + off[94]], blk)
        pm = meta["meta"][p]
        for sp, key, m0 in ((0, "aa", 0), (1, "bb", 11)):
            s = pm[key]
            assert (s["tilde_s_prod"], s["phase"], s["m"], s["nocc"]) == (mt[m0], mt[m0 + 1], int(mt[m0 + 2]),
                                                                            int(mt[m0 + 3]))
            assert s["nmo"] == 10 and s["f0h"] == tuple(mt[m0 + 4:m0 + 6]) and s["f0f"] == tuple(mt[m0 + 6:m0 + 8])
            assert s["v0"] == tuple(mt[m0 + 8:m0 + 11])
            m_seen.add(s["m"])
        assert pm["vab0"] == tuple(mt[22:26]) and pm["vba0"] == tuple(mt[26:30])
        base += off[94]
    assert base == len(slab) and 1 in m_seen   # m > 0 pairs are in the file
    # a fresh context loads the cache and reproduces the elements bit for bit
    eng2 = nwick.Engine.from_system(sysm)
    eng2.load_wicks(str(tmp_path), tol=1e-8)
    H2, S2 = eng2.hs_matrix(rec, mode=1)
    assert np.array_equal(H2, H) and np.array_equal(S2, S)
    with pytest.raises(nwick.NwickError):
        eng2.update_fock(sysm.h, sysm.h)


def test_cache_of_mixed_build_round_trip(tmp_path):
    """A reduced-precision build keeps its rank-4 blocks as FP32 on the device;
    nw_save_wicks writes them widened to FP64 (the reference's f64 file), so a
    fresh context that loads the cache reproduces the mixed build's elements
    bit for bit (m > 0 pairs included: several J slots and II blocks)."""
    sysm = synth.make_system(10, 3, 2, 3, seed=45, orthonormal_ao=False, n_aux=20, orth_refs=(2,))
    dets = basis.excited_basis(3, 10, 3, 2, per_ref=25, seed=45, include_refs=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm)
    eng.set_precision(nwick.PREC_BF16X3)
    eng.set_mixed_floor(0.0)
    meta = eng.build(tol=1e-8)
    assert all(m.fp32 == 1 for m in meta.values()) and any(m.ma > 0 for m in meta.values())
    H, S = eng.hs_matrix(rec, mode=1)
    eng.save_wicks(str(tmp_path))
    slab = np.fromfile(os.path.join(tmp_path, "wicks.bin"), dtype="<f8")
    base = 0
    for p in range(9):
        blk, off, _ = eng.pair_slab(*divmod(p, 3))
        assert np.array_equal(slab[base:base + off[94]], blk)
        base += off[94]
    eng2 = nwick.Engine.from_system(sysm)
    eng2.load_wicks(str(tmp_path), tol=1e-8)
    H2, S2 = eng2.hs_matrix(rec, mode=1)
    assert np.array_equal(H2, H) and np.array_equal(S2, S)
