"""ctypes front end of the CPU oracle (oracle/nwick_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libnwick_oracle.so")
_lib = None

D = ct.POINTER(ct.c_double)
U64 = ct.POINTER(ct.c_uint64)
I64 = ct.POINTER(ct.c_longlong)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ct.CDLL(_LIB)
        L.ow_svd.argtypes = [ct.c_int, D, D, D, D]
        L.ow_pair_offsets.argtypes = [ct.c_int, ct.c_int, ct.c_int, I64]
        L.ow_pair_offsets.restype = ct.c_longlong
        L.ow_count_zero.argtypes = [ct.c_int, ct.c_int, D, D, U64, D, U64, ct.c_double, D]
        L.ow_build_pair.argtypes = [ct.c_int, ct.c_int, D, D, D, D, D, U64, U64, D, D, U64, U64,
                                    ct.c_double, I64, D, D]
        L.ow_hs.argtypes = [D, I64, D, ct.c_int, U64, U64, ct.c_double, ct.c_double, ct.c_double, ct.c_int, D, D]
        L.ow_hs_readings.argtypes = L.ow_hs.argtypes
        L.ow_excitation_phase.argtypes = [U64, U64, U64]
        L.ow_excitation_phase.restype = ct.c_double
        L.ow_hs_naive.argtypes = [ct.c_int, ct.c_int, D, D, D, D, D, D, U64, U64, D, D, U64, U64,
                                  ct.c_double, ct.c_double, D, D]
        L.ow_update_fock_pair.argtypes = [ct.c_int, ct.c_int, D, D, D, D, D, U64, U64, D, D, U64, U64,
                                          ct.c_double, I64, D, D]
        L.ow_fs.argtypes = [D, I64, D, ct.c_int, U64, U64, ct.c_double, ct.c_double, ct.c_double, ct.c_int, D, D]
        L.ow_fs_spin.argtypes = [D, I64, D, ct.c_int, ct.c_int, U64, U64, U64, U64, ct.c_double, ct.c_double,
                                 ct.c_int, D, D]
        L.ow_set_dgemm.argtypes = [ct.c_void_p]
        L.ow_meta_len.restype = ct.c_int
        L.ow_offsets_len.restype = ct.c_int
        _lib = L
    return _lib


def use_openblas(enable=True):
    """Quarter passes of the oracle's four-index transforms through OpenBLAS
    dgemm (scipy's bundled build, reached through scipy.linalg.cython_blas),
    the GEMM library the reference links; enable=False restores the built-in
    loops.  Returns a description of the library and its thread count."""
    if not enable:
        lib().ow_set_dgemm(None)
        return None
    from scipy.linalg import cython_blas
    cap = cython_blas.__pyx_capi__["dgemm"]
    get_name = ct.pythonapi.PyCapsule_GetName
    get_name.restype, get_name.argtypes = ct.c_char_p, [ct.py_object]
    get_ptr = ct.pythonapi.PyCapsule_GetPointer
    get_ptr.restype, get_ptr.argtypes = ct.c_void_p, [ct.py_object, ct.c_char_p]
    lib().ow_set_dgemm(get_ptr(cap, get_name(cap)))
    info = "OpenBLAS (scipy)"
    try:
        from threadpoolctl import threadpool_info
        for t in threadpool_info():
            if t.get("internal_api") == "openblas" and "scipy" in t.get("filepath", ""):
                info = f"OpenBLAS {t.get('version')} (scipy), {t.get('num_threads')} threads"
    except Exception:
        pass
    return info


# Literal readings of the reference's m>0 arithmetic (oracle/nwick_oracle.c RD_*).
RD_V0, RD_VIDX, RD_JSLOT, RD_DIST, RD_MINOR, RD_ENUC = 1, 2, 4, 8, 16, 32
RD_ALL = 63
# planted index mistakes, for showing that the checks of RD_VIDX / RD_JSLOT see one
RD_PLANT_VT, RD_PLANT_JCOL = 256, 512


def _d(a):
    return a.ctypes.data_as(D)


def _u(a):
    return a.ctypes.data_as(U64)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def mask2(v: int):
    return np.array([v & 0xFFFFFFFFFFFFFFFF, (v >> 64) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)


def svd(A):
    A = _c(A)
    n = A.shape[0]
    U = np.zeros((n, n)); s = np.zeros(n); V = np.zeros((n, n))
    lib().ow_svd(n, _d(A), _d(U), _d(s), _d(V))
    return U, s, V


def excitation_phase(occ: int, holes: int, parts: int) -> float:
    return lib().ow_excitation_phase(_u(mask2(occ)), _u(mask2(holes)), _u(mask2(parts)))


def pair_offsets(ma, mb, nmo):
    off = np.zeros(lib().ow_offsets_len(), dtype=np.int64)
    size = lib().ow_pair_offsets(ma, mb, nmo, off.ctypes.data_as(I64))
    return off, int(size)


class Slab:
    """All requested ordered reference pairs, laid out as src/nonorthogonalwicks/layout.rs."""

    def __init__(self, sysm, tol=1e-8, pairs=None):
        L = lib()
        self.nref = sysm.nref
        self.nmo = sysm.nmo
        self.tol = tol
        self.enuc = sysm.enuc
        nb, nm = sysm.n_bas, sysm.nmo
        S, h, eri = _c(sysm.S), _c(sysm.h), _c(sysm.eri)
        oa, ob = mask2((1 << sysm.nocc_a) - 1), mask2((1 << sysm.nocc_b) - 1)
        if pairs is None:
            pairs = [(x, w) for x in range(self.nref) for w in range(self.nref)]
        self.pairs = list(pairs)
        self.off = {}
        self.base = {}
        self.meta = {}
        self.m = {}
        total = 0
        for (x, w) in self.pairs:
            Cx, Cw = sysm.refs[x], sysm.refs[w]
            ma = L.ow_count_zero(nb, nm, _d(S), _d(_c(Cx[0])), _u(oa), _d(_c(Cw[0])), _u(oa), tol, None)
            mb = L.ow_count_zero(nb, nm, _d(S), _d(_c(Cx[1])), _u(ob), _d(_c(Cw[1])), _u(ob), tol, None)
            off, size = pair_offsets(ma, mb, nm)
            self.off[(x, w)] = off
            self.base[(x, w)] = total
            self.m[(x, w)] = (ma, mb)
            total += size
        self.slab = np.full(total, np.nan)
        for (x, w) in self.pairs:
            Cx, Cw = sysm.refs[x], sysm.refs[w]
            meta = np.zeros(L.ow_meta_len())
            sub = self.slab[self.base[(x, w)]:]
            L.ow_build_pair(nb, nm, _d(S), _d(h), _d(eri),
                            _d(_c(Cx[0])), _d(_c(Cx[1])), _u(oa), _u(ob),
                            _d(_c(Cw[0])), _d(_c(Cw[1])), _u(oa), _u(ob),
                            tol, self.off[(x, w)].ctypes.data_as(I64), _d(sub), _d(meta))
            self.meta[(x, w)] = meta

    def update_fock(self, sysm, fa, fb):
        """update_wicks_fock (src/noci/wicks.rs:290): f0f and ff for every built pair."""
        nb, nm = sysm.n_bas, sysm.nmo
        oa, ob = mask2((1 << sysm.nocc_a) - 1), mask2((1 << sysm.nocc_b) - 1)
        for (x, w) in self.pairs:
            Cx, Cw = sysm.refs[x], sysm.refs[w]
            sub = self.slab[self.base[(x, w)]:]
            lib().ow_update_fock_pair(nb, nm, _d(_c(sysm.S)), _d(_c(fa)), _d(_c(fb)),
                                      _d(_c(Cx[0])), _d(_c(Cx[1])), _u(oa), _u(ob),
                                      _d(_c(Cw[0])), _d(_c(Cw[1])), _u(oa), _u(ob),
                                      self.tol, self.off[(x, w)].ctypes.data_as(I64), _d(sub), _d(self.meta[(x, w)]))

    def fs(self, bra, ket, mode=0):
        """(F, S) of one determinant pair; M = F - e0 S."""
        sub, off, meta = self.pair_view(bra.parent, ket.parent)
        xex = np.concatenate([mask2(bra.ha), mask2(bra.pa), mask2(bra.hb), mask2(bra.pb)])
        wex = np.concatenate([mask2(ket.ha), mask2(ket.pa), mask2(ket.hb), mask2(ket.pb)])
        F = ct.c_double(); S = ct.c_double()
        lib().ow_fs(_d(sub), off.ctypes.data_as(I64), _d(meta), self.nmo, _u(xex), _u(wex),
                    bra.pha * ket.pha, bra.phb * ket.phb, self.tol, mode, ct.byref(F), ct.byref(S))
        return F.value, S.value

    def fs_spin(self, bra, ket, spin, mode=0):
        """One spin's {S, F} factor of a pair (bra = x side), times that spin's
        excitation phases: the per-spin value the factor tables hold."""
        sub, off, meta = self.pair_view(bra.parent, ket.parent)
        if spin == 0:
            m = [mask2(bra.ha), mask2(bra.pa), mask2(ket.ha), mask2(ket.pa)]
            ph = bra.pha * ket.pha
        else:
            m = [mask2(bra.hb), mask2(bra.pb), mask2(ket.hb), mask2(ket.pb)]
            ph = bra.phb * ket.phb
        F = ct.c_double(); S = ct.c_double()
        lib().ow_fs_spin(_d(sub), off.ctypes.data_as(I64), _d(meta), self.nmo, int(spin), *[_u(v) for v in m],
                         float(ph), self.tol, mode, ct.byref(F), ct.byref(S))
        return S.value, F.value

    def pair_view(self, x, w):
        b = self.base[(x, w)]
        return self.slab[b:], self.off[(x, w)], self.meta[(x, w)]

    def _hs_call(self, fn, bra, ket, arg):
        sub, off, meta = self.pair_view(bra.parent, ket.parent)
        xex = np.concatenate([mask2(bra.ha), mask2(bra.pa), mask2(bra.hb), mask2(bra.pb)])
        wex = np.concatenate([mask2(ket.ha), mask2(ket.pa), mask2(ket.hb), mask2(ket.pb)])
        H = ct.c_double(); S = ct.c_double()
        exph = (bra.pha * ket.pha) * (bra.phb * ket.phb)
        fn(_d(sub), off.ctypes.data_as(I64), _d(meta), self.nmo, _u(xex), _u(wex),
           exph, self.enuc, self.tol, int(arg), ct.byref(H), ct.byref(S))
        return H.value, S.value

    def hs(self, bra, ket, mode=0):
        """xw_hamiltonian_overlap_prepared for one determinant pair (bra = x side).
        mode 0: the reference arithmetic; mode 1: corrected m>0 path."""
        return self._hs_call(lib().ow_hs, bra, ket, mode)

    def hs_readings(self, bra, ket, rd):
        """One determinant pair with a chosen subset `rd` of the reference's
        literal m>0 readings (RD_* flags; RD_ALL = mode 0, 0 = mode 1)."""
        return self._hs_call(lib().ow_hs_readings, bra, ket, rd)


def det_coeffs(sysm, det):
    return sysm.refs[det.parent]


def hs_naive(sysm, bra, ket, tol=1e-8, eri_asym=None):
    """calculate_hs_pair_naive (src/noci/hs.rs:365) on the excited determinants."""
    nb, nm = sysm.n_bas, sysm.nmo
    if eri_asym is None:
        eri_asym = sysm.eri_asym()
    Cl, Cg = sysm.refs[bra.parent], sysm.refs[ket.parent]
    H = ct.c_double(); S = ct.c_double()
    lib().ow_hs_naive(nb, nm, _d(_c(sysm.S)), _d(_c(sysm.h)), _d(_c(sysm.eri)), _d(_c(eri_asym)),
                      _d(_c(Cl[0])), _d(_c(Cl[1])), _u(mask2(bra.oa)), _u(mask2(bra.ob)),
                      _d(_c(Cg[0])), _d(_c(Cg[1])), _u(mask2(ket.oa)), _u(mask2(ket.ob)),
                      sysm.enuc, tol, ct.byref(H), ct.byref(S))
    return H.value, S.value


def fs_matrix(slab: Slab, dets, mode=0):
    n = len(dets)
    F = np.zeros((n, n)); S = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            f, s = slab.fs(dets[i], dets[j], mode)
            F[i, j] = F[j, i] = f
            S[i, j] = S[j, i] = s
    return F, S


def fs_naive(sysm, bra, ket, fa, fb, tol=1e-8):
    """One-body generalised Slater-Condon route for F (one_electron_scalar,
    src/noci/naive.rs:270 via calculate_m_pair_naive, src/noci/m.rs:114):
    F = S_b f_a + S_a f_b on the excited determinants."""
    import numpy as _np
    res = []
    for sp, (Fm, occ_l, occ_g) in enumerate(((fa, bra.oa, ket.oa), (fb, bra.ob, ket.ob))):
        Cl = sysm.refs[bra.parent][sp][:, [i for i in range(sysm.nmo) if occ_l >> i & 1]]
        Cg = sysm.refs[ket.parent][sp][:, [i for i in range(sysm.nmo) if occ_g >> i & 1]]
        Socc = Cl.T @ sysm.S @ Cg
        s = _np.linalg.det(Socc)
        # one-body transition element <l|f|g> = sum_ij (Cl^T F Cg)_ij adj(Socc)_ji
        A = Cl.T @ Fm @ Cg
        n = Socc.shape[0]
        cof = _np.zeros_like(Socc)
        for i in range(n):
            for j in range(n):
                mnr = _np.delete(_np.delete(Socc, i, 0), j, 1)
                cof[i, j] = (-1) ** (i + j) * (_np.linalg.det(mnr) if n > 1 else 1.0)
        res.append((s, float(_np.sum(A * cof))))
    (sa, fa_), (sb, fb_) = res
    return sb * fa_ + sa * fb_, sa * sb


def hs_matrix(slab: Slab, dets, symmetric=True, mode=0):
    """calculate_matrix_elements (src/noci/matrix.rs:39-81) + scatter (:93):
    pairs i <= j, mirrored."""
    n = len(dets)
    H = np.zeros((n, n)); S = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            h, s = slab.hs(dets[i], dets[j], mode)
            H[i, j] = h; S[i, j] = s
            H[j, i] = h; S[j, i] = s
    return H, S
