"""Python host mirror of the MI355X nonorthogonal-Wick engine.

Thin ctypes layer over the C ABI in include/nwick.h (libnwick.so, built from
noci-rs_amd/csrc).  The object model mirrors noci-rs: an AO problem
(AoData), reference determinants, a per-reference-pair build
(build_wicks_shared) and the element evaluators used by the NOCI H/S build
(calculate_matrix_elements) and batched QMC spawning
(calculate_hs_pairs_wicks_batched).

There is no fallback: if the shared library is missing or fails to load,
every entry point raises.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from .basis import DET_DTYPE  # noqa: F401  (re-export)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libnwick.so")

MODE_REFERENCE = 0
MODE_CORRECTED = 1

STATUS = {0: "NW_OK", 1: "NW_E_ARG", 2: "NW_E_SHAPE", 3: "NW_E_HIP", 4: "NW_E_MEM", 5: "NW_E_RANK",
          6: "NW_E_STATE", 7: "NW_E_NONFINITE", 8: "NW_E_ORDER"}

# Every symbol the header declares.
EXPORTS = ["nw_create", "nw_set_refs", "nw_build", "nw_hs_batched", "nw_hs_matrix", "nw_update_fock",
           "nw_fs_matrix", "nw_fs_batched", "nw_matvec_m", "nw_project_hs", "nw_noci_fock", "nw_set_precision", "nw_set_mixed_floor", "nw_save_wicks", "nw_load_wicks", "nw_pair_slab", "nw_last_timings", "nw_hs_matrix_streamed",
           "nw_last_error", "nw_destroy", "nw_batch_create", "nw_batch_add", "nw_batch_plan", "nw_batch_run",
           "nw_batch_result", "nw_batch_schedule", "nw_batch_set_schedule", "nw_batch_last_ms",
           "nw_batch_last_error", "nw_batch_destroy", "nw_mlp_create", "nw_mlp_num_params", "nw_mlp_set_params",
           "nw_mlp_get_params", "nw_mlp_forward", "nw_mlp_grad", "nw_mlp_step", "nw_mlp_train", "nw_mlp_last_error",
           "nw_mlp_destroy", "nw_factor_create", "nw_factor_refresh", "nw_factor_apply", "nw_factor_diagonals",
           "nw_factor_info", "nw_factor_ids", "nw_factor_table", "nw_factor_last_ms", "nw_factor_destroy",
           "nw_factor_set_precision", "nw_factor_apply_pair", "nw_factor_apply_bytes", "nw_factor_block_groups", "nw_hs_batched_noalloc", "nw_s_same_batched", "nw_m_batched",
           "nw_set_eri_factors", "nw_rank4_coverage", "nw_rank4_stats",
           "nw_rank4_kernel_stats", "nw_hs_naive_batched", "nw_hs_compare"]
PREC_FP64 = 0
PREC_MIXED = 1
PREC_BF16X3 = 2
BATCH_TUNE = 1
BATCH_GRAPH = 2
GROUP_CODE, GROUP_CLASS, GROUP_ALL = 0, 1, 2
FACTOR_DEVICE, FACTOR_TRANSIENT = 0, 1

_lib = None


class NwickError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class TuneEntry(ct.Structure):
    _fields_ = [("nocc", ct.c_int32), ("code", ct.c_int32), ("shape", ct.c_int32), ("pad", ct.c_int32),
                ("ms_best", ct.c_double), ("ms_default", ct.c_double)]


class PairMeta(ct.Structure):
    _fields_ = [("ma", ct.c_int32), ("mb", ct.c_int32), ("tilde_s_a", ct.c_double), ("phase_a", ct.c_double),
                ("tilde_s_b", ct.c_double), ("phase_b", ct.c_double), ("sigma_min", ct.c_double),
                ("fp32", ct.c_int32), ("pad", ct.c_int32)]


def load():
    """Load libnwick.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NwickError(3, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    try:
        # One HIP runtime per process: when torch is present its bundled
        # libamdhip64.so.7 must be the one libnwick binds to, so load it first.
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ct.CDLL(LIB_PATH)
    vp, dp, u64p = ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_uint64)
    L.nw_create.argtypes = [ct.POINTER(vp), ct.c_int, ct.c_int, vp, vp, vp, ct.c_double]
    L.nw_set_refs.argtypes = [vp, ct.c_int, ct.c_int, vp, vp, vp]
    L.nw_build.argtypes = [vp, ct.c_double, vp, vp]
    L.nw_hs_batched.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_uint64, vp, vp, ct.c_int]
    L.nw_hs_batched_noalloc.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_uint64, vp, vp, ct.c_int]
    L.nw_s_same_batched.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_uint64, ct.c_int, vp]
    L.nw_m_batched.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_uint64, ct.c_double, vp, ct.c_int]
    L.nw_hs_matrix.argtypes = [vp, vp, ct.c_uint64, vp, vp, ct.c_int]
    L.nw_hs_matrix_streamed.argtypes = [vp, ct.c_double, vp, ct.c_uint64, vp, vp, ct.c_int, ct.c_uint64, vp, vp]
    L.nw_pair_slab.argtypes = [vp, ct.c_int, ct.c_int, vp, vp, vp]
    L.nw_set_precision.argtypes = [vp, ct.c_int]
    L.nw_set_eri_factors.argtypes = [vp, vp, ct.c_int, ct.c_double, dp]
    L.nw_rank4_coverage.argtypes = [vp, ct.c_int, ct.c_int, dp]
    L.nw_rank4_stats.argtypes = [vp, dp]
    L.nw_rank4_kernel_stats.argtypes = [vp, dp]
    L.nw_hs_naive_batched.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_uint64, vp, vp]
    L.nw_hs_compare.argtypes = [vp, vp, ct.c_uint64, vp, vp, ct.c_int, dp, dp]
    L.nw_project_hs.argtypes = [vp, vp, ct.c_uint64, ct.c_uint64, vp, vp, vp, ct.c_int]
    L.nw_noci_fock.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_double, vp, vp, vp, vp]
    L.nw_save_wicks.argtypes = [vp, ct.c_char_p]
    L.nw_load_wicks.argtypes = [vp, ct.c_char_p, ct.c_double]
    L.nw_set_mixed_floor.argtypes = [vp, ct.c_double]
    L.nw_update_fock.argtypes = [vp, vp, vp]
    L.nw_fs_matrix.argtypes = [vp, vp, ct.c_uint64, vp, vp, ct.c_int]
    L.nw_fs_batched.argtypes = [vp, vp, ct.c_uint64, vp, ct.c_uint64, vp, vp, ct.c_int]
    L.nw_matvec_m.argtypes = [vp, vp, ct.c_uint64, ct.c_double, vp, vp, vp, ct.c_int]
    L.nw_last_timings.argtypes = [vp, dp]
    L.nw_last_error.argtypes = [vp]
    L.nw_last_error.restype = ct.c_char_p
    L.nw_destroy.argtypes = [vp]
    L.nw_destroy.restype = None
    L.nw_factor_create.argtypes = [vp, vp, ct.c_uint64, ct.c_int, ct.c_int, ct.POINTER(vp)]
    L.nw_factor_refresh.argtypes = [vp]
    L.nw_factor_apply.argtypes = [vp, ct.c_double, vp, vp, vp]
    L.nw_factor_diagonals.argtypes = [vp, ct.c_double, vp, vp]
    L.nw_factor_info.argtypes = [vp, ct.c_int, u64p, u64p, u64p]
    L.nw_factor_ids.argtypes = [vp, vp, vp]
    L.nw_factor_table.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, vp]
    L.nw_factor_last_ms.argtypes = [vp, dp]
    L.nw_factor_destroy.argtypes = [vp]
    L.nw_factor_destroy.restype = None
    L.nw_factor_set_precision.argtypes = [vp, ct.c_int]
    L.nw_factor_apply_pair.argtypes = [vp, ct.c_double, vp, vp, vp]
    L.nw_factor_apply_bytes.argtypes = [vp, u64p]
    L.nw_factor_block_groups.argtypes = [vp, u64p, u64p]
    L.nw_batch_create.argtypes = [ct.POINTER(vp), ct.c_int]
    L.nw_batch_add.argtypes = [vp, vp, vp, ct.c_uint64, ct.c_int, ct.POINTER(ct.c_int)]
    L.nw_batch_plan.argtypes = [vp, ct.c_int]
    L.nw_batch_run.argtypes = [vp]
    L.nw_batch_result.argtypes = [vp, ct.c_int, vp, vp]
    L.nw_batch_schedule.argtypes = [vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), vp, ct.c_int,
                                    ct.POINTER(ct.c_int)]
    L.nw_batch_set_schedule.argtypes = [vp, ct.c_int, ct.c_int, ct.c_int, vp, ct.c_int]
    L.nw_batch_last_ms.argtypes = [vp, dp, ct.POINTER(ct.c_longlong)]
    L.nw_batch_last_error.argtypes = [vp]
    L.nw_batch_last_error.restype = ct.c_char_p
    L.nw_batch_destroy.argtypes = [vp]
    L.nw_batch_destroy.restype = None
    for name in EXPORTS:
        getattr(L, name)
    _lib = L
    return L


def sync_torch(*arrays):
    """The engine runs on its own HIP stream and returns only after that stream
    has finished; torch device tensors handed to it may still be in flight on
    torch's current stream, so finish that stream first (reads of inputs and
    writes of outputs both depend on it)."""
    for a in arrays:
        if a is not None and getattr(a, "is_cuda", False):
            import torch
            torch.cuda.current_stream(a.device).synchronize()
            return


def _ptr(a):
    """Host numpy array or torch tensor (device or host) -> void pointer."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ct.c_void_p(a.data_ptr())
    return ct.c_void_p(a.ctypes.data)


class Engine:
    """One context on one GPU: AO integrals + references + slab."""

    def __init__(self, S, h, eri, enuc=0.0, device=0):
        self.L = load()
        self.n_bas = S.shape[0]
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (S, h, eri)]
        ctx = ct.c_void_p()
        rc = self.L.nw_create(ct.byref(ctx), device, self.n_bas, *[_ptr(x) for x in self._keep], float(enuc))
        if rc != 0:
            raise NwickError(rc, "nw_create failed")
        self.ctx = ctx
        self.nref = 0

    @classmethod
    def from_system(cls, sysm, device=0):
        eng = cls(sysm.S, sysm.h, sysm.eri, sysm.enuc, device)
        eng.set_refs(sysm.refs, sysm.nocc_a, sysm.nocc_b)
        return eng

    def _check(self, rc):
        if rc != 0:
            raise NwickError(rc, self.L.nw_last_error(self.ctx).decode())

    def set_refs(self, refs, nocc_a, nocc_b):
        nref = len(refs)
        nmo = refs[0][0].shape[1]
        ca = np.ascontiguousarray(np.stack([r[0] for r in refs]), dtype=np.float64)
        cb = np.ascontiguousarray(np.stack([r[1] for r in refs]), dtype=np.float64)
        occ = np.zeros((nref, 4), dtype=np.uint64)
        for i in range(nref):
            for k, no in enumerate((nocc_a, nocc_b)):
                v = (1 << no) - 1
                occ[i, 2 * k] = v & 0xFFFFFFFFFFFFFFFF
                occ[i, 2 * k + 1] = v >> 64
        self._check(self.L.nw_set_refs(self.ctx, nref, nmo, _ptr(ca), _ptr(cb), _ptr(occ)))
        self.nref, self.nmo = nref, nmo

    def set_eri_factors(self, L, check_tol=1e-10):
        """Install integral factors (ab|cd) = sum_K L[K,a,b] L[K,c,d] (L: naux x n x n,
        symmetric per K; None removes them).  The next build leaves the rank-4
        blocks of m = 0 pairs to the evaluation calls, which build the (row, row)
        blocks their determinants read.  Returns the largest sampled deviation
        from the context's dense integrals."""
        dev = ct.c_double(0.0)
        if L is None:
            self._check(self.L.nw_set_eri_factors(self.ctx, None, 0, float(check_tol), ct.byref(dev)))
            self._fac = None
            return 0.0
        if hasattr(L, "data_ptr"):
            naux = int(L.shape[0])
            sync_torch(L)
            self._fac = L
        else:
            L = np.ascontiguousarray(L, dtype=np.float64)
            naux = L.shape[0]
            self._fac = L
        self._check(self.L.nw_set_eri_factors(self.ctx, _ptr(L), naux, float(check_tol), ct.byref(dev)))
        return dev.value

    def rank4_coverage(self, x, w):
        """Fraction of the (row, row) blocks built of pair (x, w)'s alpha J, beta J and II tensors."""
        out = (ct.c_double * 3)()
        self._check(self.L.nw_rank4_coverage(self.ctx, int(x), int(w), out))
        return tuple(out)

    def rank4_stats(self):
        """On-demand rank-4 work of the last call that built blocks: device ms and algorithmic flops per stage."""
        out = (ct.c_double * 10)()
        self._check(self.L.nw_rank4_stats(self.ctx, out))
        keys = ("ms_total", "ms_factor", "ms_rect", "ms_single", "ms_exchange", "flops_rect", "flops_single",
                "flops_exchange", "flops_factor", "single_blocks")
        return dict(zip(keys, out))

    def rank4_kernel_stats(self):
        """The same call's block products by launch class (FP64 / FP32 pairs, large / narrow tiles): ms and flops."""
        out = (ct.c_double * 8)()
        self._check(self.L.nw_rank4_kernel_stats(self.ctx, out))
        keys = ("ms_f64_large", "ms_f64_narrow", "ms_f32_large", "ms_f32_narrow",
                "flops_f64_large", "flops_f64_narrow", "flops_f32_large", "flops_f32_narrow")
        return dict(zip(keys, out))

    def set_precision(self, prec):
        """PREC_FP64 or PREC_MIXED (FP32 MFMA four-index transforms)."""
        self._check(self.L.nw_set_precision(self.ctx, int(prec)))

    def set_mixed_floor(self, sigma_floor):
        """Pairs with a nonzero occupied-overlap singular value below sigma_floor keep FP64 transforms."""
        self._check(self.L.nw_set_mixed_floor(self.ctx, float(sigma_floor)))

    def build(self, tol=1e-8, pair_mask=None):
        """build_wicks_shared: returns {(x, w): PairMeta} for the built pairs."""
        meta = (PairMeta * (self.nref * self.nref))()
        mask = None
        if pair_mask is not None:
            mask = np.ascontiguousarray(pair_mask, dtype=np.uint8).reshape(-1)
        self._check(self.L.nw_build(self.ctx, float(tol), _ptr(mask), ct.cast(meta, ct.c_void_p)))
        out = {}
        for x in range(self.nref):
            for w in range(self.nref):
                if mask is None or mask[x * self.nref + w]:
                    out[(x, w)] = meta[x * self.nref + w]
        return out

    def hs_matrix(self, dets, mode=MODE_REFERENCE, out=None):
        """Full H, S (pairs i <= j, mirrored).  dets: DET_DTYPE array or a device tensor of records."""
        n = len(dets) if not hasattr(dets, "data_ptr") else dets.numel() * dets.element_size() // DET_DTYPE.itemsize
        if out is None:
            H = np.empty((n, n)); S = np.empty((n, n))
        else:
            H, S = out
        sync_torch(dets, H, S)
        self._check(self.L.nw_hs_matrix(self.ctx, _ptr(dets), n, _ptr(H), _ptr(S), int(mode)))
        return H, S

    def hs_matrix_streamed(self, dets, tol=1e-8, mode=MODE_REFERENCE, slab_bytes=0, out=None, meta=None):
        """Build and evaluate in groups of reference pairs whose slab blocks fit
        in slab_bytes (0: what device memory allows); returns H, S and the
        number of groups.  meta: a dict that receives {(x, w): PairMeta} of
        every pair built."""
        n = len(dets) if not hasattr(dets, "data_ptr") else dets.numel() * dets.element_size() // DET_DTYPE.itemsize
        if out is None:
            H = np.empty((n, n)); S = np.empty((n, n))
        else:
            H, S = out
        groups = ct.c_int()
        pm = (PairMeta * (self.nref * self.nref))()
        ct.memset(pm, 0xFF, ct.sizeof(pm))  # ma = -1 marks pairs not built
        sync_torch(dets, H, S)
        self._check(self.L.nw_hs_matrix_streamed(self.ctx, float(tol), _ptr(dets), n, _ptr(H), _ptr(S), int(mode),
                                                 int(slab_bytes), ct.cast(pm, ct.c_void_p), ct.byref(groups)))
        if meta is not None:
            for p in range(self.nref * self.nref):
                if pm[p].ma >= 0:
                    meta[divmod(p, self.nref)] = pm[p]
        return H, S, groups.value

    def hs_batched(self, dets, req, mode=MODE_REFERENCE):
        req = np.ascontiguousarray(req, dtype=np.uint64).reshape(-1, 2)
        h = np.empty(len(req)); s = np.empty(len(req))
        sync_torch(dets)
        self._check(self.L.nw_hs_batched(self.ctx, _ptr(dets), len(dets), _ptr(req), len(req), _ptr(h), _ptr(s),
                                         int(mode)))
        return h, s

    def hs_batched_device(self, dets, req, h, s, mode=MODE_REFERENCE, noalloc=False):
        """nw_hs_batched (or nw_hs_batched_noalloc) on device tensors: dets a
        uint8 tensor of records, req an int64 tensor (n, 2), h and s float64."""
        ndet = dets.numel() * dets.element_size() // DET_DTYPE.itemsize
        nreq = req.numel() // 2
        sync_torch(dets, req, h, s)
        fn = self.L.nw_hs_batched_noalloc if noalloc else self.L.nw_hs_batched
        self._check(fn(self.ctx, _ptr(dets), ndet, _ptr(req), nreq, _ptr(h), _ptr(s), int(mode)))
        return h, s

    def hs_naive_batched(self, dets, req):
        """Generalised Slater-Condon H, S (calculate_hs_pair_naive) of request
        pairs, from the references and integrals (no build needed)."""
        req = np.ascontiguousarray(req, dtype=np.uint64).reshape(-1, 2)
        dets = np.ascontiguousarray(dets)
        h = np.empty(len(req)); s = np.empty(len(req))
        self._check(self.L.nw_hs_naive_batched(self.ctx, _ptr(dets), len(dets), _ptr(req), len(req), _ptr(h),
                                               _ptr(s)))
        return h, s

    def hs_compare(self, dets, mode=MODE_REFERENCE):
        """wicks.compare: Wick's H, S of every pair plus the naive route's
        discrepancy over the upper triangle: (H, S, total, max_element)."""
        dets = np.ascontiguousarray(dets)
        n = len(dets)
        H = np.empty((n, n)); S = np.empty((n, n))
        tot = ct.c_double(); mx = ct.c_double()
        self._check(self.L.nw_hs_compare(self.ctx, _ptr(dets), n, _ptr(H), _ptr(S), int(mode), ct.byref(tot),
                                         ct.byref(mx)))
        return H, S, tot.value, mx.value

    def s_same_batched(self, dets, req, spin):
        """Per-spin overlap factors (excitation phases applied) of request pairs."""
        req = np.ascontiguousarray(req, dtype=np.uint64).reshape(-1, 2)
        s = np.empty(len(req))
        sync_torch(dets)
        self._check(self.L.nw_s_same_batched(self.ctx, _ptr(dets), len(dets), _ptr(req), len(req), int(spin),
                                             _ptr(s)))
        return s

    def m_batched(self, dets, req, e0, mode=MODE_REFERENCE):
        """M = F - e0 S of request pairs (calculate_m_pair_wicks)."""
        req = np.ascontiguousarray(req, dtype=np.uint64).reshape(-1, 2)
        m = np.empty(len(req))
        sync_torch(dets)
        self._check(self.L.nw_m_batched(self.ctx, _ptr(dets), len(dets), _ptr(req), len(req), float(e0), _ptr(m),
                                        int(mode)))
        return m

    def update_fock(self, fa, fb):
        """update_wicks_fock: generalised-Fock intermediates from spin Fock matrices."""
        fa = np.ascontiguousarray(fa, dtype=np.float64)
        fb = np.ascontiguousarray(fb, dtype=np.float64)
        self._check(self.L.nw_update_fock(self.ctx, _ptr(fa), _ptr(fb)))

    @staticmethod
    def _ndet(dets):
        return len(dets) if not hasattr(dets, "data_ptr") else dets.numel() * dets.element_size() // DET_DTYPE.itemsize

    def fs_matrix(self, dets, mode=MODE_REFERENCE, out=None):
        n = self._ndet(dets)
        F, S = (np.empty((n, n)), np.empty((n, n))) if out is None else out
        sync_torch(dets, F, S)
        self._check(self.L.nw_fs_matrix(self.ctx, _ptr(dets), n, _ptr(F), _ptr(S), int(mode)))
        return F, S

    def fs_batched(self, dets, req, mode=MODE_REFERENCE):
        req = np.ascontiguousarray(req, dtype=np.uint64).reshape(-1, 2)
        f = np.empty(len(req)); s = np.empty(len(req))
        sync_torch(dets)
        self._check(self.L.nw_fs_batched(self.ctx, _ptr(dets), self._ndet(dets), _ptr(req), len(req), _ptr(f),
                                         _ptr(s), int(mode)))
        return f, s

    def matvec_m(self, dets, e0, x, y=None, ys=None, mode=MODE_REFERENCE):
        """y = (F - e0 S) x (and ys = S x if ys is given); arrays may be torch device tensors."""
        n = self._ndet(dets)
        if y is None:
            y = np.empty(n)
        sync_torch(dets, x, y, ys)
        self._check(self.L.nw_matvec_m(self.ctx, _ptr(dets), n, float(e0), _ptr(x), _ptr(y), _ptr(ys), int(mode)))
        return y if ys is None else (y, ys)

    def factor(self, dets, mode=MODE_REFERENCE, storage=FACTOR_DEVICE):
        """Spin-factorised one-body / overlap operator over dets (nw_factor_*)."""
        return Factor(self, dets, mode, storage)

    def project_hs(self, dets, ncur, coef, v=None, s=None, mode=MODE_REFERENCE):
        """v_a = sum_i H_ai c_i, s_a = sum_i S_ai c_i for every candidate a (dets[ncur:]) over the
        current space dets[:ncur]; arrays may be torch device tensors."""
        n = self._ndet(dets)
        nc = n - ncur
        v = np.empty(nc) if v is None else v
        s = np.empty(nc) if s is None else s
        coef = coef if hasattr(coef, "data_ptr") else np.ascontiguousarray(coef, dtype=np.float64)
        sync_torch(dets, coef, v, s)
        self._check(self.L.nw_project_hs(self.ctx, _ptr(dets), n, int(ncur), _ptr(coef), _ptr(v), _ptr(s), int(mode)))
        return v, s

    def noci_fock(self, dets, coef, tol=1e-8, densities=False):
        """Spin Fock matrices (and optionally NOCI densities) of the state sum_i coef_i |dets_i>."""
        n = self._ndet(dets)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        fa = np.empty((self.n_bas, self.n_bas)); fb = np.empty((self.n_bas, self.n_bas))
        da = np.empty_like(fa) if densities else None
        db = np.empty_like(fa) if densities else None
        sync_torch(dets)
        self._check(self.L.nw_noci_fock(self.ctx, _ptr(dets), n, _ptr(coef), float(tol), _ptr(fa), _ptr(fb),
                                        _ptr(da), _ptr(db)))
        return (fa, fb, da, db) if densities else (fa, fb)

    def save_wicks(self, directory):
        """Write wicks.bin + wicks.meta (the reference's disk cache, version 2) into directory."""
        self._check(self.L.nw_save_wicks(self.ctx, os.fsencode(directory)))

    def load_wicks(self, directory, tol=1e-8):
        """Replace the build by a reference-format disk cache from directory."""
        self._check(self.L.nw_load_wicks(self.ctx, os.fsencode(directory), float(tol)))

    def pair_slab(self, x, w):
        off = np.zeros(95, dtype=np.int64)
        meta = np.zeros(32)
        self._check(self.L.nw_pair_slab(self.ctx, x, w, None, _ptr(off), _ptr(meta)))
        out = np.empty(int(off[94]))
        self._check(self.L.nw_pair_slab(self.ctx, x, w, _ptr(out), _ptr(off), _ptr(meta)))
        return out, off, meta

    def timings(self):
        t = (ct.c_double * 3)()
        self.L.nw_last_timings(self.ctx, t)
        return {"total_ms": t[0], "transform_ms": t[1], "eval_ms": t[2]}

    def close(self):
        if getattr(self, "ctx", None):
            self.L.nw_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Factor:
    """Spin-factorised one-body and overlap operators of one determinant list
    (nw_factor_*): per-parent alpha/beta components and per-spin {S, F}
    tables for every ordered parent pair, applied as y = (F + lambda S) x."""

    def __init__(self, eng, dets, mode=MODE_REFERENCE, storage=FACTOR_DEVICE):
        self.eng, self.L = eng, eng.L
        self.n = Engine._ndet(dets)
        h = ct.c_void_p()
        sync_torch(dets)
        eng._check(self.L.nw_factor_create(eng.ctx, _ptr(dets), self.n, int(mode), int(storage), ct.byref(h)))
        self.h = h
        self.precision = PREC_FP64

    def _check(self, rc):
        self.eng._check(rc)

    def refresh(self):
        self._check(self.L.nw_factor_refresh(self.h))

    def set_precision(self, precision):
        """Tables apply reads: PREC_FP64 or PREC_MIXED (FP32 tables, FP64 sums)."""
        self._check(self.L.nw_factor_set_precision(self.h, int(precision)))
        self.precision = int(precision)

    def apply(self, lam, x, y=None, ys=None):
        """y = (F + lam S) x, and ys = S x if ys is given; arrays may be torch device tensors."""
        if y is None:
            y = np.empty(self.n)
        sync_torch(x, y, ys)
        self._check(self.L.nw_factor_apply(self.h, float(lam), _ptr(x), _ptr(y), _ptr(ys)))
        return y if ys is None else (y, ys)

    def apply_pair(self, lam, x, y=None, ys=None):
        """The same operator on two vectors in one pass over the tables: x of
        complex dtype (real and imaginary parts) or shape [n, 2], contiguous;
        y, ys likewise (ys = S x if given)."""
        def flat(a):
            if a is None:
                return None
            if hasattr(a, "is_complex") and a.is_complex():
                import torch
                a = torch.view_as_real(a)
            elif isinstance(a, np.ndarray) and np.iscomplexobj(a):
                a = a.view(np.float64).reshape(-1, 2)
            if tuple(a.shape) != (self.n, 2):
                raise ValueError("apply_pair takes n complex values or an [n, 2] array")
            if not (a.is_contiguous() if hasattr(a, "is_contiguous") else a.flags.c_contiguous):
                raise ValueError("apply_pair needs contiguous arrays")
            return a
        if y is None:
            y = np.empty((self.n, 2))
        xf, yf, ysf = flat(x), flat(y), flat(ys)
        sync_torch(xf, yf, ysf)
        self._check(self.L.nw_factor_apply_pair(self.h, float(lam), _ptr(xf), _ptr(yf), _ptr(ysf)))
        return y if ys is None else (y, ys)

    def diagonals(self, lam, mdiag=None, sdiag=None):
        mdiag = np.empty(self.n) if mdiag is None else mdiag
        sdiag = np.empty(self.n) if sdiag is None else sdiag
        sync_torch(mdiag, sdiag)
        self._check(self.L.nw_factor_diagonals(self.h, float(lam), _ptr(mdiag), _ptr(sdiag)))
        return mdiag, sdiag

    def info(self, parent=-1):
        na, nb, tb = ct.c_uint64(), ct.c_uint64(), ct.c_uint64()
        self._check(self.L.nw_factor_info(self.h, int(parent), ct.byref(na), ct.byref(nb), ct.byref(tb)))
        ab, bg, bk = ct.c_uint64(), ct.c_uint64(), ct.c_uint64()
        self._check(self.L.nw_factor_apply_bytes(self.h, ct.byref(ab)))
        self._check(self.L.nw_factor_block_groups(self.h, ct.byref(bg), ct.byref(bk)))
        return {"ncomp_a": na.value, "ncomp_b": nb.value, "table_bytes": tb.value, "apply_bytes": ab.value,
                "block_groups": bg.value, "blocks": bk.value}

    def ids(self):
        a = np.empty(self.n, dtype=np.uint32); b = np.empty(self.n, dtype=np.uint32)
        self._check(self.L.nw_factor_ids(self.h, _ptr(a), _ptr(b)))
        return a, b

    def table(self, target, source, spin):
        """[source component, target component, {S, F}] factor table."""
        nt = self.info(target)["ncomp_a" if spin == 0 else "ncomp_b"]
        ns = self.info(source)["ncomp_a" if spin == 0 else "ncomp_b"]
        out = np.empty((ns, nt, 2))
        self._check(self.L.nw_factor_table(self.h, int(target), int(source), int(spin), _ptr(out)))
        return out

    def last_ms(self):
        ms = (ct.c_double * 2)()
        self._check(self.L.nw_factor_last_ms(self.h, ms))
        return {"total_ms": ms[0], "regen_ms": ms[1]}

    def close(self):
        if getattr(self, "h", None):
            # a closed engine has already freed its factors (nw_destroy)
            if getattr(self.eng, "ctx", None):
                self.L.nw_factor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several built engines evaluated together (nw_batch_*): one interleaved
    plan of persistent fused launches over every system's pairs, optionally
    tuned and captured into a HIP graph."""

    def __init__(self, device=0):
        self.L = load()
        b = ct.c_void_p()
        rc = self.L.nw_batch_create(ct.byref(b), device)
        if rc != 0:
            raise NwickError(rc, "nw_batch_create failed")
        self.b = b
        self.engines, self.sizes = [], []

    def _check(self, rc):
        if rc != 0:
            raise NwickError(rc, self.L.nw_batch_last_error(self.b).decode())

    def add(self, eng, dets, mode=MODE_REFERENCE):
        idx = ct.c_int()
        n = Engine._ndet(dets)
        sync_torch(dets)
        self._check(self.L.nw_batch_add(self.b, eng.ctx, _ptr(dets), n, int(mode), ct.byref(idx)))
        self.engines.append(eng)
        self.sizes.append(n)
        return idx.value

    def plan(self, tune=False, graph=True):
        self._check(self.L.nw_batch_plan(self.b, (BATCH_TUNE if tune else 0) | (BATCH_GRAPH if graph else 0)))

    def run(self):
        self._check(self.L.nw_batch_run(self.b))

    def result(self, index, out=None):
        n = self.sizes[index]
        H, S = (np.empty((n, n)), np.empty((n, n))) if out is None else out
        sync_torch(H, S)
        self._check(self.L.nw_batch_result(self.b, index, _ptr(H), _ptr(S)))
        return H, S

    def schedule(self):
        g, d, gp, cnt = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        self._check(self.L.nw_batch_schedule(self.b, ct.byref(g), ct.byref(d), ct.byref(gp), None, 0, ct.byref(cnt)))
        ent = (TuneEntry * max(cnt.value, 1))()
        self._check(self.L.nw_batch_schedule(self.b, None, None, None, ct.cast(ent, ct.c_void_p), cnt.value, None))
        return {"grouping": g.value, "dynamic": d.value, "grid_per_cu": gp.value,
                "buckets": [{"nocc": e.nocc, "code": e.code, "la": (e.code >> 1) // 5, "lb": (e.code >> 1) % 5,
                             "m0": e.code & 1, "shape": e.shape, "ms_best": e.ms_best, "ms_default": e.ms_default}
                            for e in ent[:cnt.value]]}

    def set_schedule(self, sched):
        bk = sched.get("buckets", [])
        ent = (TuneEntry * max(len(bk), 1))()
        for k, e in enumerate(bk):
            ent[k].nocc, ent[k].code, ent[k].shape = e["nocc"], e["code"], e["shape"]
        self._check(self.L.nw_batch_set_schedule(self.b, sched["grouping"], sched["dynamic"], sched["grid_per_cu"],
                                                 ct.cast(ent, ct.c_void_p), len(bk)))

    def last(self):
        ms, n = ct.c_double(), ct.c_longlong()
        self._check(self.L.nw_batch_last_ms(self.b, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "b", None):
            self.L.nw_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
