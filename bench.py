#!/usr/bin/env python3
"""Benchmark: nonorthogonal-Wick H,S pair elements per second on one MI355X.

Workload (the synthetic batched-Wick configuration): n_bas = nmo = 100,
10 alpha + 10 beta electrons, orthonormal AO basis, ERIs from 300 seeded
Cholesky-like vectors, 8 references C0 exp(kappa), one reference with an
alpha occupied orbital exactly orthogonal to every other reference (m = 1
pairs), 512 random singles/doubles per reference: 4096 determinants and
8,390,656 upper-triangle pairs.

One step = the whole hot path on inputs resident in HBM: the per-reference-
pair build for every ordered pair the upper triangle needs (x <= w: 36 pairs)
followed by the evaluation of all 8,390,656 H and S elements into device
matrices.  value = pairs / second of that end-to-end step.  The step runs in
dense forms (every rank-4 tensor transformed from the dense ERIs; FP64, FP32
and bf16x3 transforms) and in factored forms (the workload's own Cholesky
vectors installed with nw_set_eri_factors: the rank-4 blocks the
determinants read built on demand, FP64 and FP32 products); value = the
fastest mixed form whose NOCI energy is within 1 uHa of the dense FP64 one.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "noci-rs_amd"))

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """CPU model, logical CPUs of the machine and of this process."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff}


def openblas_quarter_pass(n, rows=None):
    """GFLOP/s of numpy's bundled OpenBLAS (the GEMM library the reference
    links through ndarray-linalg / openblas-src) on one quarter of the n^4
    transform, rows x n times n x n: the rate a BLAS-based CPU build reaches."""
    import numpy as np
    rows = rows or n ** 3
    X = np.random.default_rng(0).random((rows, n))
    c = np.random.default_rng(1).random((n, n))
    X @ c
    t = time.perf_counter()
    X @ c
    return 2.0 * rows * n * n / (time.perf_counter() - t) / 1e9


def transform_flops(nb, ma, mb):
    """Rank-4 transforms of one ordered pair's CPU build in the restatement:
    two per same-spin J slot (direct and exchange, as build.rs:267-313) and
    one per II block, 8 nb^5 flop each."""
    slots = lambda m: sum(1 for b in [(0, 0, 0, 0), (0, 0, 0, 1), (0, 0, 1, 0), (0, 0, 1, 1), (0, 1, 0, 1),
                                      (0, 1, 1, 0), (0, 1, 1, 1), (1, 0, 1, 0), (1, 0, 1, 1), (1, 1, 1, 1)]
                          if sum(b) <= m)
    nii = sum(1 for a0 in (0, 1) for aj in (0, 1) for b0 in (0, 1) for bj in (0, 1) if a0 + aj <= ma and b0 + bj <= mb)
    return (2 * (slots(ma) + slots(mb)) + nii) * 8.0 * nb ** 5


def tensors_of(ma, mb):
    """Rank-4 tensors one ordered pair's GPU build transforms: one per
    same-spin J slot (allocated when its branch sum is at most m,
    build.rs:19-30) and one per II block (a0 + aj <= m_a, b0 + bj <= m_b)."""
    slots = lambda m: sum(1 for b in [(0, 0, 0, 0), (0, 0, 0, 1), (0, 0, 1, 0), (0, 0, 1, 1), (0, 1, 0, 1),
                                      (0, 1, 1, 0), (0, 1, 1, 1), (1, 0, 1, 0), (1, 0, 1, 1), (1, 1, 1, 1)]
                          if sum(b) <= m)
    nii = sum(1 for a0 in (0, 1) for aj in (0, 1) for b0 in (0, 1) for bj in (0, 1) if a0 + aj <= ma and b0 + bj <= mb)
    return slots(ma) + slots(mb) + nii


def eval_bytes(la, lb):
    """Algorithmic bytes of one determinant-pair element of contraction ranks
    (la, lb) on an m = 0 reference pair: the X/Y entries of both spins, the
    one-column hcol0 entries, the J and II gathers, the two outputs and the
    two packed determinants (the per-element formula of the evaluation
    roofline).  Numpy arrays or ints."""
    c2 = lambda l: l * (l - 1) // 2
    return 8.0 * (2 * (la * la + lb * lb) + 2 * c2(la) ** 2 + 2 * c2(lb) ** 2 + la * la * lb * lb) + 60.0


def cpu_baseline(sysm, dets, tol, mode, with_build=True):
    """Time the CPU oracle on a bounded sample of the same workload: the
    build of one ordered reference pair (0, 1) plus the evaluation of all
    determinant pairs of that block (with_build=False: the evaluation only,
    the rate of requests against a resident slab)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as ct

    import numpy as np

    import oracle as O

    L = O.lib()
    L.ow_hs_many.argtypes = [O.D, O.I64, O.D, ct.c_int, ct.c_longlong, O.U64, O.U64, O.D, ct.c_double,
                             ct.c_double, ct.c_int, O.D, O.D]
    L.ow_num_threads.restype = ct.c_int
    threads = L.ow_num_threads()
    # the build's quarter passes on OpenBLAS dgemm, the reference's GEMM library
    blas = O.use_openblas(True) if with_build else None
    t0 = time.perf_counter()
    slab = O.Slab(sysm, tol=tol, pairs=[(0, 1)])
    t1 = time.perf_counter()
    bra = [d for d in dets if d.parent == 0]
    ket = [d for d in dets if d.parent == 1]
    npair = len(bra) * len(ket)
    xex = np.zeros((npair, 8), dtype=np.uint64)
    wex = np.zeros((npair, 8), dtype=np.uint64)
    exph = np.zeros(npair)

    def m8(d):
        return np.concatenate([O.mask2(d.ha), O.mask2(d.pa), O.mask2(d.hb), O.mask2(d.pb)])
    bm = np.stack([m8(d) for d in bra])
    km = np.stack([m8(d) for d in ket])
    bp = np.array([d.pha * d.phb for d in bra])
    kp = np.array([d.pha * d.phb for d in ket])
    xex[:] = np.repeat(bm, len(ket), axis=0)
    wex[:] = np.tile(km, (len(bra), 1))
    exph[:] = np.repeat(bp, len(ket)) * np.tile(kp, len(bra))
    H = np.zeros(npair); S = np.zeros(npair)
    sub, off, meta = slab.pair_view(0, 1)
    t2 = time.perf_counter()
    L.ow_hs_many(O._d(sub), off.ctypes.data_as(O.I64), O._d(meta), sysm.nmo, npair, O._u(xex), O._u(wex),
                 O._d(exph), sysm.enuc, tol, mode, O._d(H), O._d(S))
    t3 = time.perf_counter()
    host = host_info()
    if not with_build:
        return dict({"value": npair / (t3 - t2), "unit": "pairs/s", "cores": threads, "kind": "port",
                     "sample": f"oracle evaluation of the {len(bra)}x{len(ket)} determinant pairs of ordered "
                               f"reference pair (0,1) on its resident slab block ({t3 - t2:.2f} s; build "
                               f"{t1 - t0:.1f} s not counted), OpenMP threads={threads}"}, **host)
    secs = (t1 - t0) + (t3 - t2)
    ma, mb = slab.m[(0, 1)]
    fl = transform_flops(sysm.n_bas, ma, mb)
    return dict({"value": npair / secs, "unit": "pairs/s", "cores": threads, "kind": "port",
                 "sample": f"oracle build of ordered reference pair (0,1) ({t1 - t0:.1f} s; four-index transforms "
                           f"on {blas} dgemm) + evaluation of its {len(bra)}x{len(ket)} determinant pairs "
                           f"({t3 - t2:.2f} s), OpenMP threads={threads}",
                 "build_gemm": blas,
                 "build_transform_gflops": fl / (t1 - t0) / 1e9,
                 "openblas_quarter_pass_gflops": openblas_quarter_pass(sysm.n_bas)}, **host)


def pt2_workload(args):
    """NOCI-PT2 matvec sized like F2/cc-pVDZ: n_bas = 28, 9 + 9 electrons,
    6 references (kappa ~ N(0, 0.4)), candidates = all singles + 20,000 seeded
    doubles per reference; current space = the references.  Times the
    matrix-free (F - E0 S) x product over all candidate pairs two ways: the
    spin-factorised operator with device-resident per-spin tables (the
    reference's default PT2 path, factor_tables = RAM) and the pairwise
    on-the-fly evaluation of every pair (full_m = None without tables), with
    FP64 tables and with their FP32 copies.  Then solves the PT2 equations
    both ways (restarted GMRES, restart 30, RMS residual 1e-8): FP64, and
    with the Arnoldi matvecs on the FP32 tables and the restart residuals
    (iterative refinement) on the FP64 tables; unshifted, where this
    synthetic operator (rotated references, many negative diagonal entries)
    stagnates, and with the imaginary shift --pt2-shift (snoci imag_shifts),
    where it converges.
    value = candidate pairs represented per second by the fastest matvec
    whose shifted PT2 energy is within 1e-6 Ha of the FP64 one."""
    import numpy as np
    import torch

    import nwick
    from nwick import basis, synth
    from nwick.pt2 import PT2

    sysm = synth.make_system(28, 9, 9, 6, seed=args.seed, kappa_sigma=0.4)
    refs = basis.excited_basis(6, 28, 9, 9, per_ref=0, include_refs=True, counts={})
    rng = np.random.default_rng(args.seed)
    cands = []
    for r in range(6):
        singles = [e for e in basis.all_excitations(28, 9, 9, orders=(1,))]
        doubles = basis.random_excitations(rng, 28, 9, 9, {"aa": 5000, "bb": 5000, "ab": 10000})
        exc = sorted(singles + doubles, key=lambda e: basis.generation_key(9, 9, *e))
        cands += [basis.Det(r, (1 << 9) - 1, (1 << 9) - 1, *e) for e in exc]
    n = len(cands)
    log(f"pt2 workload: {n} candidates, {n * (n + 1) // 2} pairs per matvec")
    eng = nwick.Engine.from_system(sysm, device=0)
    eng.build()
    pt = PT2(eng, sysm, refs, cands, mode=args.mode, factor_tables="device", precision="mixed")
    pt.setup()
    assert pt.operator == "factorised"
    info = pt.fac.info()
    log(f"pt2 setup {pt.timings['setup_s']:.1f} s (factor tables {pt.timings['factor_s']:.2f} s, "
        f"{info['table_bytes'] / 2**30:.1f} GiB), E_NOCI {pt.ecurrent:.10f}, E0 {pt.e0:.10f}")
    x = torch.randn(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    yp = torch.empty_like(x)
    lam = -pt.e0
    for _ in range(args.warmup):
        pt.fac.apply(lam, x, y=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kern = []
    for k in range(args.steps):
        pt.fac.apply(lam, x, y=y)
        kern.append(pt.fac.last_ms()["total_ms"])
        log(f"factorised matvec {k}: {kern[-1]:.2f} ms device")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the same matvec on FP32 tables (FP64 sums)
    bytes64 = info["apply_bytes"]
    pt.fac.set_precision(nwick.PREC_MIXED)
    bytes32 = pt.fac.info()["apply_bytes"]
    y32 = torch.empty_like(x)
    for _ in range(args.warmup):
        pt.fac.apply(lam, x, y=y32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kern32 = []
    for k in range(args.steps):
        pt.fac.apply(lam, x, y=y32)
        kern32.append(pt.fac.last_ms()["total_ms"])
        log(f"factorised FP32-table matvec {k}: {kern32[-1]:.2f} ms device")
    torch.cuda.synchronize()
    el32 = time.perf_counter() - t0
    # two vectors (a complex vector's parts) in one pass over the tables, FP32 then FP64 tables
    xc = torch.complex(x, torch.randn_like(x))
    yc = torch.empty_like(xc)
    pair_ms = {}
    for name, prec in (("fp32", nwick.PREC_MIXED), ("fp64", nwick.PREC_FP64)):
        pt.fac.set_precision(prec)
        pt.fac.apply_pair(lam, xc, y=yc)
        ks = []
        for k in range(args.steps):
            pt.fac.apply_pair(lam, xc, y=yc)
            ks.append(pt.fac.last_ms()["total_ms"])
        pair_ms[name] = float(np.mean(ks))
        log(f"factorised two-vector matvec, {name} tables: {pair_ms[name]:.2f} ms device")
    pt.fac.set_precision(nwick.PREC_FP64)
    rel32 = float((y32 - y).abs().max() / y.abs().max())
    # pairwise on-the-fly matvec over the same operator (one warm-up, two timed)
    eng.matvec_m(pt.cand_rec, pt.e0, x, y=yp, mode=args.mode)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pk = []
    for k in range(2):
        eng.matvec_m(pt.cand_rec, pt.e0, x, y=yp, mode=args.mode)
        pk.append(eng.timings()["eval_ms"])
    torch.cuda.synchronize()
    el_p = (time.perf_counter() - t1) / 2
    rel = float((y - yp).abs().max() / yp.abs().max())
    log(f"pairwise matvec {np.mean(pk):.1f} ms device; factorised vs pairwise max rel diff {rel:.2e}; "
        f"FP32 vs FP64 tables {rel32:.2e}")
    # PT2 solves: FP64, and FP32 Arnoldi matvecs with FP64 restart residuals
    sol = {}
    shifts = (args.pt2_shift,) if getattr(args, "compact", False) else (0.0, args.pt2_shift)
    for eps in shifts:
        key = f"imag_shift_{eps:g}"
        sol[key] = {}
        for prec in ("fp64", "mixed"):
            pt.precision = prec
            r = pt.solve(res_tol=1e-8, restart=30, max_iter=args.pt2_max_iter, imag_shift=eps)
            sol[key][prec] = {"ept2": r["ept2"], "converged": bool(r["converged"]),
                              "residual_rms": r["residual_rms"], "iterations": r["iterations"],
                              "matvecs": r["matvecs"], "matvecs_fp32": r["matvecs_fp32"],
                              "matvecs_fp64": r["matvecs"] - r["matvecs_fp32"], "solve_s": pt.timings["gmres_s"]}
            log(f"PT2 solve eps={eps:g} {prec}: E_PT2 {r['ept2']:.10f} residual {r['residual_rms']:.2e} "
                f"converged {r['converged']} iterations {r['iterations']} matvecs {r['matvecs']} "
                f"(FP32 {r['matvecs_fp32']}) {pt.timings['gmres_s']:.2f} s")
    shifted = sol[f"imag_shift_{args.pt2_shift:g}"]
    de = abs(shifted["mixed"]["ept2"] - shifted["fp64"]["ept2"])
    mixed_ok = shifted["mixed"]["converged"] and shifted["fp64"]["converged"] and de < 1e-6
    pairs = n * (n + 1) // 2
    use32 = mixed_ok and el32 < el
    el_v, kern_v, bytes_v = (el32, kern32, bytes32) if use32 else (el, kern, bytes64)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = pt2_cpu_baseline(info, pt)
    out = {"metric": "NOCI-PT2 matrix-free M = F - E0 S matvecs per second (spin-factorised operator)",
           "value": args.steps / el_v, "unit": "matvecs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * el_v / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32-tables+f64" if use32 else "f64", "data": "synthetic",
           "config": {"workload": "NOCI-PT2 matvec, n_bas=28, nocc=9+9, 6 refs, all singles + 20000 doubles "
                                  "per ref", "candidates": n, "mode": args.mode, "operator": "spin-factorised"},
           "ms_per_matvec": 1e3 * el_v / args.steps,
           "roofline": {"bound": "hbm", "kernel": "k_fac_dual + k_fac_mat + k_fac_rows (factorised matvec)",
                        "achieved": bytes_v / (float(np.mean(kern_v)) * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                        "unit": "GB/s",
                        "frac": bytes_v / (float(np.mean(kern_v)) * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "traffic": None, "traffic_source": "profiles/factor_pt2_c4_hbm_traffic.txt",
                        "algorithmic_bytes": bytes_v,
                        "bytes_meaning": "table bytes one apply reads (each table row once; dual passes read a "
                                         "parent pair's table once for both directions)"},
           "kernel_ms": float(np.mean(kern_v)), "kernel_ms_fp64_tables": float(np.mean(kern)),
           "kernel_ms_fp32_tables": float(np.mean(kern32)), "fp32_vs_fp64_tables_max_rel": rel32,
           "kernel_ms_two_vectors_fp64_tables": pair_ms["fp64"], "kernel_ms_two_vectors_fp32_tables": pair_ms["fp32"],
           "table_build_s": pt.timings["factor_s"], "table_bytes": info["table_bytes"],
           "table_bytes_fp32": info["table_bytes"] // 2, "apply_bytes_fp64": bytes64, "apply_bytes_fp32": bytes32,
           "pairwise_kernel_ms": float(np.mean(pk)), "factorised_vs_pairwise_max_rel": rel,
           "pt2_solve": sol, "dE_pt2_mixed_minus_fp64": shifted["mixed"]["ept2"] - shifted["fp64"]["ept2"],
           "pt2_solve_speedup_mixed": shifted["fp64"]["solve_s"] / shifted["mixed"]["solve_s"],
           "e_noci": pt.ecurrent, "e0": pt.e0, "cpu_baseline": cpu}
    pt.fac.close()
    eng.close()
    return out


def pt2_cpu_baseline(info, pt):
    """CPU rate of the reference's default PT2 operator, the factorised apply
    over RAM tables (apply_one_body, src/noci/factorise/onebody.rs:290-345):
    its cost is one pass over the {S, F} tables.  Bounded sample: numpy
    (OpenBLAS) applies one parent pair's table of the workload's component
    counts (random values of the same shape) to a vector; the rate, scaled
    to the resident table bytes of one apply (the GPU's apply_bytes), gives
    matvecs per second.  The table build on the CPU is not counted."""
    import numpy as np
    i0 = pt.fac.info(0)
    nc = max(i0["ncomp_a"], i0["ncomp_b"])
    tab = np.random.default_rng(0).random((nc, 2 * nc))
    x = np.random.default_rng(1).random(nc)
    x @ tab
    reps, t = 0, time.perf_counter()
    while time.perf_counter() - t < 2.0:
        x @ tab
        reps += 1
    dt = (time.perf_counter() - t) / reps
    gbs = tab.nbytes / dt / 1e9
    apply_bytes = info["apply_bytes"]
    return dict({"value": gbs * 1e9 / apply_bytes, "unit": "matvecs/s", "cores": host_info()["affinity_cpus"],
                 "kind": "port",
                 "sample": f"numpy (OpenBLAS) apply of one {nc}x{nc} {{S,F}} float64 table ({tab.nbytes / 1e6:.0f} MB) "
                           f"at {gbs:.1f} GB/s, scaled to the {apply_bytes / 1e9:.2f} GB of FP64 tables one apply reads"},
                **host_info())


def batch_workload(args):
    """Uneven multi-system batch: eight seeded systems with n_bas = 5n and
    n_occ = n/2 per spin for n = 4, 6, 8, 10 (n_bas 20-50), 4-16 references
    and 64-1024 singles/doubles per reference (58,062,543 upper-triangle
    pairs).  Every system is built in FP64 and in the mixed precision; the
    pairs of all systems run interleaved in one tuned plan of persistent
    fused launches captured into a HIP graph.  One step = the builds of every
    system (per-step device time) + one replay of the batch graph."""
    import numpy as np
    import torch

    import nwick
    from nwick import basis, synth

    tol = 1e-8
    t = time.perf_counter()
    systems = synth.batch_systems(seed=3030)
    log(f"batch setup: {len(systems)} systems in {time.perf_counter() - t:.1f} s")
    dev = torch.device("cuda:0")

    def build_all(prec):
        """Every system's context builds on its own stream; the builds of one
        step are issued from one host thread per system and overlap on the
        GPU.  build_ms = wall time of a step's builds (all systems)."""
        from concurrent.futures import ThreadPoolExecutor
        engs, masks = [], []
        for sysm, dets in systems:
            eng = nwick.Engine.from_system(sysm, device=0)
            eng.set_precision(prec)
            masks.append(np.triu(np.ones((sysm.nref, sysm.nref), dtype=np.uint8)))
            engs.append(eng)
        serial = 0.0
        tstat["transform_ms"] = tstat["flops"] = 0.0
        for eng, mask, (sysm, _) in zip(engs, masks, systems):
            for w in range(args.warmup + 1):
                meta = eng.build(tol=tol, pair_mask=mask)
            tm = eng.timings()
            serial += tm["total_ms"]
            # FP64 transforms only (the mixed build keeps a varying share in FP64)
            tstat["transform_ms"] += tm["transform_ms"]
            tstat["flops"] += sum(tensors_of(pm.ma, pm.mb) for pm in meta.values()) * 8.0 * sysm.n_bas ** 5
        pool = ThreadPoolExecutor(max_workers=len(engs))
        walls = []
        for k in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            list(pool.map(lambda em: em[0].build(tol=tol, pair_mask=em[1]), zip(engs, masks)))
            torch.cuda.synchronize()
            if k >= args.warmup:
                walls.append(1e3 * (time.perf_counter() - t0))
        pool.shutdown()
        log(f"builds: {np.mean(walls):.1f} ms per step concurrent (one system after another: {serial:.1f} ms device)")
        return engs, float(np.mean(walls)), serial

    def energies(b):
        out = []
        for k, (sysm, dets) in enumerate(systems):
            n = len(dets)
            H = torch.empty((n, n), dtype=torch.float64, device=dev)
            S = torch.empty((n, n), dtype=torch.float64, device=dev)
            b.result(k, out=(H, S))
            w, U = torch.linalg.eigh(S)
            keep = w > 1e-8 * w[-1]
            X = U[:, keep] / torch.sqrt(w[keep])
            A = X.T @ H @ X
            out.append(float(torch.linalg.eigvalsh(0.5 * (A + A.T))[0]))
            del H, S, U, X, A
        return out

    res = {}
    tstat = {}
    tstats = {}
    for prec, name in ((nwick.PREC_FP64, "fp64"), (nwick.PREC_MIXED, "mixed")):
        engs, build_ms, build_serial = build_all(prec)
        tstats[name] = dict(tstat)
        b = nwick.Batch()
        for eng, (sysm, dets) in zip(engs, systems):
            b.add(eng, basis.to_records(dets), mode=args.mode)
        t0 = time.perf_counter()
        b.plan(tune=True, graph=True)
        tune_s = time.perf_counter() - t0
        sched = b.schedule()
        for w in range(args.warmup):
            b.run()
        ms = []
        for k in range(args.steps):
            b.run()
            ms.append(b.last()[0])
            log(f"[{name}] batch step {k}: {ms[-1]:.2f} ms (builds {build_ms:.1f} ms)")
        npairs = b.last()[1]
        # the same plan untuned (16x16 tiles, register-class grouping, dynamic, 4 workgroups/CU)
        b2 = nwick.Batch()
        for eng, (sysm, dets) in zip(engs, systems):
            b2.add(eng, basis.to_records(dets), mode=args.mode)
        b2.plan(tune=False, graph=True)
        b2.run()
        t_untuned = []
        for _ in range(max(args.steps, 1)):
            b2.run()
            t_untuned.append(b2.last()[0])
        ms_untuned = min(t_untuned)
        b2.close()
        e = energies(b)
        res[name] = {"build_ms": build_ms, "build_ms_serial_device": build_serial,
                     "eval_ms": float(np.mean(ms)), "eval_ms_untuned": ms_untuned,
                     "npairs": npairs, "energies": e, "schedule": sched, "tune_s": tune_s}
        log(f"[{name}] eval {np.mean(ms):.2f} ms (untuned {ms_untuned:.2f} ms), builds {build_ms:.1f} ms, "
            f"tuning {tune_s:.1f} s, grouping {sched['grouping']} dyn {sched['dynamic']} grid/CU {sched['grid_per_cu']}")
        b.close()
        for eng in engs:
            eng.close()
        del engs
        torch.cuda.empty_cache()
    r64, rmx = res["fp64"], res["mixed"]
    dE = [a - c for a, c in zip(rmx["energies"], r64["energies"])]
    use_mixed = max(abs(x) for x in dE) < 1e-6
    r = rmx if use_mixed else r64
    npairs = r["npairs"]
    step_ms = r["build_ms"] + r["eval_ms"]
    # algorithmic bytes of the evaluation (m = 0 formula per contraction-rank bin)
    total_bytes = 0.0
    for sysm, dets in systems:
        ra = np.array([d.rank_a for d in dets]); rb = np.array([d.rank_b for d in dets])
        par = np.array([d.parent for d in dets])
        for x in range(sysm.nref):
            for y in range(x, sysm.nref):
                ia, ib = ra[par == x], rb[par == x]
                ja, jb = ra[par == y], rb[par == y]
                la = (ia[:, None] + ja[None, :]).ravel(); lb = (ib[:, None] + jb[None, :]).ravel()
                c2 = lambda l: l * (l - 1) // 2
                B = 8.0 * (2 * (la * la + lb * lb) + 2 * c2(la) ** 2 + 2 * c2(lb) ** 2 + la * la * lb * lb) + 60.0
                total_bytes += B.sum() * (0.5 if x == y else 1.0)
    achieved = total_bytes / (r["eval_ms"] * 1e-3) / 1e9
    t64_tflops = tstats["fp64"]["flops"] / (tstats["fp64"]["transform_ms"] * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        sysm, dets = systems[6]
        log("cpu baseline: oracle on system 6 (n_bas 50), pair (0,1) ...")
        cpu = cpu_baseline(sysm, dets, tol, args.mode)
        cpu["sample"] = "n_bas=50 system: " + cpu["sample"]
    out = {"metric": "Wick H,S pair elements/s over an uneven multi-system batch (end-to-end: builds + one "
                     "batched evaluation of every system)",
           "value": npairs / (step_ms * 1e-3), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64" if use_mixed else "f64", "data": "synthetic",
           "config": {"workload": "uneven batch: 8 systems, n_bas=5n, nocc=n/2 (n=4,6,8,10), 4-16 refs, "
                                  "64-1024 singles/doubles per ref, one tuned graph-captured plan",
                      "systems": [[s.n_bas, s.nocc_a, s.nref, len(d)] for s, d in systems], "pairs": npairs,
                      "precision": "mixed" if use_mixed else "fp64"},
           "eval_only_pairs_per_s": npairs / (r["eval_ms"] * 1e-3),
           "eval_only_pairs_per_s_untuned": npairs / (r["eval_ms_untuned"] * 1e-3),
           "fp64": {k: r64[k] for k in ("build_ms", "build_ms_serial_device", "eval_ms", "eval_ms_untuned",
                                        "energies", "tune_s")},
           "mixed": {k: rmx[k] for k in ("build_ms", "build_ms_serial_device", "eval_ms", "eval_ms_untuned",
                                         "energies", "tune_s")},
           "dE_mixed_minus_fp64": dE, "max_abs_dE": max(abs(x) for x in dE),
           "schedule": {k: r["schedule"][k] for k in ("grouping", "dynamic", "grid_per_cu")},
           "tuned_buckets": [[e["nocc"], e["la"], e["lb"], e["m0"], e["shape"], round(e["ms_default"] / max(e["ms_best"], 1e-9), 3)]
                             for e in r["schedule"]["buckets"]],
           "roofline": {"bound": "mfma", "kernel": "k_qtransform_mfma_pipe (FP64 four-index transforms of the "
                                                 "builds, the step's largest device time)",
                        "achieved": t64_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": t64_tflops / PEAK_FP64_TFLOPS, "traffic": None,
                        "flops_per_step": tstats["fp64"]["flops"], "transform_ms_per_step": tstats["fp64"]["transform_ms"],
                        "counters": "profiles/qtransform_mfma_counters.txt"},
           "roofline_eval": {"bound": "hbm", "kernel": "k_eval_fused (batched determinant-pair evaluation)",
                             "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                             "traffic": None, "algorithmic_bytes": total_bytes},
           "cpu_baseline": cpu}
    return out


def screen_workload(args):
    """Selected NOCI with the learned screener: 16 references, n_bas = 60,
    8 + 8 electrons, 1,000,000 seeded singles/doubles (62,500 per reference).
    Five selection iterations add 64 candidates each by the exact score
    |<D|H - E S|Psi>|^2 / |E - H_DD|.  Exhaustive: every candidate scored exactly
    each iteration.  Learned: iterations 0-1 exhaustive and used as training
    data for the 24-128-128-1 network (GPU forward/backward/Adam kernels);
    iterations 2-4 score exactly only the network's ranking until a chunk
    holds no candidate near the 64th best, plus a 1% random audit, with an
    exhaustive fallback on an audit miss."""
    import numpy as np

    import nwick
    from nwick import screen, synth

    t = time.perf_counter()
    sysm = synth.make_system(60, 8, 8, 16, seed=args.seed)
    cand = screen.random_candidates(16, 60, 8, 8, 62500, seed=args.seed)
    log(f"screen setup: system + {len(cand)} candidates in {time.perf_counter() - t:.1f} s")
    eng = nwick.Engine.from_system(sysm, device=0)
    eng.build()
    log(f"build of all 256 ordered reference pairs: {eng.timings()['total_ms']:.1f} ms")
    kw = dict(iters=5, k_add=64, seed=args.seed, mode=args.mode, log=log)
    ex = screen.select(eng, sysm, cand, learned=False, **kw)
    ln = screen.select(eng, sysm, cand, learned=True, learn_iters=2, top_frac=0.05, audit_frac=0.01, epochs=8,
                       batch=4096, lr=2e-3, measure_recall=True, **kw)
    n = len(cand)
    t_ex = float(np.mean(ex.screen_s[2:]))
    t_ln = float(np.mean(ln.screen_s[2:]))
    out = {"metric": "selected-NOCI candidate screening: candidates ranked per second (learned screener, "
                     "exact-verified)", "value": n / t_ln, "unit": "candidates/s", "n_gpus": 1,
           "steps": 3, "warmup": 2, "ms_per_step": 1e3 * t_ln, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (exact scores) + bf16/f32 (network)", "data": "synthetic",
           "config": {"workload": "SNOCI screening: 16 refs, n_bas=60, nocc=8+8, 1M seeded singles/doubles, "
                                  "5 iterations x 64 additions", "candidates": n},
           "exhaustive_candidates_per_s": n / t_ex,
           "energy_exhaustive": ex.energies[-1], "energy_learned": ln.energies[-1],
           "dE_learned_minus_exhaustive": ln.energies[-1] - ex.energies[-1],
           "exact_evaluations_exhaustive": ex.exact_evals, "exact_evaluations_learned": ln.exact_evals,
           "screen_s_exhaustive": ex.screen_s, "screen_s_learned": ln.screen_s,
           "recall_top64": ln.recall, "captured_score_fraction": ln.captured, "audit_misses": ln.audit_misses,
           "fallbacks": ln.fallbacks, "training": ln.train}
    # algorithmic bytes of the exact scores of the screened iterations: every
    # exactly scored candidate against every current determinant (references
    # and the candidates selected so far) plus its diagonal element
    ra, rb = cand.ranks()
    cur_a, cur_b = [np.zeros(16, int)], [np.zeros(16, int)]
    for it, pick in enumerate(ln.selected):
        cur_a.append(ra[np.asarray(pick)]); cur_b.append(rb[np.asarray(pick)])
    mean_b = []
    for it in range(2, len(ln.selected)):
        ca, cb = np.concatenate(cur_a[:it + 1]), np.concatenate(cur_b[:it + 1])
        sub = np.random.default_rng(it).choice(n, size=4096, replace=False)
        B = eval_bytes(ra[sub][:, None] + ca[None, :], rb[sub][:, None] + cb[None, :]).sum(1) + \
            eval_bytes(2 * ra[sub], 2 * rb[sub])
        mean_b.append(float(B.mean()) * ln.exact_evals[it])
    alg = float(np.mean(mean_b))
    out["roofline"] = {"bound": "hbm", "kernel": "k_eval_blocks<PROJECT> + diagonal k_eval_reqs (exact scores; "
                                                 "the screened iteration's time includes ranking and picking)",
                       "achieved": alg / t_ln / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": alg / t_ln / 1e9 / PEAK_HBM_GBS, "traffic": None, "algorithmic_bytes": alg,
                       "bytes_meaning": "per exactly scored candidate: the element bytes against each current "
                                        "determinant and of its diagonal (m = 0 formula; sampled over 4096 "
                                        "candidates per iteration)"}
    out["cpu_baseline"] = None if args.no_cpu_baseline else screen_cpu_baseline(sysm, cand, args.mode)
    eng.close()
    return out


def screen_cpu_baseline(sysm, cand, mode):
    """Exact candidate scores on the CPU oracle for a bounded sample: the
    first 20,000 candidates of reference 0 against the 16 references (the
    first selection iteration's current space) plus their diagonal elements,
    after the oracle builds the 16 ordered reference pairs (0, r) (counted)."""
    import ctypes as ct

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    L = O.lib()
    L.ow_hs_many.argtypes = [O.D, O.I64, O.D, ct.c_int, ct.c_longlong, O.U64, O.U64, O.D, ct.c_double,
                             ct.c_double, ct.c_int, O.D, O.D]
    L.ow_num_threads.restype = ct.c_int
    blas = O.use_openblas(True)
    t0 = time.perf_counter()
    slab = O.Slab(sysm, tol=1e-8, pairs=[(0, r) for r in range(sysm.nref)])
    t1 = time.perf_counter()
    idx = np.nonzero(cand.parent == 0)[0][:20000]
    m = len(idx)

    def masks(h):
        """(lo, hi) 64-bit words of the orbital sets (n_bas = 60: lo only)."""
        out = np.zeros((m, 2), dtype=np.uint64)
        for k in range(h.shape[1]):
            ok = h[idx, k] >= 0
            out[ok, 0] |= np.uint64(1) << h[idx, k][ok].astype(np.uint64)
        return out
    xex = np.ascontiguousarray(np.concatenate([masks(cand.ah), masks(cand.ap), masks(cand.bh), masks(cand.bp)], 1))
    ph = np.ones(m)
    wex0 = np.zeros((m, 8), dtype=np.uint64)
    H = np.zeros(m); S = np.zeros(m)
    t2 = time.perf_counter()
    for r in range(sysm.nref):
        sub, off, meta = slab.pair_view(0, r)
        L.ow_hs_many(O._d(sub), off.ctypes.data_as(O.I64), O._d(meta), sysm.nmo, m, O._u(xex), O._u(wex0),
                     O._d(ph), sysm.enuc, 1e-8, mode, O._d(H), O._d(S))
    sub, off, meta = slab.pair_view(0, 0)
    L.ow_hs_many(O._d(sub), off.ctypes.data_as(O.I64), O._d(meta), sysm.nmo, m, O._u(xex), O._u(xex),
                 O._d(ph), sysm.enuc, 1e-8, mode, O._d(H), O._d(S))
    t3 = time.perf_counter()
    threads = L.ow_num_threads()
    return dict({"value": m / ((t1 - t0) + (t3 - t2)), "unit": "candidates/s", "cores": threads, "kind": "port",
                 "sample": f"oracle build of the 16 ordered pairs (0, r) ({t1 - t0:.1f} s; transforms on {blas} "
                           f"dgemm) + exact elements of {m} candidates of reference 0 against the 16 references "
                           f"and their diagonals ({t3 - t2:.2f} s), OpenMP threads={threads}"}, **host_info())


def spawn_workload(args):
    """QMC spawning elements (find_hs_batched, src/stochastic/common.rs:88-112,
    over calculate_hs_pairs_wicks_batched, src/noci/hs.rs:72-130): H and S of
    uniformly random (bra, ket) determinant requests against a resident slab.
    System: n_bas = 60, 8 + 8 electrons, 8 references (all 64 ordered
    reference pairs built), 512 singles/doubles per reference.  One step is one
    nw_hs_batched call of 2^21 requests with host arrays in and out, as the
    caller passes them; the engine bins requests by contraction rank and
    reference pair, uploads them and evaluates every bin on the GPU."""
    import numpy as np

    import nwick
    from nwick import basis, synth

    tol = 1e-8
    sysm = synth.make_system(60, 8, 8, 8, seed=args.seed)
    dets = basis.excited_basis(8, 60, 8, 8, per_ref=512, seed=args.seed, include_refs=True)
    rec = basis.to_records(dets)
    eng = nwick.Engine.from_system(sysm, device=0)
    eng.build(tol=tol)
    log(f"spawn setup: {len(dets)} determinants, 64 ordered reference pairs built in "
        f"{eng.timings()['total_ms']:.1f} ms")
    nreq = 1 << 21
    rng = np.random.default_rng(args.seed)
    req = rng.integers(0, len(dets), size=(nreq, 2)).astype(np.uint64)
    for _ in range(args.warmup):
        eng.hs_batched(rec, req, mode=args.mode)
    calls, evals = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        h, s = eng.hs_batched(rec, req, mode=args.mode)
        calls.append(time.perf_counter() - t0)
        evals.append(eng.timings()["eval_ms"])
        log(f"spawn step {k}: call {1e3 * calls[-1]:.2f} ms, GPU evaluation {evals[-1]:.2f} ms")
    t_call = float(np.mean(calls))
    t_eval = float(np.mean(evals)) * 1e-3
    # the same requests with every array already in device memory (the rate
    # reported as value; the host-array call above adds the PCIe copies)
    import torch
    rec_d = torch.from_numpy(rec.view(np.uint8).copy()).to("cuda:0")
    req_d = torch.from_numpy(req.view(np.int64).copy()).to("cuda:0")
    h_d = torch.empty(nreq, dtype=torch.float64, device="cuda:0")
    s_d = torch.empty_like(h_d)
    eng.hs_batched_device(rec_d, req_d, h_d, s_d, mode=args.mode)
    torch.cuda.synchronize()
    dcalls = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        eng.hs_batched_device(rec_d, req_d, h_d, s_d, mode=args.mode)
        dcalls.append(time.perf_counter() - t0)
    t_dev = float(np.mean(dcalls))
    assert np.array_equal(h_d.cpu().numpy(), h) and np.array_equal(s_d.cpu().numpy(), s)
    log(f"spawn device-array call {1e3 * t_dev:.2f} ms")
    # algorithmic bytes of the requests (evaluation kernels k_eval_reqs)
    ra = np.array([d.rank_a for d in dets]); rb = np.array([d.rank_b for d in dets])
    par = np.array([d.parent for d in dets])
    oa = np.array([d.oa for d in dets], dtype=object); ob = np.array([d.ob for d in dets], dtype=object)
    a, b = req[:, 0].astype(np.int64), req[:, 1].astype(np.int64)
    same = par[a] == par[b]
    zero = np.zeros(nreq, dtype=bool)
    si = np.nonzero(same)[0]
    diff = np.array([bin(oa[a[q]] ^ oa[b[q]]).count("1") + bin(ob[a[q]] ^ ob[b[q]]).count("1") for q in si])
    zero[si[diff > 4]] = True
    B = np.where(zero, 80.0, eval_bytes(ra[a] + ra[b], rb[a] + rb[b]) + 16.0)
    alg = float(B.sum())
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sysm, dets, tol, args.mode, with_build=False)
        log(f"cpu baseline: {cpu['value']:.1f} pairs/s ({cpu['sample']})")
    out = {"metric": "QMC spawning elements: H,S of random determinant-pair requests per second (resident slab)",
           "value": nreq / t_dev, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * t_dev, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "random H,S requests: n_bas=60, nocc=8+8, 8 refs (64 ordered pairs built), "
                                  "512 singles/doubles per ref, 2^21 requests per call", "determinants": len(dets),
                      "requests": nreq},
           "gpu_evaluation_pairs_per_s": nreq / t_eval, "gpu_evaluation_ms": 1e3 * t_eval,
           "host_array_call_pairs_per_s": nreq / t_call, "host_array_call_ms": 1e3 * t_call,
           "nonzero_fraction": float(np.mean(s != 0.0)), "exact_zero_fraction": float(zero.mean()),
           "roofline": {"bound": "hbm", "kernel": "k_eval_reqs (per contraction-rank bin)",
                        "achieved": alg / t_eval / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": alg / t_eval / 1e9 / PEAK_HBM_GBS, "traffic": None, "algorithmic_bytes": alg,
                        "bytes_meaning": "per request: the element bytes of its contraction-rank bin plus the "
                                         "16-byte request; 80 bytes for exact zeros"},
           "cpu_baseline": cpu}
    eng.close()
    return out


def copy_workload(args):
    """Calibration for HBM traffic counters: a plain device-to-device copy of
    256 MiB (torch's elementwise copy kernel), repeated steps times.  The
    per-dispatch FETCH_SIZE / WRITE_SIZE of this kernel, compared with the
    268,435,456 bytes it moves each way, scale the same counters read for the
    engine's kernels."""
    import torch
    n = (256 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out = {"metric": "device copy bandwidth (counter calibration)", "value": 2 * n * 8 * args.steps / el / 1e9,
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": 1e3 * el / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "copy of 256 MiB", "bytes_each_way": n * 8}}
    return out


def wick_workload(args):
    """The headline: batched Wick elements at n_bas = 100 (module docstring),
    FP64 then the mixed forms; returns the result dict."""
    import numpy as np
    import torch

    import nwick
    from nwick import basis, synth

    tol = 1e-8
    t = time.perf_counter()
    sysm = synth.wick_batch_n100(seed=args.seed)
    dets = basis.excited_basis(sysm.nref, sysm.nmo, sysm.nocc_a, sysm.nocc_b, per_ref=512, seed=args.seed)
    rec = basis.to_records(dets)
    ndet = len(dets)
    npairs = ndet * (ndet + 1) // 2
    log(f"setup: system + {ndet} determinants in {time.perf_counter() - t:.1f} s")
    eng = nwick.Engine.from_system(sysm, device=0)
    mask = np.triu(np.ones((sysm.nref, sysm.nref), dtype=np.uint8))
    dev = torch.device("cuda:0")
    Hd = torch.empty((ndet, ndet), dtype=torch.float64, device=dev)
    Sd = torch.empty((ndet, ndet), dtype=torch.float64, device=dev)
    rec_d = torch.from_numpy(rec.view(np.uint8).copy()).to(dev)
    torch.cuda.synchronize()

    def step():
        meta = eng.build(tol=tol, pair_mask=mask)
        tb = eng.timings()
        eng.hs_matrix(rec_d, mode=args.mode, out=(Hd, Sd))
        te = eng.timings()
        return meta, tb, te

    def lowest_energy():
        L = torch.linalg.cholesky(Sd)
        A = torch.linalg.solve_triangular(L, Hd, upper=False)
        A = torch.linalg.solve_triangular(L, A.T.contiguous(), upper=False)
        return float(torch.linalg.eigvalsh(0.5 * (A + A.T))[0])

    def phase(prec, factored=False):
        """Steps of one form.  Dense forms: the build transforms every rank-4
        tensor (transforms = its transform time).  Factored forms (integral
        factors installed): the build leaves the rank-4 blocks to hs_matrix,
        which builds the blocks the determinants read (transforms = that
        time; stages from nw_rank4_stats)."""
        eng.set_precision(prec)
        name = ("factored " if factored else "") + PREC_NAME[prec]
        stages = []
        for w in range(args.warmup):
            meta, tb, te = step()
            log(f"[{name}] warmup {w}: build {tb['total_ms']:.1f} ms "
                f"(transforms {tb['transform_ms']:.1f} ms), eval {te['eval_ms']:.2f} ms")
        torch.cuda.synchronize()
        builds, transforms, evals = [], [], []
        t0 = time.perf_counter()
        for k in range(args.steps):
            meta, tb, te = step()
            r4 = te["transform_ms"] if factored else tb["transform_ms"]
            builds.append(tb["total_ms"] + (r4 if factored else 0.0)); transforms.append(r4); evals.append(te["eval_ms"])
            if factored:
                stages.append({**eng.rank4_stats(), **eng.rank4_kernel_stats()})
            log(f"[{name}] step {k}: build {tb['total_ms']:.1f} ms, rank-4 {r4:.1f} ms, eval {te['eval_ms']:.2f} ms")
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        energy = float("nan") if args.no_energy else lowest_energy()
        log(f"[{name}] {npairs * args.steps / elapsed:.4e} pairs/s, lowest NOCI energy {energy:.12f}")
        st = {k: float(np.mean([x[k] for x in stages])) for k in stages[0]} if stages else None
        return {"elapsed": elapsed, "meta": meta, "builds": builds, "transforms": transforms, "evals": evals,
                "energy": energy, "rate": npairs * args.steps / elapsed, "stages": st}

    PREC_NAME = {nwick.PREC_FP64: "fp64", nwick.PREC_MIXED: "mixed", nwick.PREC_BF16X3: "bf16x3"}
    res64 = phase(nwick.PREC_FP64)
    mixed_runs = {}
    if args.precision == "mixed":
        for prec in (nwick.PREC_MIXED, nwick.PREC_BF16X3):
            r = phase(prec)
            r["dE"] = None if args.no_energy else r["energy"] - res64["energy"]
            mixed_runs[prec] = r
    # the same steps with the workload's integral factors (n_aux = 300): rank-4
    # blocks built on demand, FP64 and FP32 products
    dev_f = eng.set_eri_factors(sysm.chol)
    resf64 = phase(nwick.PREC_FP64, factored=True)
    resf64["dE"] = None if args.no_energy else resf64["energy"] - res64["energy"]
    resfmx = resfbf = None
    if args.precision == "mixed":
        # FP32 pairs' block products on FP32 MFMA (mixed) and on BF16 MFMA with split operands (bf16x3)
        resfmx = phase(nwick.PREC_MIXED, factored=True)
        resfmx["dE"] = None if args.no_energy else resfmx["energy"] - res64["energy"]
        mixed_runs["factored"] = resfmx
        resfbf = phase(nwick.PREC_BF16X3, factored=True)
        resfbf["dE"] = None if args.no_energy else resfbf["energy"] - res64["energy"]
        mixed_runs["factored_bf16x3"] = resfbf
    eng.set_eri_factors(None)
    # report the fastest mixed form whose energy is within 1 uHa of the dense FP64 one
    ok = [p for p, r in mixed_runs.items() if args.no_energy or abs(r["dE"]) < 1e-6]
    best = max(ok, key=lambda p: mixed_runs[p]["rate"]) if ok else None
    resmx = mixed_runs.get(nwick.PREC_MIXED)
    dE = None if resmx is None else resmx["dE"]
    use_mixed = best is not None
    res = mixed_runs[best] if use_mixed else res64
    meta = res["meta"]
    ms_per_step = 1e3 * res["elapsed"] / args.steps
    value = res["rate"]
    transforms, evals, builds = res["transforms"], res["evals"], res["builds"]
    # roofline of the dense four-index transforms: the fastest dense mixed
    # form within 1 uHa (or the FP64 transforms)
    dok = [p for p in ok if p not in ("factored", "factored_bf16x3")]
    dbest = max(dok, key=lambda p: mixed_runs[p]["rate"]) if dok else None
    dense_mixed = dbest is not None
    dres = mixed_runs[dbest] if dense_mixed else res64
    # dominant kernel: the rotating quarter transform, 2 * n_bas * rows * nmo = 2 n^5 flop per tensor pass;
    # the build launches up to g tensors per dispatch (g tensors of scratch fit in 1 GiB)
    n = sysm.n_bas
    slots = lambda m: sum(1 for b in [(0, 0, 0, 0), (0, 0, 0, 1), (0, 0, 1, 0), (0, 0, 1, 1), (0, 1, 0, 1),
                                      (0, 1, 1, 0), (0, 1, 1, 1), (1, 0, 1, 0), (1, 0, 1, 1), (1, 1, 1, 1)]
                          if sum(b) <= m)
    ntensor = n32 = 0
    for (x, w), pm in res64["meta"].items():
        nii = sum(1 for a0 in (0, 1) for aj in (0, 1) for b0 in (0, 1) for bj in (0, 1)
                  if a0 + aj <= pm.ma and b0 + bj <= pm.mb)
        k = slots(pm.ma) + slots(pm.mb) + nii
        ntensor += k
        if dense_mixed and dres["meta"][(x, w)].fp32:
            n32 += k
    flops_tensor = 4 * 2.0 * n ** 5
    t64 = float(np.mean(res64["transforms"])) / ntensor  # ms per FP64 tensor
    group = lambda elem: max(1, min(32, (1 << 30) // (n ** 4 * elem)))
    if dense_mixed:
        t_k = float(np.mean(dres["transforms"])) - (ntensor - n32) * t64
        nk, g = n32, group(4)
    else:
        t_k = float(np.mean(res64["transforms"]))
        nk, g = ntensor, group(8)
    dispatches = 4 * ((nk + g - 1) // g)
    achieved = nk * flops_tensor / (t_k * 1e-3) / 1e12
    avg_ms = t_k / dispatches
    eval_rate = npairs / (float(np.mean(evals)) * 1e-3)
    if dense_mixed and dbest == nwick.PREC_BF16X3:
        # six BF16 products per FP32-equivalent product: the emulated peak is dense BF16 MFMA / 6
        kname, peak = "k_qtransform_bf3", PEAK_BF16_TFLOPS / 6.0
    elif dense_mixed:
        kname, peak = "k_qtransform_f32_pipe", PEAK_FP32_TFLOPS
    else:
        kname, peak = "k_qtransform_mfma_pipe", PEAK_FP64_TFLOPS
    traffic = None
    pname = ("qtransform_bf3" if dbest == nwick.PREC_BF16X3 else "qtransform_f32") if dense_mixed else "qtransform"
    tfile = os.path.join(ROOT, "profiles", f"{pname}_n100_hbm_traffic.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("TRAFFIC_BYTES"):
                traffic = float(line.split()[1])
    # algorithmic bytes per dispatch: FP64 passes read and write n^4 doubles; a
    # reduced-precision pass reads and writes n^4 floats (the last pass too:
    # its rank-4 slab blocks are stored in FP32), g tensors per dispatch
    alg_bytes = g * 8.0 * n ** 4 if dense_mixed else 16.0 * n ** 4
    flops_disp = g * flops_tensor / 4
    ridge = peak / (PEAK_HBM_GBS / 1e3)  # flop per byte where the MFMA and HBM roofs meet
    hbm_bound = flops_disp / alg_bytes < ridge
    if hbm_bound:
        roof = {"bound": "hbm", "achieved": alg_bytes / (avg_ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": alg_bytes / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, "mfma_tflops": achieved,
                "mfma_peak_tflops": peak}
    else:
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "hbm_gbs": alg_bytes / (avg_ms * 1e-3) / 1e9}
    # evaluation roofline: algorithmic bytes of every upper-triangle element by
    # contraction-rank bin (runs of equal parent and ranks), over the measured
    # evaluation time of the step
    runs = []
    for d in dets:
        key = (d.parent, d.rank_a, d.rank_b)
        if runs and runs[-1][0] == key:
            runs[-1][1] += 1
        else:
            runs.append([key, 1])
    eval_alg = eval_alg_m0 = 0.0
    n_m0 = 0
    for i, (ka, na) in enumerate(runs):
        for j in range(i, len(runs)):
            kb, nb = runs[j]
            cnt = na * (na + 1) // 2 if i == j else na * nb
            bts = cnt * eval_bytes(ka[1] + kb[1], ka[2] + kb[2])
            eval_alg += bts
            pm = res64["meta"][(ka[0], kb[0])]
            if pm.ma == 0 and pm.mb == 0:
                eval_alg_m0 += bts
                n_m0 += cnt
    eval_s = float(np.mean(evals)) * 1e-3
    roof_eval = {"bound": "hbm", "kernel": "k_eval_blocks<HS, L_a, L_b, m0> (every contraction-rank bin)",
                 "achieved": eval_alg / eval_s / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                 "frac": eval_alg / eval_s / 1e9 / PEAK_HBM_GBS, "traffic": None,
                 "traffic_source": "profiles/eval_blocks_n100_hbm_traffic.txt",
                 "algorithmic_bytes": eval_alg, "algorithmic_bytes_m0_pairs": eval_alg_m0, "pairs_m0": n_m0,
                 "bytes_meaning": "per element: X/Y entries of both spins, hcol0, J and II gathers, H and S out, two "
                                  "packed determinants (m = 0 formula; a lower bound for the m = 1 pairs, which "
                                  "read the D1 blocks and more J slots as well)"}
    tfile_e = os.path.join(ROOT, "profiles", "eval_blocks_n100_hbm_traffic.txt")
    if os.path.exists(tfile_e):
        for line in open(tfile_e):
            if line.startswith("TRAFFIC_BYTES"):
                roof_eval["traffic"] = float(line.split()[1])
    roof_dense = dict(roof, **{"kernel": kname + " (rotating quarter transform)", "traffic": traffic,
                      "traffic_source": os.path.relpath(tfile, ROOT) + " (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                      "algorithmic_bytes_per_dispatch": alg_bytes,
                      "dispatches_per_step": dispatches, "tensors_per_dispatch": g, "avg_launch_ms": avg_ms,
                      "flops_per_dispatch": flops_disp})

    def stage_roofs(r, mixed, bf3=False):
        """MFMA rooflines of the on-demand rank-4 stages of a factored run:
        algorithmic flops (nw_rank4_stats) over the stage's device time.  A
        mixed run's products (factor and block) run on FP32 MFMA except those
        of pairs under the sigma floor (FP64), which the FP32 peak leaves in,
        so its fractions are lower bounds.  With bf3 the FP32 pairs' block
        products run on BF16 MFMA as six products of split operands: their
        peak is the BF16 dense peak / 6 in FP32-equivalent flops."""
        st = r["stages"]
        peak_b = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
        peak_k = PEAK_BF16_TFLOPS / 6.0 if bf3 else peak_b
        kb = "k_r4_bf3" if bf3 else "k_r4_gemm"
        out_s = {}
        for key, ms, fl, kern, peak in (
                ("factor", "ms_factor", "flops_factor", "k_fac_gemm (factor products X^T L_K Y)", peak_b),
                ("rect", "ms_rect", "flops_rect", kb + " (rectangle products)", peak_k),
                ("single", "ms_single", "flops_single", kb + " (single blocks)", peak_k),
                ("exchange", "ms_exchange", "flops_exchange", kb + " (exchange products)", peak_k)):
            if st[ms] <= 0.0:
                continue
            ach = st[fl] / (st[ms] * 1e-3) / 1e12
            out_s[key] = {"bound": "mfma", "kernel": kern, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                          "frac": ach / peak, "ms_per_step": st[ms], "flops_per_step": st[fl]}
        # the block products by kernel launch class (nw_rank4_kernel_stats): the
        # FP32 pairs' large tiles are one kernel (k_r4_bf3 or k_r4_gemm<float>)
        for key, kern, peak in (
                ("f32_large", ("k_r4_bf3<128,128> / <64,64>" if bf3 else "k_r4_gemm<float> large tiles")
                 + " (block products of the FP32 pairs)", peak_k),
                ("f32_narrow", "k_r4_gemm<float,128,32> (ket-particle pair products of the FP32 pairs)",
                 PEAK_FP32_TFLOPS),
                ("f64_large", "k_r4_gemm<double> large tiles (block products of the FP64 pairs)", PEAK_FP64_TFLOPS),
                ("f64_narrow", "k_r4_gemm<double,128,32> (ket-particle pair products of the FP64 pairs)",
                 PEAK_FP64_TFLOPS)):
            ms, fl = st["ms_" + key], st["flops_" + key]
            if ms <= 0.0:
                continue
            ach = fl / (ms * 1e-3) / 1e12
            out_s["kernel_" + key] = {"bound": "mfma", "kernel": kern, "achieved": ach, "peak": peak,
                                      "unit": "TFLOP/s", "frac": ach / peak, "ms_per_step": ms, "flops_per_step": fl}
        return out_s
    stages_f64 = stage_roofs(resf64, False)
    stages_fmx = stage_roofs(resfmx, True) if resfmx is not None else None
    stages_fbf = stage_roofs(resfbf, True, bf3=True) if resfbf is not None else None
    if best in ("factored", "factored_bf16x3"):
        # the dominant kernel: the FP32 pairs' large-tile block products
        st_best = stages_fmx if best == "factored" else stages_fbf
        kern_roofs = {k: v for k, v in st_best.items() if k.startswith("kernel_")}
        dom = max(kern_roofs.items(), key=lambda kv: kv[1]["ms_per_step"])
        headline_roof = dict(dom[1], **{"traffic": None, "stage": dom[0],
                                        "profile": "profiles/r4_factored_c2_kernel_stats.txt"})
        # HBM bytes per dispatch of k_r4_bf3 (copy-calibrated FETCH_SIZE + WRITE_SIZE, 3 dispatches per step)
        tfile_r4 = os.path.join(ROOT, "profiles", "r4_factored_c2_counters.txt")
        if best == "factored_bf16x3" and dom[0] == "kernel_f32_large" and os.path.exists(tfile_r4):
            for line in open(tfile_r4):
                if line.startswith("TRAFFIC_BYTES_PER_DISPATCH k_r4_bf3"):
                    headline_roof["traffic"] = float(line.split()[2])
                    headline_roof["traffic_source"] = ("profiles/r4_factored_c2_counters.txt (rocprofv3 FETCH_SIZE/"
                                                       "WRITE_SIZE, copy-calibrated; per dispatch, 3 per step)")
    else:
        headline_roof = roof_dense
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle on pair (0,1) ...")
        cpu = cpu_baseline(sysm, dets, tol, args.mode)
        log(f"cpu baseline: {cpu['value']:.1f} pairs/s ({cpu['sample']})")
    out = {
        "metric": "Wick H,S pair elements/s: mixed vs FP64 at |\u0394E|<1\u00b5Ha; % roofline",
        "metric_detail": "end to end: the per-reference-pair build of the 36 upper-triangle pairs + all "
                         "upper-triangle H,S elements; value = the fastest mixed form within 1 uHa of FP64",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ({nwick.PREC_BF16X3: "bf16x3+f64", nwick.PREC_MIXED: "f32+f64", "factored": "f32+f64",
                   "factored_bf16x3": "bf16x3+f32+f64"}[best]
                  if use_mixed else "f64"),
        "data": "synthetic",
        "config": {"workload": "batched Wick elements, n_bas=nmo=100, nocc=10+10, 8 refs (1 with an exactly "
                               "orthogonal alpha orbital), 512 singles/doubles per ref, 4096 dets, 8390656 pairs",
                   "mode": "reference" if args.mode == 0 else "corrected", "pairs_built": len(meta),
                   "precision": ({"factored": "factored mixed", "factored_bf16x3": "factored bf16x3"}.get(best) or PREC_NAME[best]) if use_mixed else "fp64",
                   "integrals": "dense ERIs; factored forms also take the workload's 300 Cholesky vectors",
                   "rank4_tensors": ntensor, "fp32_tensors": n32},
        "fp64_pairs_per_s": res64["rate"],
        "mixed_pairs_per_s": None if resmx is None else resmx["rate"],
        "mixed_over_fp64": None if resmx is None else resmx["rate"] / res64["rate"],
        "bf16x3_pairs_per_s": mixed_runs[nwick.PREC_BF16X3]["rate"] if nwick.PREC_BF16X3 in mixed_runs else None,
        "bf16x3_over_fp64": (mixed_runs[nwick.PREC_BF16X3]["rate"] / res64["rate"]
                             if nwick.PREC_BF16X3 in mixed_runs else None),
        "energy_fp64": res64["energy"],
        "energy_mixed": None if resmx is None else resmx["energy"],
        "dE_mixed_minus_fp64": dE,
        "dE_bf16x3_minus_fp64": mixed_runs[nwick.PREC_BF16X3]["dE"] if nwick.PREC_BF16X3 in mixed_runs else None,
        "value_over_fp64": value / res64["rate"],
        "factored_fp64_pairs_per_s": resf64["rate"],
        "factored_fp64_over_fp64": resf64["rate"] / res64["rate"],
        "dE_factored_fp64_minus_fp64": resf64["dE"],
        "factored_mixed_pairs_per_s": None if resfmx is None else resfmx["rate"],
        "factored_mixed_over_fp64": None if resfmx is None else resfmx["rate"] / res64["rate"],
        "factored_mixed_over_factored_fp64": None if resfmx is None else resfmx["rate"] / resf64["rate"],
        "dE_factored_mixed_minus_fp64": None if resfmx is None else resfmx["dE"],
        "factored_bf16x3_pairs_per_s": None if resfbf is None else resfbf["rate"],
        "factored_bf16x3_over_fp64": None if resfbf is None else resfbf["rate"] / res64["rate"],
        "dE_factored_bf16x3_minus_fp64": None if resfbf is None else resfbf["dE"],
        "factor_check_deviation": dev_f,
        "factored_rank4_ms": {"fp64": float(np.mean(resf64["transforms"])),
                              "mixed": None if resfmx is None else float(np.mean(resfmx["transforms"])),
                              "bf16x3": None if resfbf is None else float(np.mean(resfbf["transforms"]))},
        "eval_only_pairs_per_s": eval_rate,
        "build_ms": float(np.mean(builds)),
        "transform_ms": float(np.mean(transforms)),
        "eval_ms": float(np.mean(evals)),
        "roofline": headline_roof,
        "roofline_stages": {"factored_fp64": stages_f64, "factored_mixed": stages_fmx,
                            "factored_bf16x3": stages_fbf},
        "roofline_dense_transform": roof_dense,
        "roofline_eval": roof_eval,
        "cpu_baseline": cpu,
    }
    eng.close()
    del Hd, Sd
    return out




def stream_workload(args):
    """The headline recipe with 16 references (136 upper-triangle pairs, about
    340 GB of FP64 slab: more than one device holds) through the streamed
    build -> evaluate -> reuse schedule, FP64, bf16x3, and FP32 with the
    workload's integral factors (rank-4 blocks built on demand per group);
    returns the result dict."""
    import numpy as np
    import torch

    import nwick
    from nwick import basis, synth

    tol, nref = 1e-8, 16
    t = time.perf_counter()
    # reference 15 has an alpha orbital orthogonal to the occupied spaces of references 0-6
    sysm = synth.make_system(100, 10, 10, nref, args.seed, n_aux=300, orth_refs=(nref - 1,), orth_count=1,
                             orth_against=tuple(range(7)))
    dets = basis.excited_basis(nref, 100, 10, 10, per_ref=512, seed=args.seed)
    rec = basis.to_records(dets)
    ndet = len(dets)
    npairs = ndet * (ndet + 1) // 2
    log(f"stream setup: {nref} refs, {ndet} determinants in {time.perf_counter() - t:.1f} s")
    eng = nwick.Engine.from_system(sysm, device=0)
    dev = torch.device("cuda:0")
    Hd = torch.empty((ndet, ndet), dtype=torch.float64, device=dev)
    Sd = torch.empty((ndet, ndet), dtype=torch.float64, device=dev)
    rec_d = torch.from_numpy(rec.view(np.uint8).copy()).to(dev)
    torch.cuda.synchronize()

    def lowest_energy():
        L = torch.linalg.cholesky(Sd)
        A = torch.linalg.solve_triangular(L, Hd, upper=False)
        A = torch.linalg.solve_triangular(L, A.T.contiguous(), upper=False)
        return float(torch.linalg.eigvalsh(0.5 * (A + A.T))[0])

    runs = {}
    forms = (("fp64", nwick.PREC_FP64, False), ("bf16x3", nwick.PREC_BF16X3, False),
             ("factored_mixed", nwick.PREC_MIXED, True), ("factored_bf16x3", nwick.PREC_BF16X3, True))
    for name, prec, factored in forms:
        if factored:  # the workload's 300 Cholesky vectors: rank-4 blocks built on demand per group
            eng.set_eri_factors(sysm.chol)
        eng.set_precision(prec)
        meta = {}
        for w in range(args.warmup):
            _, _, groups = eng.hs_matrix_streamed(rec_d, tol=tol, mode=args.mode, out=(Hd, Sd), meta=meta)
        torch.cuda.synchronize()
        totals, builds, evals = [], [], []
        t0 = time.perf_counter()
        for k in range(args.steps):
            _, _, groups = eng.hs_matrix_streamed(rec_d, tol=tol, mode=args.mode, out=(Hd, Sd), meta=meta)
            tm = eng.timings()
            totals.append(tm["total_ms"]); builds.append(tm["transform_ms"]); evals.append(tm["eval_ms"])
            log(f"[stream {name}] step {k}: {groups} groups, {tm['total_ms']:.0f} ms (builds {tm['transform_ms']:.0f}"
                f" ms, evaluations {tm['eval_ms']:.1f} ms)")
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        stages = {**eng.rank4_stats(), **eng.rank4_kernel_stats()} if factored else None
        if factored:
            eng.set_eri_factors(None)
        energy = float("nan") if args.no_energy else lowest_energy()
        ntensor = n32 = 0
        for pm in meta.values():
            k = tensors_of(pm.ma, pm.mb)
            ntensor += k
            n32 += k if pm.fp32 else 0
        runs[name] = {"rate": npairs * args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
                      "groups": groups, "energy": energy, "build_ms": float(np.mean(builds)),
                      "eval_ms": float(np.mean(evals)), "pairs_built": len(meta), "rank4_tensors": ntensor,
                      "fp32_tensors": n32,
                      "slab_bytes_fp64_estimate": sum(8 * (58 * 100 ** 2 + tensors_of(pm.ma, pm.mb) * 100 ** 4)
                                                    for pm in meta.values()),
                      "rank4_stages_last_step": stages}
        log(f"[stream {name}] {runs[name]['rate']:.4e} pairs/s, {groups} groups, energy {energy:.12f}")
    eng.close()
    del Hd, Sd
    r64, rbf, rfac, rfbf = runs["fp64"], runs["bf16x3"], runs["factored_mixed"], runs["factored_bf16x3"]
    dE = None if args.no_energy else rbf["energy"] - r64["energy"]
    dE_fac = None if args.no_energy else rfac["energy"] - r64["energy"]
    dE_fbf = None if args.no_energy else rfbf["energy"] - r64["energy"]
    # the fastest form within 1 uHa of FP64
    ok = [n for n in ("bf16x3", "factored_mixed", "factored_bf16x3")
          if args.no_energy or abs(runs[n]["energy"] - r64["energy"]) < 1e-6]
    best = max(ok, key=lambda n: runs[n]["rate"]) if ok else "fp64"
    use_bf = best == "bf16x3"
    res = runs[best]
    # dominant kernel: the four-index transforms inside the builds; the build time
    # also holds the SVDs, co-densities and small GEMMs, so this is a lower bound
    n = 100
    rd = r64 if best.startswith("factored") else res  # the dense form measured
    bytes_tensor = 32.0 * n ** 4  # four passes, each reading and writing n^4 floats
    alg = rd["fp32_tensors"] * bytes_tensor + (rd["rank4_tensors"] - rd["fp32_tensors"]) * 2 * bytes_tensor
    achieved = alg / (rd["build_ms"] * 1e-3) / 1e9
    roof_dense = {"bound": "hbm", "kernel": "k_qtransform_bf3 / k_qtransform_mfma_pipe (whole builds: a lower "
                  "bound for the transform kernels)", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                  "frac": achieved / PEAK_HBM_GBS, "traffic": None, "algorithmic_bytes": alg,
                  "bytes_meaning": "32 n^4 bytes per reduced-precision rank-4 tensor (four passes over FP32), "
                                   "64 n^4 per FP64 tensor", "form": "bf16x3" if rd is rbf else "fp64"}
    roof = roof_dense
    st = runs[best]["rank4_stages_last_step"] if best.startswith("factored") else None
    if best == "factored_bf16x3" and st and st["ms_f32_large"] > 0.0:
        # the dominant kernel: the FP32 pairs' large-tile block products (k_r4_bf3), all groups of the last step
        ach = st["flops_f32_large"] / (st["ms_f32_large"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": "k_r4_bf3 (block products of the FP32 pairs, BF16 MFMA on split operands)",
                "achieved": ach, "peak": PEAK_BF16_TFLOPS / 6.0, "unit": "TFLOP/s",
                "frac": ach / (PEAK_BF16_TFLOPS / 6.0), "traffic": None, "ms_per_step": st["ms_f32_large"],
                "flops_per_step": st["flops_f32_large"],
                "note": "FP32-equivalent flops; peak = BF16 dense peak / 6 (six BF16 products per FP32 product)"}
    elif best == "factored_mixed" and st:
        # dominant stage of the on-demand rank-4 blocks (all groups of the last step)
        cand = [("factor", "ms_factor", "flops_factor", "k_fac_gemm (factor products X^T L_K Y)"),
                ("rect", "ms_rect", "flops_rect", "k_r4_gemm (rectangle products)"),
                ("single", "ms_single", "flops_single", "k_r4_gemm (single blocks)"),
                ("exchange", "ms_exchange", "flops_exchange", "k_r4_gemm (exchange products)")]
        key, ms, fl, kern = max(cand, key=lambda c: st[c[1]])
        ach = st[fl] / (st[ms] * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": kern, "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP32_TFLOPS, "traffic": None, "stage": key, "ms_per_step": st[ms],
                "flops_per_step": st[fl],
                "note": "FP32 peak; the products of FP64 pairs (sigma floor) are included, so a lower bound"}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sysm, dets, tol, args.mode)
    return {
        "metric": "Wick H,S pair elements/s with the slab streamed through device memory (build -> evaluate -> "
                  "reuse per group of reference pairs)",
        "value": res["rate"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3+f64", "factored_mixed": "f32+f64", "factored_bf16x3": "bf16x3+f32+f64",
                  "fp64": "f64"}[best], "data": "synthetic",
        "config": {"workload": "batched Wick elements, n_bas=nmo=100, nocc=10+10, 16 refs (1 with an alpha orbital "
                               "orthogonal to 7 others), 512 singles/doubles per ref, 8192 dets, "
                               f"{npairs} pairs, 136 upper-triangle reference pairs", "mode": args.mode,
                   "form": best},
        "fp64": r64, "bf16x3": rbf, "factored_mixed": rfac, "factored_bf16x3": rfbf, "dE_bf16x3_minus_fp64": dE,
        "dE_factored_mixed_minus_fp64": dE_fac, "dE_factored_bf16x3_minus_fp64": dE_fbf,
        "roofline": roof, "roofline_dense_build": roof_dense,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", type=int, default=0, help="0: reference arithmetic, 1: corrected m>0 path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="mixed: also run the mixed-precision build and report it when |dE| < 1 uHa")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--pt2-max-iter", type=int, default=300,
                    help="pt2 workload: GMRES iteration limit of each PT2 solve")
    ap.add_argument("--pt2-shift", type=float, default=2.0,
                    help="pt2 workload: imaginary shift of the shifted PT2 solves")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the NOCI energies (eigensolver) -- for profiler runs; then |dE| is not checked")
    ap.add_argument("--workload", choices=["all", "wick", "pt2", "batch", "screen", "spawn", "stream", "copy"],
                    default="all",
                    help="all (default): the headline wick line carrying compact runs of the other workloads; "
                         "wick: batched H,S elements (headline); pt2: NOCI-PT2 matvec; batch: uneven "
                         "multi-system batch; screen: learned SNOCI screening; spawn: QMC spawning requests; "
                         "stream: 16-reference slab streamed through device memory; copy: counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path in this build "
              "(sharding is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    args.compact = False
    one = {"wick": wick_workload, "pt2": pt2_workload, "batch": batch_workload, "screen": screen_workload,
           "spawn": spawn_workload, "stream": stream_workload, "copy": copy_workload}
    if args.workload != "all":
        print(json.dumps(one[args.workload](args)), flush=True)
        return
    import torch
    out = wick_workload(args)
    torch.cuda.empty_cache()
    # the other workloads, compact (steps and warm-ups as given, PT2 shifted solves only), each with its own
    # roofline and CPU baseline, inside the same driver-timed run
    args.compact = True
    out["workloads"] = {}
    for name in ("batch", "pt2", "screen", "spawn", "stream"):
        t0 = time.perf_counter()
        log(f"--- workload {name}")
        r = one[name](args)
        r["wall_s"] = time.perf_counter() - t0
        out["workloads"][name] = r
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
