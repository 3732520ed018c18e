"""Benchmark of the MRHS even-odd Wilson hopping term on one MI355X.

Workload (the headline line): SU(3), 24^3 x 48 lattice, nrhs = 16, FP32
storage and arithmetic, packed RHS-innermost fields resident in HBM; a step
is one application of the hopping term to all 16 right-hand sides, parities
alternating. value = sites*RHS per second = S * nrhs / t_step.

The JSON line also carries
  roofline:     algorithmic bytes per launch S*[k*8N*(b_in+b_out) + 16 N^2 b_U]
                over the kernel's average launch time (HIP events on the
                launch stream) against 8 TB/s; `traffic` is the rocprofv3
                FETCH_SIZE + WRITE_SIZE per launch (scaled by a copy-kernel
                calibration) when a measured summary is committed under profiles/,
                with the tile it was measured on (`traffic_tile`);
  cpu_baseline: the FP32 C oracle (OpenMP) on the host cores, same workload,
                a bounded number of applications;
  extra:        FP16 / BF16 / FP64 storage and two-row and eight-real SU(3)
                links at the same size, a cache-resident 16^3x32 case, the batched graph,
                one process's share of a split lattice (interior, face pack,
                boundary terms), the clover term, and the qcu.h end-to-end
                call;
  cg, clover_cg: mixed-precision vs FP64 CG time to a 1e-10 true residual
                at 32^3x64, nrhs = 16, Wilson and Wilson-clover.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
_GAUGE_CACHE = {}
RETUNE = False  # --retune: tune the timed shapes instead of using the pinned tiles


def bytes_alg(S, k, N, b_in, b_out, b_u):
    """Compulsory HBM bytes of one application (every field touched once)."""
    return S * (k * 8 * N * (b_in + b_out) + 16 * N * N * b_u)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG comparison")
    ap.add_argument("--cg-only", action="store_true", help="only the CG comparison (printed as one JSON line)")
    ap.add_argument("--cg-L", type=int, nargs=4, default=[32, 32, 32, 64])
    ap.add_argument("--cg-clover", type=float, default=None, help="with --cg-only: Wilson-clover with this c_sw")
    ap.add_argument("--cg-solver", choices=["both", "fp64", "mixed", "fp8"], default="both",
                    help="with --cg-only: run one solver only (profiling)")
    ap.add_argument("--profile", action="store_true", help="headline kernel only (for rocprofv3 runs)")
    ap.add_argument("--retune", action="store_true",
                    help="tune the timed shapes instead of using the pinned tiles shipped with the library")
    ap.add_argument("--tile", type=str, default=None,
                    help="fixed tile R,block,lds,tblock[,zblock,hints,occupancy,traversal] for the headline kernel "
                         "(skips the tuner)")
    ap.add_argument("--faces-only", action="store_true",
                    help="only the split-lattice face exchange and clover timings")
    ap.add_argument("--stream", type=float, nargs="*", default=None,
                    help="only the stream-bandwidth ceiling over buffers of these sizes in MiB")
    ap.add_argument("--link-block", action="store_true",
                    help="only the VALU vs MFMA link x half-spinor block benchmark (one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of the headline field size (counter calibration)")
    ap.add_argument("--precision", type=str, default="fp32", choices=["fp16", "bf16", "fp32", "fp64", "fp8"],
                    help="storage precision of the headline kernel (profiling runs)")
    ap.add_argument("--L", type=int, nargs=4, default=[24, 24, 24, 48])
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed launches of the headline operator before its warmup steps, in wall-clock ms "
                         "(the GPU's clocks settle after ~70 launches of it; 0 = none)")
    return ap.parse_args()


def time_dslash(lib, _lib, torch, geom, prec, k, steps, warmup, L, N, seed=0, tile=None, settle_ms=0.0):
    """Tune (unless a tile is given), then time `steps` applications;
    returns (ms_per_step, tile). settle_ms > 0: before the `warmup`
    launches, launch the same operator (untimed) for that much wall time, so
    that the timed launches see the GPU's steady clocks: from a cold start
    the first launch of the headline kernel takes ~2.4 ms (code load),
    launches 2-4 ~340 us, launches 8-14 ~390 us (the clocks dropping as the
    power draw rises) and the kernel settles at ~333 us only after ~70
    launches."""
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    from mrhs_qcu_amd import fields
    import numpy as np

    S = fields.half_volume(L)
    key = (tuple(L), N, seed)
    if key not in _GAUGE_CACHE:
        _GAUGE_CACHE.clear()
        _GAUGE_CACHE[key] = torch.view_as_real(fields.gauge_field_device(L, N, seed=seed, device=dev)).reshape(-1)
    gu = _GAUGE_CACHE[key]
    # FP8 spinors go with binary16 links (gauge bytes and packing as HALF).
    dt = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16, 5: torch.float16}[prec]
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, prec) // dt.itemsize, dtype=dt, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), prec, gu.data_ptr(), _lib.DOUBLE, geom, sp))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    # One fixed Gaussian input and one output: step i applies the operator of
    # parity i & 1 to the same input, so every timed launch sees the same
    # finite data (feeding outputs back in would grow the field ~4x per
    # application) and the fields touched are the two a ping-pong would touch.
    a = torch.randn(S * 4 * N * k * 2, generator=gen, device=dev, dtype=torch.float32)
    if prec == _lib.FP8:
        a8 = torch.empty(S * 4 * N * k * 2, dtype=torch.uint8, device=dev)
        _lib.check(lib.qcuMrhsConvert(a8.data_ptr(), _lib.FP8, a.data_ptr(), _lib.SINGLE, S * 4 * N * k, sp))
        a = a8
    else:
        a = a.to(dt)
    ins = [a, a]
    o = torch.empty_like(a)
    outs = [o, o]
    if tile is None:
        # The pinned tile of this shape (tiles_<arch>.txt beside the library),
        # else the default; --retune times every candidate here instead.
        best = _lib.Tile()
        if RETUNE:
            us = ctypes.c_double()
            _lib.check(lib.qcuMrhsTune(outs[0].data_ptr(), ins[0].data_ptr(), gn.data_ptr(), prec, k, geom,
                                       ctypes.byref(best), ctypes.byref(us), sp))
        else:
            _lib.check(lib.qcuMrhsGetTile(geom, prec, k, ctypes.byref(best)))
        tp = None
    else:
        best = _lib.tile(*tile)
        tp = ctypes.byref(best)
    args = [(outs[p].data_ptr(), ins[p].data_ptr(), gn.data_ptr(), prec, k, p, 0, geom, tp, sp) for p in (0, 1)]
    if settle_ms > 0:
        t0 = time.perf_counter()
        while (time.perf_counter() - t0) * 1e3 < settle_ms:
            for i in range(16):
                _lib.check(lib.qcuMrhsDslash(*args[i & 1]))
            torch.cuda.synchronize()
    for i in range(warmup):
        _lib.check(lib.qcuMrhsDslash(*args[i & 1]))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rcs = [0] * steps
    e0.record(st)
    for i in range(steps):
        rcs[i] = lib.qcuMrhsDslash(*args[i & 1])
    e1.record(st)
    torch.cuda.synchronize()
    for rc in rcs:
        _lib.check(rc)
    ms = e0.elapsed_time(e1) / steps
    return ms, best.as_tuple()


def time_dslash12(lib, _lib, torch, geom, prec, k, steps, warmup, L, seed=0, reals=12):
    """The same workload on compressed SU(3) links, two rows (reals = 12,
    qcuMrhsDslash12) or eight reals (reals = 8, qcuMrhsDslash8), with the
    tile of the native layout; returns ms_per_step."""
    from mrhs_qcu_amd import fields

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    S = fields.half_volume(L)
    key = (tuple(L), 3, seed)
    if key not in _GAUGE_CACHE:
        _GAUGE_CACHE.clear()
        _GAUGE_CACHE[key] = torch.view_as_real(fields.gauge_field_device(L, 3, seed=seed, device=dev)).reshape(-1)
    gu = _GAUGE_CACHE[key]
    dt = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16}[prec]
    nbytes, pack, run = ((lib.qcuMrhsGauge12Bytes, lib.qcuMrhsPackGauge12, lib.qcuMrhsDslash12) if reals == 12 else
                         (lib.qcuMrhsGauge8Bytes, lib.qcuMrhsPackGauge8, lib.qcuMrhsDslash8))
    g12 = torch.empty(nbytes(geom, prec) // _lib.BYTES[prec] * 2, dtype=dt, device=dev)
    # The generator folds antiperiodic time into the links: t_sign = -1.
    _lib.check(pack(g12.data_ptr(), prec, gu.data_ptr(), _lib.DOUBLE, geom, -1, sp))
    a = torch.randn(S * 4 * 3 * k * 2, device=dev, dtype=torch.float32).to(dt)
    ins = [a, a]
    o = torch.empty_like(a)
    outs = [o, o]
    args = [(outs[p].data_ptr(), ins[p].data_ptr(), g12.data_ptr(), prec, k, p, 0, -1, geom, None, sp) for p in (0, 1)]
    for i in range(warmup):
        _lib.check(run(*args[i & 1]))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rcs = [0] * steps
    e0.record(st)
    for i in range(steps):
        rcs[i] = run(*args[i & 1])
    e1.record(st)
    torch.cuda.synchronize()
    for rc in rcs:
        _lib.check(rc)
    return e0.elapsed_time(e1) / steps


def cg_benchmark(lib, _lib, torch, L=(32, 32, 32, 64), N=3, k=16, kappa=0.1350, tol=1e-10, seed=0,
                 solvers=("fp64", "mixed_fp16"), repeats=3, csw=None):
    """Even-odd preconditioned MRHS CG on M^+M^ to a true relative residual
    tol: pure FP64 vs mixed precision (FP16 dslash storage, FP32 Krylov
    vectors, FP64 reductions and reliable updates), same GPU, same inputs.
    Sources: RHS 0..min(k,4N)-1 unit deltas (spin-colour component r) at
    seeded even sites, the rest Gaussian noise. Each solver: one untimed
    2-iteration solve (tuning, work vectors), then `repeats` timed full solves
    from x = 0; seconds = the fastest, seconds_all = every one. With csw the
    operator is Wilson-clover, M^ = A_ee - kappa^2 D_eo A_oo^-1 D_oe with
    A = 1 - (kappa csw / 2) sigma.F built from the same links."""
    import numpy as np

    from mrhs_qcu_amd import fields

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    geom = _lib.geom(L, N)
    S = fields.half_volume(L)
    gu = torch.view_as_real(fields.gauge_field_device(L, N, seed=seed, device=dev)).reshape(-1)
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.DOUBLE) // 8, dtype=torch.float64, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.DOUBLE, gu.data_ptr(), _lib.DOUBLE, geom, sp))
    cl = cli = None
    if csw is not None:
        nc = lib.qcuMrhsCloverBytes(geom, _lib.DOUBLE) // 8
        cl = torch.empty(nc, dtype=torch.float64, device=dev)
        cli = torch.empty(nc, dtype=torch.float64, device=dev)
        _lib.check(lib.qcuMrhsCloverCompute(cl.data_ptr(), cli.data_ptr(), _lib.DOUBLE, gu.data_ptr(), _lib.DOUBLE,
                                            -kappa * csw / 2, geom, sp))
    del gu

    def solve(x, prm, res):
        if cl is None:
            return lib.qcuMrhsCgSolve(x.data_ptr(), b.data_ptr(), gn.data_ptr(), k, geom, ctypes.byref(prm), res, sp)
        return lib.qcuMrhsCgSolveClover(x.data_ptr(), b.data_ptr(), gn.data_ptr(), cl.data_ptr(), cli.data_ptr(), k,
                                        geom, ctypes.byref(prm), res, sp)

    b = torch.zeros(S, 4, N, k, 2, dtype=torch.float64, device=dev)
    npt = min(k, 4 * N)
    rng = np.random.Generator(np.random.PCG64(seed + 11))
    sites = rng.integers(0, S, size=npt)
    for r in range(npt):
        b[int(sites[r]), (r % (4 * N)) // N, r % N, r, 0] = 1.0
    if k > npt:
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed + 12)
        b[:, :, :, npt:, :] = torch.randn(S, 4, N, k - npt, 2, generator=gen, device=dev, dtype=torch.float64)
    op = "cg_mdagm" if csw is None else f"clover_csw{csw}_cg_mdagm"
    out = {"workload": f"su{N}_eo_{op}_{'x'.join(map(str, L))}_nrhs{k}_kappa{kappa}", "tol": tol,
           "sources": f"{npt} point sources + {k - npt} Gaussian"}
    sols = {}
    for name, sloppy in (("fp64", _lib.DOUBLE), ("mixed_fp16", _lib.HALF), ("mixed_fp8", _lib.FP8)):
        if name not in solvers:
            continue
        x = torch.empty_like(b)
        prm = _lib.CgParam()
        res = (ctypes.c_double * k)()
        # Untimed warm-up: tunes the operator tiles and allocates the solver's
        # work vectors (both kept for the timed solve).
        prm.kappa, prm.tol, prm.max_iter, prm.reliable_interval, prm.sloppy_prec = kappa, tol, 2, 50, sloppy
        _lib.check(solve(x, prm, res))
        # Timed: the full solve from x = 0, `repeats` times; the fastest is
        # reported (every repeat does the same work and must converge).
        walls = []
        for _ in range(repeats):
            prm = _lib.CgParam()
            prm.kappa, prm.tol, prm.max_iter, prm.reliable_interval, prm.sloppy_prec = kappa, tol, 20000, 50, sloppy
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _lib.check(solve(x, prm, res))
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
            if not max(res) <= tol:
                break
        out[name] = {"seconds": min(walls), "seconds_all": walls, "iterations": prm.iterations,
                     "reliable_updates": prm.reliable_updates, "max_true_rel_residual": max(res),
                     "converged": bool(max(res) <= tol)}
        sols[name] = x
    if len(sols) < 2:
        return out
    d = (sols["mixed_fp16"] - sols["fp64"]).permute(3, 0, 1, 2, 4).reshape(k, -1).norm(dim=1)
    n = sols["fp64"].permute(3, 0, 1, 2, 4).reshape(k, -1).norm(dim=1)
    out["max_rel_solution_diff"] = float((d / n).max())
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed_fp16"]["seconds"]
    if "mixed_fp8" in sols:
        d8 = (sols["mixed_fp8"] - sols["fp64"]).permute(3, 0, 1, 2, 4).reshape(k, -1).norm(dim=1)
        out["mixed_fp8"]["max_rel_solution_diff"] = float((d8 / n).max())
        out["mixed_fp8"]["speedup_vs_fp64"] = out["fp64"]["seconds"] / out["mixed_fp8"]["seconds"]
    return out


BATCH_WORKLOAD = [((16, 16, 16, 16), 2, 8), ((24, 24, 24, 48), 3, 12), ((16, 16, 16, 32), 4, 24), ((12, 12, 12, 24), 5, 6)]


def faces_benchmark(lib, _lib, torch, L=(24, 24, 24, 24), grid=(1, 1, 1, 2), N=3, k=16, steps=200, warmup=20):
    """One process's share of 24^3x48 split in two along t (24^3x24, FP32,
    16 RHS): interior operator (boundary terms skipped), face pack and the
    boundary terms from received faces (here the process's own faces), each
    timed with HIP events on the launch stream. Face bytes: pack reads the
    spinors of both faces and the links of one and writes two half-spinor
    faces; the boundary launch reads the faces and one face's links and
    reads and writes the output on both faces."""
    from mrhs_qcu_amd import fields

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    geom = _lib.geom(L, N, grid=grid)
    S = fields.half_volume(L)
    prec = _lib.SINGLE
    gu = torch.view_as_real(fields.gauge_field_device(L, N, seed=0, device=dev)).reshape(-1)
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, prec) // 4, dtype=torch.float32, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), prec, gu.data_ptr(), _lib.DOUBLE, geom, sp))
    del gu
    a = torch.randn(S * 4 * N * k * 2, device=dev, dtype=torch.float32)
    b = torch.empty_like(a)
    if RETUNE:
        _lib.check(lib.qcuMrhsTune(b.data_ptr(), a.data_ptr(), gn.data_ptr(), prec, k, geom, None, None, sp))
    tb, tf = [0] * 4, [0] * 4
    keep = []
    pack_bytes = faces_bytes = 0
    for mu in range(4):
        if grid[mu] > 1:
            nb = lib.qcuMrhsFaceBytes(geom, k, prec, mu)
            keep += [torch.empty(nb // 4, dtype=torch.float32, device=dev) for _ in range(2)]
            tb[mu], tf[mu] = keep[-2].data_ptr(), keep[-1].data_ptr()
            F = nb // (2 * N * k * 8)
            spinor, half, link = 4 * N * k * 8, 2 * N * k * 8, N * N * 8
            pack_bytes += F * (2 * spinor + link + 2 * half)
            faces_bytes += F * (2 * half + link + 2 * 2 * spinor)
    pb, kb = _lib.ptr_array(tb)
    pf, kf = _lib.ptr_array(tf)

    def timed(fn):
        for _ in range(warmup):
            _lib.check(fn())
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        rcs = [0] * steps
        e0.record(st)
        for i in range(steps):
            rcs[i] = fn()
        e1.record(st)
        torch.cuda.synchronize()
        for rc in rcs:
            _lib.check(rc)
        return e0.elapsed_time(e1) / steps

    ms_int = timed(lambda: lib.qcuMrhsDslash(b.data_ptr(), a.data_ptr(), gn.data_ptr(), prec, k, 0, 0, geom, None, sp))
    ms_pack = timed(lambda: lib.qcuMrhsPackFaces(pb, pf, a.data_ptr(), gn.data_ptr(), prec, k, 0, 0, geom, sp))
    ms_faces = timed(lambda: lib.qcuMrhsDslashFaces(b.data_ptr(), pf, pb, gn.data_ptr(), prec, k, 0, 0, geom, sp))
    return {"workload": f"su3_eo_wilson_dslash_{'x'.join(map(str, L))}_grid{''.join(map(str, grid))}_nrhs{k}_fp32",
            "interior_ms": ms_int, "pack_faces_ms": ms_pack, "dslash_faces_ms": ms_faces,
            "face_bytes_per_direction": lib.qcuMrhsFaceBytes(geom, k, prec, 3),
            "pack_GBs": pack_bytes / (ms_pack * 1e-3) / 1e9,
            "faces_GBs": faces_bytes / (ms_faces * 1e-3) / 1e9}


def clover_benchmark(lib, _lib, torch, L=(24, 24, 24, 48), N=3, k=16, steps=200, warmup=20):
    """Clover term A(x) psi(x) on one parity, 24^3x48, 16 RHS, FP32 (field
    built once from the links, coeff -0.1 arbitrary): time per application
    and its bytes (read in, write out, read the 8N^2 clover entries of each
    site once) over the time."""
    from mrhs_qcu_amd import fields

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    geom = _lib.geom(L, N)
    S = fields.half_volume(L)
    gu = torch.view_as_real(fields.gauge_field_device(L, N, seed=0, device=dev)).reshape(-1)
    cl = torch.empty(lib.qcuMrhsCloverBytes(geom, _lib.SINGLE) // 4, dtype=torch.float32, device=dev)
    _lib.check(lib.qcuMrhsCloverCompute(cl.data_ptr(), None, _lib.SINGLE, gu.data_ptr(), _lib.DOUBLE, -0.1, geom, sp))
    del gu
    a = torch.randn(S * 4 * N * k * 2, device=dev, dtype=torch.float32)
    b = torch.empty_like(a)
    for i in range(warmup):
        _lib.check(lib.qcuMrhsCloverApply(b.data_ptr(), a.data_ptr(), cl.data_ptr(), _lib.SINGLE, k, i & 1, geom, sp))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rcs = [0] * steps
    e0.record(st)
    for i in range(steps):
        rcs[i] = lib.qcuMrhsCloverApply(b.data_ptr(), a.data_ptr(), cl.data_ptr(), _lib.SINGLE, k, i & 1, geom, sp)
    e1.record(st)
    torch.cuda.synchronize()
    for rc in rcs:
        _lib.check(rc)
    ms = e0.elapsed_time(e1) / steps
    nbytes = S * (k * 4 * N * 8 * 2 + 8 * N * N * 8)
    return {"workload": f"su{N}_clover_apply_{'x'.join(map(str, L))}_nrhs{k}_fp32", "ms_per_step": ms,
            "sites_rhs_per_s": S * k / (ms * 1e-3), "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def batch_benchmark(lib, _lib, torch, steps=50, warmup=5):
    """Uneven batched workload in one HIP graph (SU(2) 16^4 nrhs=8, SU(3)
    24^3x48 nrhs=12, SU(4) 16^3x32 nrhs=24, SU(5) 12^3x24 nrhs=6, FP32, tuned
    tile per problem): time per graph replay against the same launches issued
    one after another on one stream."""
    from mrhs_qcu_amd import fields

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    probs = (_lib.Problem * len(BATCH_WORKLOAD))()
    keep, tot_sites_rhs, tot_bytes = [], 0, 0
    for i, (L, N, k) in enumerate(BATCH_WORKLOAD):
        geom = _lib.geom(L, N)
        S = fields.half_volume(L)
        gu = torch.view_as_real(fields.gauge_field_device(L, N, seed=100 + i, device=dev)).reshape(-1)
        gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.SINGLE) // 4, dtype=torch.float32, device=dev)
        _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.SINGLE, gu.data_ptr(), _lib.DOUBLE, geom, sp))
        del gu
        a = torch.randn(S * 4 * N * k * 2, device=dev, dtype=torch.float32)
        o = torch.empty_like(a)
        probs[i].geom = geom
        probs[i].nrhs, probs[i].parity, probs[i].dagger = k, i & 1, 0
        probs[i].out, probs[i].in_, probs[i].gauge = o.data_ptr(), a.data_ptr(), gn.data_ptr()
        keep += [gn, a, o]
        tot_sites_rhs += S * k
        tot_bytes += bytes_alg(S, k, N, 4, 4, 4)
    h = ctypes.c_void_p()
    _lib.check(lib.qcuMrhsBatchCreate(probs, len(BATCH_WORKLOAD), ctypes.byref(h), sp))
    try:
        for _ in range(warmup):
            _lib.check(lib.qcuMrhsBatchRun(h, sp))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        rcs = []
        e0.record(st)
        for _ in range(steps):
            rcs.append(lib.qcuMrhsBatchRun(h, sp))
        e1.record(st)
        torch.cuda.synchronize()
        ms_graph = e0.elapsed_time(e1) / steps
        e0.record(st)
        for _ in range(steps):
            for p in probs:
                rcs.append(lib.qcuMrhsDslash(p.out, p.in_, p.gauge, _lib.SINGLE, p.nrhs, p.parity, p.dagger, p.geom,
                                             None, sp))
        e1.record(st)
        torch.cuda.synchronize()
        for rc in rcs:
            _lib.check(rc)
        ms_seq = e0.elapsed_time(e1) / steps
    finally:
        lib.qcuMrhsBatchDestroy(h)
    return {"workload": "batch su2_16^4_k8 + su3_24^3x48_k12 + su4_16^3x32_k24 + su5_12^3x24_k6 fp32",
            "ms_per_graph_replay": ms_graph, "ms_sequential_launches": ms_seq,
            "sites_rhs_per_s": tot_sites_rhs / (ms_graph * 1e-3),
            "hbm_frac": tot_bytes / (ms_graph * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_alg_per_replay": tot_bytes}


def qcu_h_end_to_end(lib, _lib, torch, L=(24, 24, 24, 48), N=3, k=16, reps=20):
    """The reference's call protocol through include/qcu.h, FP32 user buffers:
    per call k pushBackFermions (input copied at push) + one startDslash
    (synchronous). Wall time per call, host included."""
    from mrhs_qcu_amd import fields

    dev = torch.device("cuda:0")
    S = fields.half_volume(L)
    grid, param = _lib.QcuGrid(), _lib.QcuParam()
    for d in range(4):
        grid.grid_size[d] = 1
        param.lattice_size[d] = L[d]
    lib.initGridSize(ctypes.byref(grid), ctypes.byref(param), N)
    gu = fields.gauge_field_device(L, N, seed=3, device=dev)
    lib.loadQcuGauge(gu.data_ptr())
    lib.getDslash(0, 0.0, N, k, 1)
    ins = [torch.randn(S * 4 * N * 2, device=dev, dtype=torch.float32) for _ in range(k)]
    outs = [torch.empty_like(t) for t in ins]
    torch.cuda.synchronize()
    for i in range(3):
        for r in range(k):
            lib.pushBackFermions(outs[r].data_ptr(), ins[r].data_ptr())
        lib.startDslash(i & 1, 0)
    t0 = time.perf_counter()
    for i in range(reps):
        for r in range(k):
            lib.pushBackFermions(outs[r].data_ptr(), ins[r].data_ptr())
        lib.startDslash(i & 1, 0)
    wall = (time.perf_counter() - t0) / reps
    del gu
    # What a push could do with one input: copy it whole (what it does), or
    # write it into its slot of a 16-RHS packed field (8 bytes per 128-byte
    # line), the alternative to a copy plus one full pack at startDslash.
    nbytes = S * 4 * N * 8
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    src = ins[0]
    dst = torch.empty(nbytes * 16 // 4, dtype=torch.float32, device=dev)
    per_rhs = {}
    for mode, name in ((2, "contiguous_copy_us"), (4, "into_16_rhs_packed_slot_us")):
        gbs = ctypes.c_double()
        _lib.check(lib.qcuMrhsStreamBandwidth(mode, dst.data_ptr(), src.data_ptr(), nbytes, 20, ctypes.byref(gbs), sp))
        per_rhs[name] = 2 * nbytes / (gbs.value * 1e9) * 1e6
    return {"workload": f"qcu.h push x{k} + startDslash, su{N} {'x'.join(map(str, L))} fp32 user layout",
            "ms_per_call": wall * 1e3, "sites_rhs_per_s": S * k / wall, "one_rhs_device_time": per_rhs}


def link_block_benchmark(lib, _lib, torch, cases=((3, 16), (4, 24), (5, 6)), reps=2000):
    """The hop's link x half-spinor block products, Y (N x 2k) = U H, on the
    vector ALUs (the production kernels' form) and on the matrix cores
    (v_mfma_f32_16x16x4_f32 on the real embedding), operands reloaded from
    cache per block or held in registers: useful FP32 TFLOP/s of each and
    their errors against an FP64 product."""
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for N, k in cases:
        row = {}
        for resident, name in ((0, "operands_from_cache"), (1, "operands_in_registers")):
            tf = (ctypes.c_double * 2)()
            err = (ctypes.c_double * 2)()
            _lib.check(lib.qcuMrhsLinkBlockBench(N, k, reps, resident, tf, err, sp))
            row[name] = {"valu_tflops": tf[0], "mfma_tflops": tf[1]}
            row["valu_rel_err"], row["mfma_rel_err"] = err[0], err[1]
        out[f"su{N}_nrhs{k}"] = row
    return out


def cpu_model():
    """Host CPU model name from /proc/cpuinfo (or 'unknown')."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


STREAM_MODES = ("read", "write", "copy", "read_write")


def stream_benchmark(lib, _lib, torch, sizes_mb=(510,), reps=20):
    """Streaming ceiling of this GPU: the library's 16-byte-access stream
    kernels (read, write, copy, in-place read+write) over buffers of the given
    sizes; GB/s counts bytes read plus bytes written."""
    dev = torch.device("cuda:0")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for mb in sizes_mb:
        nbytes = int(mb * 2**20) // 16 * 16
        a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        b = torch.zeros_like(a)
        row = {}
        for mode, name in enumerate(STREAM_MODES):
            gbs = ctypes.c_double()
            _lib.check(lib.qcuMrhsStreamBandwidth(mode, b.data_ptr(), a.data_ptr(), nbytes, reps, ctypes.byref(gbs),
                                                  sp))
            row[name] = round(gbs.value, 1)
        out[f"{mb}MiB"] = row
        del a, b
    return out


def cpu_baseline(L, N, k, budget_s=12.0):
    """FP32 C oracle (OpenMP) on the host: same workload, bounded repetitions."""
    import numpy as np

    from mrhs_qcu_amd import fields
    from oracle import oracle

    S = fields.half_volume(L)
    rng = np.random.default_rng(7)
    g = fields.gauge_field(L, N, seed=0).astype(np.complex64)
    psi = (rng.standard_normal((k, S, 4, N), dtype=np.float32)
           + 1j * rng.standard_normal((k, S, 4, N), dtype=np.float32)).astype(np.complex64)
    out = np.empty_like(psi)
    oracle.dslash(g, psi, L, 0, 0, single=True, out=out)  # warm-up
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        oracle.dslash(g, psi, L, len(times) & 1, 0, single=True, out=out)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 50:
            break
    t = float(np.median(times))
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    return {"value": S * k / t, "unit": "sites*RHS/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"FP32 C restatement of the reference applyWilson (oracle/oracle.c), OpenMP, "
                      f"{'x'.join(map(str, L))} SU({N}) nrhs={k}, median of {len(times)} applications "
                      f"({sum(times):.1f} s)"}


def traffic_from_profile(name):
    """(bytes per launch, tile it was measured with) from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary, or (None, None)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None, None
    try:
        with open(p) as f:
            d = json.load(f)
        return float(d["hbm_bytes_per_launch"]), d.get("tile")
    except Exception:
        return None, None


def main():
    args = parse()
    global RETUNE
    RETUNE = args.retune
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    import torch

    from mrhs_qcu_amd import _lib, fields

    lib = _lib.load()
    L, N, k = tuple(args.L), 3, args.nrhs
    geom = _lib.geom(L, N)
    S = fields.half_volume(L)
    if args.copy_calibration:
        # One device-to-device copy of exactly the headline input field size:
        # FETCH_SIZE / WRITE_SIZE of this kernel calibrate the counters.
        nbytes = S * k * 4 * N * 8
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)
        return
    if args.stream is not None:
        sizes = tuple(args.stream) or (16, 64, 128, 510, 2048)
        print(json.dumps({"stream_GBs": stream_benchmark(lib, _lib, torch, sizes_mb=sizes)}), flush=True)
        return
    if args.faces_only:
        print(json.dumps(faces_benchmark(lib, _lib, torch)), flush=True)
        print(json.dumps(clover_benchmark(lib, _lib, torch)), flush=True)
        return
    if args.link_block:
        print(json.dumps(link_block_benchmark(lib, _lib, torch)), flush=True)
        return
    if args.cg_only:
        solvers = {"both": ("fp64", "mixed_fp16"), "fp64": ("fp64",), "mixed": ("mixed_fp16",),
                   "fp8": ("fp64", "mixed_fp16", "mixed_fp8")}[args.cg_solver]
        print(json.dumps(cg_benchmark(lib, _lib, torch, L=tuple(args.cg_L), solvers=solvers, csw=args.cg_clover)),
              flush=True)
        return
    pcode = {"fp16": _lib.HALF, "bf16": _lib.BF16, "fp32": _lib.SINGLE, "fp64": _lib.DOUBLE,
             "fp8": _lib.FP8}[args.precision]
    pb = {"fp16": 2, "bf16": 2, "fp32": 4, "fp64": 8, "fp8": 1}[args.precision]
    tile_arg = tuple(int(v) for v in args.tile.split(",")) if args.tile else None
    ms, tile = time_dslash(lib, _lib, torch, geom, pcode, k, args.steps, args.warmup, L, N, tile=tile_arg,
                           settle_ms=args.settle_ms)
    value = S * k / (ms * 1e-3)
    ba = bytes_alg(S, k, N, pb, pb, 2 if pcode == _lib.FP8 else pb)  # FP8 spinors use binary16 links
    achieved = ba / (ms * 1e-3) / 1e9
    wl = f"su3_mrhs_eo_wilson_dslash_{'x'.join(map(str, L))}_nrhs{k}_{args.precision}"
    traffic, traffic_tile = traffic_from_profile(f"{wl}_traffic.json")
    if traffic_tile != ",".join(str(v) for v in tile):
        traffic = None  # measured on another tile: not evidence for this one
    line = {
        "metric": "MRHS dslash sites\u00b7RHS/s + %HBM roofline; mixed vs FP64 CG time to 1e-10",
        "value": value,
        "unit": "sites*RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp16": "f16 storage, f32 arithmetic", "bf16": "bf16 storage, f32 arithmetic", "fp32": "f32",
                  "fp64": "f64", "fp8": "e4m3 spinor storage, f16 links, f32 arithmetic"}[args.precision],
        "data": "synthetic (seeded SU(3) links, Gaussian spinors)",
        "config": {"workload": wl, "lattice": list(L), "nColors": N, "nrhs": k, "storage": args.precision,
                   "settle_ms": args.settle_ms,
                   "tile": dict(zip(("rhs_per_lane", "block", "link_lds", "t_block", "z_block", "cache_hints",
                                     "occupancy", "traversal"), tile)),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_tile": traffic_tile,
                     "bytes_alg_per_launch": ba},
    }
    if not args.profile and not args.no_extra:
        extra = {}
        for name, prec, b in (("fp16", _lib.HALF, 2), ("bf16", _lib.BF16, 2), ("fp64", _lib.DOUBLE, 8),
                              ("fp8_spinors_fp16_links", _lib.FP8, 1)):
            ms2, t2 = time_dslash(lib, _lib, torch, geom, prec, k, max(20, args.steps // 4), 10, L, N)
            bb = bytes_alg(S, k, N, b, b, 2 if prec == _lib.FP8 else b)
            extra[name] = {"ms_per_step": ms2, "sites_rhs_per_s": S * k / (ms2 * 1e-3),
                           "hbm_frac": bb / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBS, "tile": t2}
        for name, prec, b, reals in (("two_row_links_fp32", _lib.SINGLE, 4, 12), ("two_row_links_fp16", _lib.HALF, 2, 12),
                                     ("eight_real_links_fp32", _lib.SINGLE, 4, 8),
                                     ("eight_real_links_fp16", _lib.HALF, 2, 8)):
            ms12 = time_dslash12(lib, _lib, torch, geom, prec, k, max(20, args.steps // 4), 10, L, reals=reals)
            bb12 = S * (k * 8 * N * 2 * b + 8 * reals * b)
            extra[name] = {"ms_per_step": ms12, "sites_rhs_per_s": S * k / (ms12 * 1e-3),
                           "hbm_frac": bb12 / (ms12 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                           "hbm_frac_vs_full_link_bytes": bytes_alg(S, k, N, b, b, b) / (ms12 * 1e-3) / 1e9 /
                           HBM_PEAK_GBS}
        Lc2, kc2 = (16, 16, 16, 32), 12
        g2 = _lib.geom(Lc2, N)
        ms3, t3 = time_dslash(lib, _lib, torch, g2, _lib.SINGLE, kc2, max(50, args.steps // 2), 10, Lc2, N)
        S2 = fields.half_volume(Lc2)
        extra["c2_16x16x16x32_nrhs12_fp32"] = {
            "ms_per_step": ms3, "sites_rhs_per_s": S2 * kc2 / (ms3 * 1e-3),
            "hbm_frac_cache_resident": bytes_alg(S2, kc2, N, 4, 4, 4) / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "tile": t3}
        extra["stream_ceiling_GBs"] = stream_benchmark(lib, _lib, torch, sizes_mb=(510, 2048))
        extra["batch"] = batch_benchmark(lib, _lib, torch)
        extra["link_block_valu_vs_mfma"] = link_block_benchmark(lib, _lib, torch)
        extra["faces"] = faces_benchmark(lib, _lib, torch)
        extra["clover"] = clover_benchmark(lib, _lib, torch)
        extra["qcu_h_end_to_end"] = qcu_h_end_to_end(lib, _lib, torch)
        line["extra"] = extra
        if not args.no_cg:
            line["cg"] = cg_benchmark(lib, _lib, torch, L=tuple(args.cg_L),
                                      solvers=("fp64", "mixed_fp16", "mixed_fp8"))
            line["clover_cg"] = cg_benchmark(lib, _lib, torch, L=tuple(args.cg_L), repeats=1, csw=1.0)
    if not args.no_cpu and not args.profile:
        line["cpu_baseline"] = cpu_baseline(L, N, k)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
