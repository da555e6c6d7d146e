"""Benchmark of the ED impurity-solver hot path on one MI355X.

A step is one y <- y + H x over the 4x4 Hubbard cluster (t=1, t'=-0.3, U=8, mu=4) on its
half-filled N=16, Sz=0 sector: D = 12870^2 = 165,636,900 FP64 amplitudes, algorithmic
traffic 24 D bytes (read x, read y, write y). Inputs are resident in HBM before timing.

Prints one JSON line (rank 0, N=1):
  value        whole-job H.v throughput in GB/s of algorithmic bytes
  roofline     achieved GB/s of the H.v kernel(s) from HIP events on the library stream,
               the HBM bytes per launch from the committed counter profile, and the
               achievable read-modify-write bandwidth measured in the same run (y += a x)
  config       also the p10 / p50 / p90 of 100 single timed launches
  cpu_baseline the oracle's H.v (reference loop order, one core) on the same sector
  cpu_parallel the same H.v with the oracle's OpenMP-over-rows form on up to 16 cores
  lanczos      ms and GB/s per ground-state Lanczos iteration on the same sector (48 D bytes:
               the product plus the vector update), fused and as separate passes, with
               3 iterations of the oracle's recurrence timed on 1 and up to 16 host cores
  ground_state the workload's Lanczos ground state (E0, iterations, seconds) on the
               factorized sector and on irrep 0 of the cluster's two-mirror basis
  cpt          the CPT k-sum of the 3x4 cluster (128^2 k-points x 512 frequencies), device
               time and FP64 rate on 41.5 kflop per (k, w)
  cdmft        the second half of the metric: converged CDMFT solves per second on the
               2x2 plaquette + 8 bath orbitals (16 bath parameters), n seeded points
               (U ~ U[0.5, 12], mu ~ U[0, U], t' ~ U[-0.5, 0.5]), solved by the batched
               engine from the fixed cold start and from the learned warm start (the MLP
               trained on the GPU on the committed training set, BF16 inference), at the
               same convergence tests and tolerances. A "solve" is one parameter point
               converged by the CDMFT loop. A problem that fails from the learned start is
               run again from the cold start afterwards (the plain loop is the fallback),
               reported as warm_with_fallback with its time added; warm solves that reach
               another fixed point than the plain loop's are reported separately
               (speedup_at_plain_fixed_point). --cdmft-jacobian adds a leg that also
               starts Broyden from the inverse Jacobian of the nearest training run.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--variant", type=int, default=-1, help="H.v kernel variant (-1: library default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--sweep", action="store_true", help="time every H.v variant (stderr)")
    ap.add_argument("--cdmft", type=int, default=64, help="CDMFT points for the warm-start leg (0: skip)")
    ap.add_argument("--cdmft-batch", type=int, default=32, help="CDMFT problems active at a time")
    ap.add_argument("--lanczos", type=int, default=1, help="time ground-state Lanczos iterations (0: skip)")
    ap.add_argument("--cpt", type=int, default=1, help="time the CPT k-sum on the 3x4 cluster (0: skip)")
    ap.add_argument("--mlp-ensemble", type=int, default=1, help="warm start: mean prediction of this many MLPs")
    ap.add_argument("--cdmft-jacobian", action="store_true",
                    help="add a warm leg that also starts Broyden from the nearest training run's inverse Jacobian")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched Lanczos steps as direct launches instead of hipGraph replays")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 2 GiB five times (for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the solver runs on one GPU only (the single-GPU H.v path); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return

    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import Sector

    _lib.call("qed_init", 0)
    vals = {"t": 1.0, "tp": -0.3, "U": 8.0, "mu": 4.0}
    t0 = time.time()
    S = Sector(16, 8, 8, cluster_hamiltonian_4x4(vals))
    log("sector built: D=%d (%.2f s)" % (S.dim, time.time() - t0))
    D = S.dim
    rng = np.random.default_rng(20261202)
    x = rng.standard_normal(D)
    x /= np.linalg.norm(x)
    y = rng.standard_normal(D)
    dx = DeviceArray.from_host(x)
    dy = DeviceArray.from_host(y)
    del y
    bytes_per = 24.0 * D

    import ctypes
    ms = np.zeros(1)
    if args.sweep:
        for v in (0, 1, 4, 5, 12):
            _lib.call("qed_hv_timed", S.h, v, dx.ptr, dy.ptr, 3, 10, _lib.dp(ms))
            log("variant %d: %.3f ms per H.v -> %.0f GB/s" % (v, ms[0], bytes_per / ms[0] / 1e6))

    variant = args.variant
    # warmup
    for _ in range(args.warmup):
        S.hv_dev(dx, dy, None if variant < 0 else variant)
    _lib.call("qed_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.hv_dev(dx, dy, None if variant < 0 else variant)
    _lib.call("qed_sync")
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = bytes_per * args.steps / wall / 1e9
    # kernel-only time with HIP events on the library stream
    _lib.call("qed_hv_timed", S.h, variant, dx.ptr, dy.ptr, 2, max(10, args.steps), _lib.dp(ms))
    achieved = bytes_per / (ms[0] * 1e-3) / 1e9
    log("H.v: %.3f ms/step wall, %.3f ms kernel -> %.1f GB/s (%.1f%% of 8 TB/s)"
        % (ms_step, ms[0], achieved, 100 * achieved / PEAK_HBM_GBS))
    # spread of single launches (HIP events around each)
    one = []
    for _ in range(100):
        _lib.call("qed_hv_timed", S.h, variant, dx.ptr, dy.ptr, 0, 1, _lib.dp(ms))
        one.append(ms[0])
    p10, p50, p90 = (float(np.percentile(one, q)) for q in (10, 50, 90))
    log("H.v single launches: p10 %.3f p50 %.3f p90 %.3f ms" % (p10, p50, p90))
    _lib.call("qed_stream_rmw_timed", dx.ptr, dy.ptr, D - (D % 2), 20, _lib.dp(ms))
    stream = 24.0 * (D - (D % 2)) / (ms[0] * 1e-3) / 1e9
    log("stream y += a x (24 B per element): %.0f GB/s" % stream)

    lanczos = lanczos_leg(S, dx) if args.lanczos else None

    cpu = cpu_par = None
    if not args.no_cpu:
        # one full H.v of the same sector through the oracle (reference loop order, 1 core)
        from oracle import oracle as O
        nn, nnn = O.square_bonds(4, 4)
        model, ovals = O.hubbard_cluster(16, nn, tp_bonds=nnn, **{k: vals[k] for k in ("tp", "U", "mu")})
        So = O.Sector(model, ovals, 8, 8)
        yc = np.zeros(D)
        xh = x
        t0 = time.perf_counter()
        O.lib().or_hv(So.h, O._dp(xh), O._dp(yc))
        tc = time.perf_counter() - t0
        cpu = {"value": round(bytes_per / tc / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": "one full y += H x on the same N=16 Sz=0 sector (D=165,636,900) through the "
                         "oracle's restatement of Hamiltonian_Factorized::mult_add, %.1f s" % tc}
        log("cpu baseline: %.2f s per H.v" % tc)
        nt = min(16, os.cpu_count() or 1)
        yc[:] = 0.0
        t0 = time.perf_counter()
        O.lib().or_hv_omp(So.h, O._dp(xh), O._dp(yc), nt)
        tp = time.perf_counter() - t0
        cpu_par = {"value": round(bytes_per / tp / 1e9, 4), "unit": "GB/s", "cores": nt,
                   "kind": "port, OpenMP over rows (a parallelization of ours; the reference's loop is serial)",
                   "sample": "one full y += H x on the same sector, %.2f s" % tp}
        log("cpu parallel (%d threads): %.2f s per H.v" % (nt, tp))
        if lanczos is not None:
            # a bounded sample of the same Lanczos iteration on the host: 3 iterations
            ksteps = 3
            for name, threads in (("cpu_1core", 0), ("cpu_parallel", nt)):
                t0 = time.perf_counter()
                O.lib().or_lanczos_steps(So.h, O._dp(xh), ksteps, threads, None, None)
                tl = (time.perf_counter() - t0) / ksteps
                lanczos[name] = {"ms_per_iteration": round(tl * 1e3, 1), "cores": max(1, threads),
                                 "kind": "port", "sample": "%d iterations of the oracle's restatement of "
                                 "the reference recurrence from the same start vector" % ksteps}
                log("cpu Lanczos (%d threads): %.2f s per iteration" % (max(1, threads), tl))
        del So

    dx.free()
    dy.free()
    del S
    gs = ground_state_leg(cluster_hamiltonian_4x4(vals)) if args.lanczos else None
    cpt = cpt_leg(cpu=not args.no_cpu) if args.cpt else None
    cdmft = cdmft_leg(args.cdmft, args.cdmft_batch, not args.no_graph, args.cdmft_jacobian, args.mlp_ensemble) if args.cdmft > 0 else None

    out = {
        "metric": "H·v GB/s (16-site ED, % HBM peak); CDMFT solves/s learned warm-start vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "4x4 Hubbard cluster t=1 t'=-0.3 U=8 mu=4, sector N=16 Sz=0 "
                        "(D=165,636,900), one y += H x per step, 24*D algorithmic bytes",
            "sector_dim": D,
            "bytes_per_step": bytes_per,
            "hv_variant": variant,
            "hv_ms_p10_p50_p90": [round(p10, 4), round(p50, 4), round(p90, 4)],
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_from_profile(),
                     "stream_rmw_measured": round(stream, 1), "frac_of_stream": round(achieved / stream, 4)},
        "cpu_baseline": cpu,
        "cpu_parallel": cpu_par,
        "lanczos": lanczos,
        "ground_state": gs,
        "cpt": cpt,
        "cdmft": cdmft,
    }
    print(json.dumps(out), flush=True)


def cluster_hamiltonian_4x4(vals):
    """The 4x4 cluster's operators from the square-lattice tiling (intra-cluster NN hop t,
    NNN hop tp, chemical potential mu) and the on-site Hubbard U, in the flat form the
    sector constructor takes (0-based spin orbitals: site i up = i, down = i + 16)."""
    from pyqcm_amd import lattice as LT
    from pyqcm_amd.sector import flatten_ops
    T = LT.square_tiling(4, 4)
    ops = [(name, "one-body", [(r - 1, c - 1, v) for (r, c, v) in T.cluster_operator_elements(name)])
           for name in ("t", "tp", "mu")]
    ops.append(("U", "interaction", [(i, i + 16, 1.0) for i in range(16)]))
    return flatten_ops(ops, vals)


def lanczos_leg(S, dx, steps=20):
    """Ground-state Lanczos iterations on the bench sector from the normalized x: the fused
    form (product with the rescaling and <w, Hw> folded in, then one axpy + norm pass) and
    the separate passes, ms per iteration from HIP events. Algorithmic bytes per iteration:
    48 D (the product's 24 D plus 24 D for the update w' = y - a w with its norm)."""
    from pyqcm_amd import _lib
    ms = np.zeros(2)
    _lib.call("qed_lanczos_steps_timed", S.h, dx.ptr, 2, _lib.dp(ms))  # warm-up
    _lib.call("qed_lanczos_steps_timed", S.h, dx.ptr, steps, _lib.dp(ms))
    per = 48.0 * S.dim
    fused = float(ms[0]) if ms[0] > 0 else None
    log("Lanczos iteration: fused %s ms, separate passes %.3f ms" %
        ("%.3f" % fused if fused else "n/a", ms[1]))
    return {"workload": "ground-state Lanczos iterations on the same sector, %d timed" % steps,
            "bytes_per_iteration": per,
            "ms_per_iteration": round(fused if fused else float(ms[1]), 4),
            "ms_per_iteration_separate_passes": round(float(ms[1]), 4),
            "GB_s": round(per / ((fused if fused else ms[1]) * 1e-3) / 1e9, 1),
            "fused": fused is not None}


def ground_state_leg(flat):
    """The workload's ground state by Lanczos (seed 0, accur 1e-12) on the factorized sector
    and on irrep 0 of the symmetry-adapted basis of the cluster's two mirrors (C2v; the
    ground state is fully symmetric): E0, iterations and wall seconds of each, and the
    mirror basis's product time (HIP events)."""
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import MixedSector, Sector
    out = {}
    S = Sector(16, 8, 8, flat)
    t0 = time.perf_counter()
    E0, psi, n, *_ = S.lanczos_gs(seed=0, accur=1e-12, max_iter=600)
    out["factorized"] = {"dim": S.dim, "E0": E0, "iterations": n, "seconds": round(time.perf_counter() - t0, 3)}
    psi.free()
    S.close()
    mx = [(3 - i % 4) + 4 * (i // 4) for i in range(16)]
    my = [i % 4 + 4 * (3 - i // 4) for i in range(16)]
    t0 = time.perf_counter()
    M = MixedSector(16, 8, 8, flat, 16, [mx, my], 0, False)
    tb = time.perf_counter() - t0
    rng = np.random.default_rng(3)
    dx = DeviceArray.from_host(rng.standard_normal(M.dim))
    dy = DeviceArray.from_host(rng.standard_normal(M.dim))
    ms = np.zeros(1)
    _lib.call("qed_hv_timed", M.h, -1, dx.ptr, dy.ptr, 2, 10, _lib.dp(ms))
    dx.free()
    dy.free()
    t0 = time.perf_counter()
    E1, psi, n1, *_ = M.lanczos_gs(seed=0, accur=1e-12, max_iter=600)
    out["mirror_basis_irrep0"] = {"dim": M.dim, "E0": E1, "iterations": n1,
                                  "seconds": round(time.perf_counter() - t0, 3), "build_seconds": round(tb, 3),
                                  "hv_ms": round(float(ms[0]), 3),
                                  "hv_GB_s": round(24.0 * M.dim / (ms[0] * 1e-3) / 1e9, 1)}
    psi.free()
    M.close()
    out["E0_difference"] = abs(E1 - E0)
    log("ground state: factorized E0 %.12f (%d it, %.2f s); mirror basis irrep 0 E0 %.12f (%d it, %.2f s, "
        "H.v %.2f ms)" % (E0, n, out["factorized"]["seconds"], E1, n1, out["mirror_basis_irrep0"]["seconds"], ms[0]))
    return out


def copy_calibration():
    """A plain copy kernel of known size: 2 GiB read + 2 GiB written per launch."""
    import torch
    n = (2 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 5 launches of %d bytes read + %d bytes written" % (8 * n, 8 * n))


def traffic_from_profile():
    """HBM bytes per H.v launch measured with rocprofv3 counters (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "hv_16site_traffic.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return float(ln.split(":", 1)[1])
    return None


PEAK_FP64_TFLOPS = 78.6
CPT_FLOP_PER_POINT = 41.5e3  # two complex 12 x 12 products and one LU inverse per (k, w)


def cpt_leg(side=128, nw=512, reps=3, cpu=True):
    """CPT spectral trace on the 3x4 Hubbard cluster (t=1, U=8, mu=4, sector N=12 S=0,
    D = 853,776): G_cpt = G_c (1 - V(k) G_c)^-1 over a side^2 k-grid (k_i = (i+0.5)/side
    in the superdual basis) and nw frequencies w = linspace(-10, 10) + 0.1i, one device
    launch; device time from HIP events, algorithmic flops 41.5 kflop per (k, w).
    With cpu: the same cluster's Lanczos ground state from the seed-0 start vector on the
    device and through the oracle's recurrence (one core), and the oracle's dense
    G_c (1 - V G_c)^-1 on a sample of 16 k-points x nw frequencies."""
    from pyqcm_amd import _lib
    from pyqcm_amd import lattice as LT
    T = LT.square_tiling(3, 4)
    M = LT.LatticeModel("cpt_bench", T)
    vals = {"t": 1.0, "U": 8.0, "mu": 4.0}
    inst = M.instance(vals, "R0:N12:S0")
    t0 = time.perf_counter()
    E0, sec = inst.ground_state()
    z = np.linspace(-10.0, 10.0, nw) + 0.1j
    _, Geff = inst._geff(z)
    t_ed = time.perf_counter() - t0
    nv, mats = T.fourier_terms(vals)
    V0 = inst.v0()
    k = T.kgrid(side)
    LT.cpt_batch(Geff, V0, nv, mats, k[:1024], mode="trace")  # warm-up
    walls, devs = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        LT.cpt_batch(Geff, V0, nv, mats, k, mode="trace")
        walls.append(time.perf_counter() - t0)
        ms = np.zeros(1)
        _lib.call("qed_ksum_last_ms", _lib.dp(ms))
        devs.append(float(ms[0]))
    dev = float(np.median(devs))
    flops = CPT_FLOP_PER_POINT * nw * len(k)
    tf = flops / (dev * 1e-3) / 1e12
    log("CPT 3x4: ED + G_c at %d frequencies %.2f s; k-sum over %d x %d points: %.2f ms device, %.2f ms wall "
        "-> %.2f TFLOP/s" % (nw, t_ed, nw, len(k), dev, float(np.median(walls)) * 1e3, tf))
    extra = {}
    if cpu:
        from oracle import oracle as O
        from pyqcm_amd import _lib
        from pyqcm_amd.sector import Sector
        # Lanczos ground state of the cluster: device and oracle (one core), same start vector
        ops = [(nm, "one-body", [(r - 1, c - 1, v) for (r, c, v) in T.cluster_operator_elements(nm)])
               for nm in ("t", "mu")]
        ops.append(("U", "interaction", [(i, i + 12, 1.0) for i in range(12)]))
        from pyqcm_amd.sector import flatten_ops
        S = Sector(12, 6, 6, flatten_ops(ops, vals))
        x0 = S.start_vector(0).to_host()
        t0 = time.perf_counter()
        Eg, _, ng, _, _ = S.lanczos_gs(start=x0)
        _lib.call("qed_sync")
        tg = time.perf_counter() - t0
        S.close()
        nn, _ = O.square_bonds(3, 4)
        model, ovals = O.hubbard_cluster(12, nn, t=1.0, U=8.0, mu=4.0)
        So = O.Sector(model, ovals, 6, 6)
        t0 = time.perf_counter()
        Eo, _, no, _, _ = So.lanczos_gs(start=x0)
        tl = time.perf_counter() - t0
        log("Lanczos 3x4: device %.3f s (%d iterations, E0 %.12f); oracle 1 core %.2f s (%d, %.12f)"
            % (tg, ng, Eg, tl, no, Eo))
        extra["lanczos"] = {"E0_device": Eg, "E0_oracle": Eo, "iterations_device": ng, "iterations_oracle": no,
                            "seconds_device": round(tg, 4), "seconds_oracle_1core": round(tl, 3),
                            "kind": "port", "cores": 1}
        # CPT: the oracle's dense restatement on 16 k-points x nw frequencies
        Gc = inst.green(z)
        tc = T.cluster_matrix(vals)
        ks = k[:: len(k) // 16][:16]
        tk = T.tk(vals, ks)
        t0 = time.perf_counter()
        tr = np.array([[np.trace(O.cpt_green(Gc[a], tk[b] - tc)) for b in range(len(ks))] for a in range(nw)])
        tcpu = time.perf_counter() - t0
        dev_tr = LT.cpt_batch(Geff, V0, nv, mats, ks, mode="trace")
        agree = float(np.abs(dev_tr - tr).max() / np.abs(tr).max())
        pts = nw * len(ks)
        log("CPT 3x4 oracle: %d points in %.2f s (1 core) -> %.0f points/s; device %.3g points/s; "
            "max rel. difference %.1e" % (pts, tcpu, pts / tcpu, nw * len(k) / (dev * 1e-3), agree))
        extra["cpu_baseline"] = {"value": round(pts / tcpu, 1), "unit": "points/s", "cores": 1, "kind": "port",
                                 "sample": "%d k-points x %d frequencies of the oracle's G_c (1 - V G_c)^-1 "
                                           "trace" % (len(ks), nw),
                                 "device_points_per_s": round(nw * len(k) / (dev * 1e-3), 1),
                                 "max_rel_difference": agree}
    inst.erase()
    return {"workload": "3x4 Hubbard cluster t=1 U=8 mu=4, sector N=12 S=0 (D=853,776), CPT trace of "
                        "G_c (1 - V(k) G_c)^-1 on a %dx%d k-grid x %d frequencies" % (side, side, nw),
            **extra,
            "E0": round(E0, 10), "points": nw * len(k), "ksum_ms_device": round(dev, 3),
            "ksum_ms_wall": round(float(np.median(walls)) * 1e3, 3), "ed_and_gc_s": round(t_ed, 3),
            "roofline": {"bound": "fp64", "achieved": round(tf, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tf / PEAK_FP64_TFLOPS, 4),
                         "flop_per_point": CPT_FLOP_PER_POINT}}


def cdmft_leg(n, batch, graph=True, jacobian=False, ensemble=1):
    """Plain (cold start) vs learned warm start, batched CDMFT on n seeded points with
    `batch` problems active at a time (a converged or failed problem is replaced by the
    next queued one). With `jacobian`, a third leg also starts each Broyden iteration from
    the inverse Jacobian that the nearest training run reached (the reference's CDMFT
    `alpha` accepts "the inverse Jacobian from a previous run")."""
    from pyqcm_amd import cdmft as C
    from pyqcm_amd import warmstart as W
    X, Y = W.load_training_set()
    nets = [W.MLP(3, 128, 16, seed=s) for s in range(max(1, ensemble))]
    t0 = time.perf_counter()
    hist = [net.fit(X, Y, epochs=1500, batch=64, lr=2e-3) for net in nets][0]
    t_train = time.perf_counter() - t0
    log("MLP trained on %d points in %.1f s (loss %.3g -> %.3g)" % (len(X), t_train, hist[0], hist[-1]))
    P = W.plaquette_problem("plaq_bench")
    P.graph = graph
    pts = W.plaquette_points(n, 20261202)
    feats = np.array([W.plaquette_features(v) for v in pts])
    cold = [W.PLAQUETTE_COLD_START] * n
    t0 = time.perf_counter()
    warm = list(C.canonical_bath(np.mean([net.predict(feats, bf16=True) for net in nets], axis=0)))
    t_pred = time.perf_counter() - t0
    legs = [("plain", cold, None), ("warm", warm, None)]
    FJ, t_pred_j = None, 0.0
    if jacobian:
        FJ, JJ = W.load_jacobians()
        if FJ is None:
            raise SystemExit("no training file holds inverse Jacobians")
        t0 = time.perf_counter()
        alphas = W.nearest_jacobians(feats, FJ, JJ, X.std(0))
        t_pred_j = time.perf_counter() - t0
        legs.append(("warm_jacobian", warm, alphas))
    res = {}
    for name, x0, alpha in legs:
        t0 = time.perf_counter()
        B = C.BatchedCDMFT(P, pts, x0, solver="batched", capacity=batch, alpha_list=alpha)
        dt = time.perf_counter() - t0 + (t_pred if name != "plain" else 0.0) + \
            (t_pred_j if alpha is not None else 0.0)
        good = B.converged & ~B.failed
        res[name] = {"converged": int(good.sum()), "failed": int((~good).sum()), "seconds": round(dt, 2),
                     "solves_per_s": round(good.sum() / dt, 4),
                     "mean_iterations": round(float(B.niter[good].mean()) if good.any() else 0.0, 2),
                     "impurity_solves": int(B.n_solves.sum())}
        res[name + "_x"] = np.array(B.x)
        res[name + "_ok"] = good
        log("CDMFT %s: %s" % (name, res[name]))
    # the plain loop is the fallback: the problems that failed from the learned start run
    # again from the cold start, and the warm leg with that fallback is reported beside it
    bad = np.where(~res["warm_ok"])[0]
    fb = {"retried_from_cold_start": int(len(bad)), "converged": res["warm"]["converged"],
          "seconds": res["warm"]["seconds"]}
    if len(bad):
        t0 = time.perf_counter()
        B2 = C.BatchedCDMFT(P, [pts[i] for i in bad], [cold[i] for i in bad], solver="batched", capacity=batch)
        ok2 = B2.converged & ~B2.failed
        fb["converged"] += int(ok2.sum())
        fb["seconds"] = round(fb["seconds"] + time.perf_counter() - t0, 2)
    fb["solves_per_s"] = round(fb["converged"] / fb["seconds"], 4)
    log("CDMFT warm, then cold for its failures: %s" % fb)
    both = res["plain_ok"] & res["warm_ok"]
    dx = np.abs(C.canonical_bath(res["plain_x"][both]) - C.canonical_bath(res["warm_x"][both]))
    same = dx.max(1) < 1e-2 if both.any() else np.zeros(0, bool)
    # a warm solve counted only where it reaches the plain loop's fixed point (or where the
    # plain loop did not converge)
    agree_warm = int(res["warm_ok"].sum() - (both.sum() - same.sum()))
    res["warm"]["converged_at_plain_fixed_point"] = agree_warm
    res["warm"]["solves_per_s_at_plain_fixed_point"] = round(agree_warm / res["warm"]["seconds"], 4)
    return {
        "workload": "2x2 plaquette + 8 bath orbitals, %d seeded points (%d active at a time), sectors "
                    "N=8..16 S=0, Legendre (1,10,5,10,5) grid, 32x32 k-grid, accur 1e-4 on parameters, "
                    "Broyden maxiter 32" % (n, batch),
        "plain": res["plain"], "warm": res["warm"],
        "speedup_solves_per_s": round(res["warm"]["solves_per_s"] / max(res["plain"]["solves_per_s"], 1e-12), 3),
        "speedup_at_plain_fixed_point": round(res["warm"]["solves_per_s_at_plain_fixed_point"] /
                                              max(res["plain"]["solves_per_s"], 1e-12), 3),
        "warm_with_fallback": fb,
        "speedup_with_fallback": round(fb["solves_per_s"] / max(res["plain"]["solves_per_s"], 1e-12), 3),
        "same_fixed_point": "%d of %d points converged by both" % (int(same.sum()), int(both.sum())),
        "different_fixed_point": int(both.sum() - same.sum()),
        **({"warm_jacobian": res["warm_jacobian"], "jacobian_runs": int(len(FJ)),
            "speedup_jacobian": round(res["warm_jacobian"]["solves_per_s"] /
                                      max(res["plain"]["solves_per_s"], 1e-12), 3)} if jacobian else {}),
        "mlp": {"train_points": int(len(X)), "networks": len(nets), "train_s": round(t_train, 2),
                "final_loss": float(hist[-1]),
                "inference": "bf16 v_mfma_f32_16x16x32_bf16"},
    }


if __name__ == "__main__":
    main()
