"""Benchmark of the ED impurity-solver hot path on one MI355X.

A step is one y <- y + H x over the 4x4 Hubbard cluster (t=1, t'=-0.3, U=8, mu=4) on its
half-filled N=16, Sz=0 sector: D = 12870^2 = 165,636,900 FP64 amplitudes, algorithmic
traffic 24 D bytes (read x, read y, write y). Inputs are resident in HBM before timing.

Prints one JSON line (rank 0, N=1):
  value        whole-job H.v throughput in GB/s of algorithmic bytes
  roofline     achieved GB/s of the H.v kernel(s) from HIP events on the library stream
  cpu_baseline the oracle's H.v (reference loop order, one core) on the same sector
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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the single-GPU H.v path); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from oracle import oracle as O
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import Sector

    _lib.call("qed_init", 0)
    nn, nnn = O.square_bonds(4, 4)
    model, vals = O.hubbard_cluster(16, nn, tp_bonds=nnn, tp=-0.3, U=8.0, mu=4.0)
    flat = model.flat(vals)
    t0 = time.time()
    S = Sector(16, 8, 8, flat)
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
        for v in (0, 1, 4, 5):
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

    cpu = None
    if not args.no_cpu:
        # one full H.v of the same sector through the oracle (reference loop order, 1 core)
        So = O.Sector(model, vals, 8, 8)
        yc = np.zeros(D)
        xh = x
        t0 = time.perf_counter()
        O.lib().or_hv(So.h, O._dp(xh), O._dp(yc))
        tc = time.perf_counter() - t0
        cpu = {"value": round(bytes_per / tc / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": "one full y += H x on the same N=16 Sz=0 sector (D=165,636,900) through the "
                         "oracle's restatement of Hamiltonian_Factorized::mult_add, %.1f s" % tc}
        log("cpu baseline: %.2f s per H.v" % tc)

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
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
