"""Autotuner of the solve's launch shapes on the GPU it runs on.

Writes nonequilibriumgreensfunction_amd/tuning/table.json (or --out), which the
Python layer installs in the library at load time (capi._load_tuning):

  gemm_tile_fp64, gemm_tile_c64 -- rules [n_max, tile] for the batched complex
      product kernel (the RGF chains, the decimation and the Gauss-Jordan trailing
      updates): tile 0 = 64x64 workgroup tiles, 1 = 32x32.  For each block size of
      the benchmark workloads (144 ... 2304) both tiles are timed alone
      (negf_bench_gemm_tile, 3M form) at the batch sizes one lane launches there,
      and the faster one (summed over the batches) wins.
  lanes -- rules [n_max, lanes]: for each batching cross-section (blocks 256 ...
      1600) a short sweep of a seeded wire (Npx = 24, 8 energies, mixed form) is
      timed with 1, 2 and 4 concurrent lanes, with the tile rules above installed.

Consecutive block sizes with the same choice merge into one rule; block sizes
above the last rule keep the library's defaults (64x64 tiles, 4 lanes).  Run on
the GPU:  python -m nonequilibriumgreensfunction_amd.autotune [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi
from .hamiltonian import build_wire, energy_grid

TILE_BLOCKS = (144, 256, 576, 1024, 1600, 2304)
LANE_BLOCKS = ((8, 256), (12, 576), (16, 1024), (20, 1600))  # (Npy = Npz, block)


def batches_for(n):
    """Batch sizes a lane launches products with at block n (energies per lane)."""
    return (4, 16) if n <= 576 else (2, 8)


def rules_from(choices):
    """[(n, value)] in increasing n -> merged [n_max, value] rules."""
    rules = []
    for n, v in choices:
        if rules and rules[-1][1] == v:
            rules[-1][0] = n
        else:
            rules.append([n, v])
    return rules


def tune_tiles(log):
    out = {}
    for c64, key in ((False, "gemm_tile_fp64"), (True, "gemm_tile_c64")):
        choices = []
        for n in TILE_BLOCKS:
            t = [0.0, 0.0]
            for tile in (0, 1):
                for b in batches_for(n):
                    t[tile] += capi.bench_gemm_tile(n, b, tile, complex64=c64, m3=True, reps=6)
            best = int(np.argmin(t))
            choices.append((n, best))
            log(f"{key} n={n}: 64x64 {t[0]:.3f} ms, 32x32 {t[1]:.3f} ms -> tile {best}")
        rules = rules_from(choices)
        while rules and rules[-1][1] == 0:  # trailing 64x64 rules are the default
            rules.pop()
        out[key] = rules
    return out


def tune_lanes(table, log):
    capi._load_tuning(capi.lib(), _write_tmp(table))
    choices = []
    for ny, n in LANE_BLOCKS:
        w = build_wire(24, ny, ny, disorder=0.5, seed=7, flux=0.4, zeeman=0.02)
        e = energy_grid(-0.4, 0.4, 8)
        t = {}
        for lanes in (1, 2, 4):
            with capi.Context(w.npx, w.npy, w.npz, 4, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                              precision=capi.NEGF_PREC_MIXED, lanes=lanes) as ctx:
                ctx.set_hamiltonian(w)
                for rep in range(2):  # the first sweep allocates the arena
                    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(e))
                    t0 = time.perf_counter()
                    ctx.submit(e, np.arange(len(e)), np.full(len(e), e[1] - e[0]))
                    ctx.flush((-0.1, 0.0), 0.0, acc)
                    t[lanes] = time.perf_counter() - t0
        best = min(t, key=t.get)
        choices.append((n, best))
        log(f"lanes n={n}: " + ", ".join(f"{k}: {v:.3f}s" for k, v in t.items()) + f" -> {best}")
    rules = rules_from(choices)
    while rules and rules[-1][1] == 4:
        rules.pop()
    return rules


def _write_tmp(table):
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(table, f)
    return path


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=capi.TUNING_TABLE)
    ap.add_argument("--no-lanes", action="store_true")
    args = ap.parse_args()
    if capi.device_count() < 1:
        raise SystemExit("autotune needs a GPU")

    def log(msg):
        print(f"[autotune] {msg}", flush=True)

    table = tune_tiles(log)
    table["lanes"] = [] if args.no_lanes else tune_lanes(table, log)
    table["note"] = ("written by nonequilibriumgreensfunction_amd.autotune: gemm_tile rules "
                     "[n_max, tile] (0 = 64x64, 1 = 32x32), lanes rules [n_max, lanes]")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"wrote {args.out}: {json.dumps({k: v for k, v in table.items() if k != 'note'})}")


if __name__ == "__main__":
    sys.exit(main())
