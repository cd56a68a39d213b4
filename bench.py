"""Throughput benchmark of the per-energy NEGF solve on one MI355X.

A step is one batch of energy points of one wire solved end to end on the GPU
(Sancho-Rubio contacts, forward/backward RGF, Gn/Gp site blocks, T12, modes,
density and spin).  Default workload: the plain-FP64 mid-size wire (4-band TI
model, Npx=100, Npy=Npz=16 -> block n=1024, 512 energies in [-0.4, 0.4], Anderson
disorder W=0.5, seed 1); each step takes a strided slice of that energy grid so
every step samples the whole range.

Prints one JSON line.  `roofline` is for the dominant kernel (the FP64 MFMA
complex GEMM, trailing Gauss-Jordan updates included), timed with HIP events on
the stream it runs on; `cpu_baseline` times the complex128 CPU oracle (MKL) on
the host on a bounded sample of the same wire, extrapolated to Npx=100.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nonequilibriumgreensfunction_amd import capi  # noqa: E402
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire, energy_grid  # noqa: E402

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32_MFMA = 157.3
# dominant kernel: the FP64 complex product in the 3M form (RGF column, substitution
# and G*Gamma products), k_gemm<double, N, N, plain epilogue, 64x64, 3M>
DOMINANT = "fp64_gemm_3m"

CONFIGS = {
    "c1": dict(npx=20, npy=6, npz=6, ne=64, emin=-0.3, emax=0.3, W=0.0, seed=1, per_step=64),
    "c2": dict(npx=100, npy=16, npz=16, ne=512, emin=-0.4, emax=0.4, W=0.5, seed=1, per_step=64),
    "c4": dict(npx=200, npy=24, npz=24, ne=1024, emin=-0.3, emax=0.3, W=0.3, seed=1, per_step=16),
}
ETA, SR_TOL, SR_MAX = 1e-6, 1e-10, 200
MU, MUT = (-0.1, 0.0), 0.0


def f_useful(n, L, k_left, k_right):
    """Irreducible dense flops of one energy point: 2L inverses, 2(L-1) column
    GEMMs, 2L G*Gamma GEMMs, 1 for T12 and (2 + 9K) per lead for Sancho-Rubio,
    each n^3 complex multiply-adds = 8 n^3 flops."""
    return 8.0 * n**3 * (6 * L + 3 + 9 * (k_left + k_right))


def cpu_baseline(cfg, energy, threads, budget_s=20.0):
    """Oracle (C++ complex128 + MKL, the reference's operation order) on the host,
    on a bounded sample: one energy of the same cross-section at two short lengths
    Npx = a < b (the time of one energy is affine in Npx: Sancho-Rubio plus L
    slices), extrapolated to the workload's Npx.  b grows until the sample has
    used about budget_s seconds of CPU time, or reaches the workload's Npx."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402  (checker only)
    O.set_threads(threads)

    def timed(L):
        w = build_wire(L, cfg["npy"], cfg["npz"], disorder=cfg["W"], seed=cfg["seed"])
        t0 = time.perf_counter()
        O.energy_sweep(w, energy, 0.0, MU, MUT, ETA, SR_TOL, SR_MAX, True)
        return time.perf_counter() - t0

    a = 2
    t = {a: timed(a)}
    b = a
    used = t[a]
    while b < cfg["npx"] and used < 0.5 * budget_s:
        b = min(cfg["npx"], max(b + 2, 2 * b))
        t[b] = timed(b)
        used += t[b]
    if b == cfg["npx"]:
        t_full = t[b]
    else:
        per_slice = max(t[b] - t[a], 0.0) / (b - a)
        t_full = t[a] + per_slice * (cfg["npx"] - a)
    lens = sorted(t)
    return dict(value=1.0 / t_full, unit="energy-points/s", cores=threads, kind="port",
                sample=(f"oracle, 1 energy (E={energy:.4f}), Npy=Npz={cfg['npy']}, MKL {threads} "
                        f"threads, at Npx=" + ", ".join(f"{L} ({t[L]:.1f}s)" for L in lens) +
                        f"; {used:.1f}s of CPU work, extrapolated affinely to Npx={cfg['npx']}: "
                        f"{t_full:.1f}s per energy"))


def traffic_from_profiles(config):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    --pmc summary (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy-kernel
    calibration), if one exists for this workload; else None."""
    path = os.path.join(ROOT, "profiles", f"k_gemm_fp64_traffic_{config}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def copy_calibration(mib):
    """A plain copy of known size: reads and writes mib MiB per launch."""
    import torch
    x = torch.ones(mib * 2**20 // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(8):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read_per_launch": mib * 2**20,
                      "copy_calibration_bytes_written_per_launch": mib * 2**20, "launches": 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--per-step", type=int, default=0, help="energies per step (default per config)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--flags", type=int, default=0, help="NEGF_FLAG_* algorithm selection")
    ap.add_argument("--profile-every", type=int, default=32,
                    help="time one MFMA GEMM launch in this many with HIP events")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB 8 times (the known-size "
                         "reference for rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: this benchmark runs on exactly one GPU "
                                   "(energies shard across GPUs with no communication until the "
                                   "final sums; only --gpus 1 is measured)"}))
        sys.exit(2)

    cfg = CONFIGS[args.config]
    per = args.per_step or cfg["per_step"]
    n = cfg["npy"] * cfg["npz"] * 4
    L = cfg["npx"]
    grid = energy_grid(cfg["emin"], cfg["emax"], cfg["ne"])
    dE = grid[1] - grid[0]
    nsteps = args.warmup + args.steps
    stride = max(1, min(nsteps, cfg["ne"] // per))

    def step_indices(k):
        idx = (np.arange(per) * stride + k) % cfg["ne"]
        return idx

    print(f"[bench] building wire {args.config}: Npx={L} Npy=Npz={cfg['npy']} n={n}", file=sys.stderr,
          flush=True)
    w = build_wire(L, cfg["npy"], cfg["npz"], disorder=cfg["W"], seed=cfg["seed"])
    ctx = capi.Context(L, cfg["npy"], cfg["npz"], 4, eta=ETA, sr_tol=SR_TOL, sr_max_iter=SR_MAX,
                       max_batch=per, flags=args.flags)
    ctx.set_hamiltonian(w)
    acc = capi.Accumulators(L, cfg["npy"], cfg["npz"], cfg["ne"])

    def run_step(k):
        idx = step_indices(k)
        ctx.submit(grid[idx], idx, np.full(len(idx), dE))
        st, it = ctx.flush(MU, MUT, acc)
        return st, it

    for k in range(args.warmup):
        t0 = time.perf_counter()
        run_step(k)
        print(f"[bench] warmup step {k} {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)

    capi.profile_begin(args.profile_every)
    times, iters, status = [], [], []
    t_start = time.perf_counter()
    for k in range(args.warmup, nsteps):
        t0 = time.perf_counter()
        st, it = run_step(k)  # flush() returns after the device stream has drained
        times.append(time.perf_counter() - t0)
        iters.append(it)
        status.append(st)
        print(f"[bench] step {k} {times[-1]:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    prof = capi.profile_end()
    ctx.close()

    iters = np.concatenate(iters)
    npts = per * args.steps
    value = npts / elapsed
    useful = sum(f_useful(n, L, a, b) for a, b in iters)
    g_flops, g_ms, g_launches = prof[DOMINANT]
    gemm_tflops = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    kernels = {k: {"tflops": (f / (m * 1e-3) / 1e12 if m > 0 else 0.0), "launches": c,
                   "avg_launch_ms": m / max(1, c), "busy_ms": m}
               for k, (f, m, c) in prof.items() if c}
    result = {
        "metric": "energy-points/s (plain FP64 form, one GPU)",
        "value": value,
        "unit": "energy-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 4-band TI wire, host-built Hamiltonian)",
        "config": {"workload": (f"{args.config}: Npx={L}, Npy=Npz={cfg['npy']} (block {n}), "
                                f"{cfg['ne']} energies in [{cfg['emin']},{cfg['emax']}], "
                                f"Anderson W={cfg['W']} seed {cfg['seed']}, eta={ETA}, "
                                f"SR tol {SR_TOL} max {SR_MAX}"),
                   "energies_per_step": per, "precision": "fp64"},
        "latency_ms": {"p50": 1e3 * float(np.percentile(times, 50)),
                       "p99": 1e3 * float(np.percentile(times, 99)),
                       "note": "energies of a step complete together: per-energy solve latency = step time"},
        "sr_iters_mean": float(iters.mean()),
        "status_nonzero": int(sum(int(np.count_nonzero(s)) for s in status)),
        "path_tflops_useful": useful / elapsed / 1e12,
        "path_frac_fp64_peak": useful / elapsed / 1e12 / PEAK_FP64_MFMA,
        "roofline": {"bound": "mfma",
                     "kernel": ("k_gemm<double,0,0,0,2,2,0,1> (FP64 MFMA complex product, 3M form, "
                                "n x n x n per matrix)"),
                     "achieved": gemm_tflops, "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                     "frac": gemm_tflops / PEAK_FP64_MFMA, "traffic": traffic_from_profiles(args.config),
                     "launches": g_launches, "avg_launch_ms": g_ms / max(1, g_launches),
                     "note": ("achieved = algorithmic 8 n^3 flops per complex product x batch / HIP-event "
                              "duration on the launching stream (one launch in --profile-every sampled); the "
                              "3M form executes 6 n^3 on the matrix cores; two lanes run concurrently, so "
                              "durations include the other lane's kernels sharing the CUs")},
        "kernels": kernels,
        "flags": args.flags,
    }
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(cfg, float(grid[cfg["ne"] // 3]), threads)
        print(f"[bench] cpu baseline {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
