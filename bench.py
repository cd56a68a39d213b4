"""Throughput benchmark of the per-energy NEGF solve on one MI355X.

A step is one batch of energy points of one wire solved end to end on the GPU
(Sancho-Rubio contacts, RGF, Gn/Gp site blocks, T12, modes, density and spin).

Default workload (the configuration the energy-points/s target is stated on):
the 4-band TI wire with Npx=200, Npy=Npz=24 (block n=2304), 1024 energies in
[-0.3, 0.3], Anderson disorder W=0.3 seed 1, eta=1e-6, Sancho-Rubio tol 1e-10
/ max 200; each step takes a strided slice of that grid so every step samples
the whole range.  The timed form is the mixed-precision form (FP64 contacts,
inverses and the T12 chain; complex64 first-column substitution and site-block
products), the fastest form meeting the <= 1e-8 relative transmission error;
the plain FP64 form is timed on the same energies right after, and the mixed
form's error against it is reported for every output (the plain form itself is
pinned to the complex128 CPU oracle by the test suite).

Prints one JSON line.  `roofline` is for the dominant kernel, timed with HIP
events on the stream it runs on (a uniform sample of its launches inside the
timed region); its `traffic` (and the kernel's stand-alone rate, `alone`) come
from the committed rocprofv3 counter summary of that kernel when one covers the
workload.  `cpu_baseline` times the complex128 CPU oracle (MKL) on the host on a
bounded sample of the same wire.

Counter runs: `--gemm-only` runs only the dominant kernel at the shape the sweep
launches it with; `--copy-calibration MIB` runs a plain copy of known size, whose
counter readings scale the kernel's.  `--config c3` / `--config c5` run the
batched and serving workloads through the server interface.
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

CONFIGS = {
    # small, CPU-runnable: 64 energies of a clean 6x6 wire
    "c1": dict(npx=20, npy=6, npz=6, ne=64, emin=-0.3, emax=0.3, W=0.0, seed=1, per_step=64,
               budget_gb=0.0),
    # plain FP64 mid-size: block 1024, Npx=100, W=0.5
    "c2": dict(npx=100, npy=16, npz=16, ne=512, emin=-0.4, emax=0.4, W=0.5, seed=1, per_step=64,
               budget_gb=0.0),
    # precision target size: block 2304, Npx=200, W=0.3.  245 GB of energies in flight
    # (the library caps it at 90% of free HBM) holds two energies per lane on four lanes
    "c4": dict(npx=200, npy=24, npz=24, ne=1024, emin=-0.3, emax=0.3, W=0.3, seed=1, per_step=8,
               budget_gb=245.0),
}
# Many-wire workloads run through the serving / batching layer (negf_server):
#  c3 -- a seeded sample of the batching configuration (256 disorder realisations x
#        32 energies, Npy=Npz in {8,12,16,20}, Npx in [40,120], W in [0,1], g*B in
#        [0,0.05], flux in [0,2]): `instances` realisations x `energies_per` of their 32
#        energies, all submitted at once;
#  c5 -- a serving stream with Poisson arrivals at `rate` requests/s, per request
#        Npy=Npz in `ns`, Npx in `L`, 16..`ne_max` energies in [-0.4,0.4], flux in
#        [0,2], W in [0,1], own seed; scaled down from the configuration's
#        Npy=Npz in {8..24}, Npx in [50,300], 16..256 energies so that the stream
#        fits about a minute of one GPU (the full distribution's mean request is
#        ~75 s at FP64 peak).
SERVER_WORKLOADS = {
    "c3": dict(instances=16, energies_per=8, ns=(8, 12, 16, 20), L=(40, 120), seed=3000),
    "c5": dict(requests=20, rate=0.3, ns=(8, 12, 16), L=(50, 150), ne_max=64, seed=5000),
}
ETA, SR_TOL, SR_MAX = 1e-6, 1e-10, 200
MU, MUT = (-0.1, 0.0), 0.0
LANES = [0]  # set from --lanes
OUTPUTS = ("T12", "Modec1", "Modec2", "DOSc1", "DOSc2", "EDensity", "HDensity", "RhoN", "RhoP",
           "SigXExp", "SigYExp", "SigZExp")


def f_useful(n, L, k_left, k_right):
    """A fixed yardstick of dense flops per energy point, the three-pass RGF's
    count: 2L inverses, 2(L-1) column GEMMs, 2L G*Gamma GEMMs, 1 for T12 and
    (2 + 9K) per lead for Sancho-Rubio, each n^3 complex multiply-adds = 8 n^3
    flops.  The solve does less than this (L inverses, no G*Gamma products: the
    column chains carry the factor of Gamma), so the rate it gives rises with
    such savings."""
    return 8.0 * n**3 * (6 * L + 3 + 9 * (k_left + k_right))


def cpu_baseline(cfg, energy, threads, k_mean):
    """The complex128 CPU oracle (MKL, the reference's operation order) timed on the
    host on a bounded sample of one energy of the same wire, extrapolated to the
    workload's Npx.  The time of one energy is 2 x (Sancho-Rubio of one lead) +
    Npx x (per-slice RGF, Gn/Gp and observables).  For blocks up to 1024 a whole
    energy at Npx = 2 and 4 is timed and the line through them extrapolated; for
    larger blocks the pieces are timed: one lead's decimation at 1 and 3
    iterations (per-iteration and fixed cost, scaled to the GPU run's mean
    iteration count k_mean) and the RGF/Gnp/observables stages at Npx = 2."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402  (checker only)
    O.set_threads(threads)
    n = cfg["npy"] * cfg["npz"] * 4

    def wire(L):
        return build_wire(L, cfg["npy"], cfg["npz"], disorder=cfg["W"], seed=cfg["seed"])

    if n <= 1024:
        t = {}
        for L in (2, 4):
            w = wire(L)
            t0 = time.perf_counter()
            O.energy_sweep(w, energy, 0.0, MU, MUT, ETA, SR_TOL, SR_MAX, True)
            t[L] = time.perf_counter() - t0
        per_slice = max(t[4] - t[2], 0.0) / 2.0
        t_full = t[2] + per_slice * (cfg["npx"] - 2)
        sample = (f"oracle, 1 energy (E={energy:.4f}), Npy=Npz={cfg['npy']}, MKL {threads} threads, "
                  f"at Npx=2 ({t[2]:.1f}s) and 4 ({t[4]:.1f}s); extrapolated affinely to "
                  f"Npx={cfg['npx']}: {t_full:.1f}s per energy")
        return dict(value=1.0 / t_full, unit="energy-points/s", cores=threads, kind="port",
                    sample=sample)
    w = wire(2)
    alpha = (energy + 1j * ETA) * np.eye(n) - w.htotd[:, :, 0]
    beta = -w.hopx.conj().T
    ts = {}
    for it in (1, 3):
        t0 = time.perf_counter()
        O.sancho_rubio(alpha, beta, ETA, SR_TOL, it, True)
        ts[it] = time.perf_counter() - t0
    per_it = max(ts[3] - ts[1], 0.0) / 2.0
    fixed = max(ts[1] - per_it, 0.0)
    c = O.contact_self_energy(w, energy, ETA, SR_TOL, 1, True)  # Sigma values do not change the cost
    t0 = time.perf_counter()
    A = O.assemble_invg(w, c["SigmaL"], c["SigmaR"], energy, ETA)
    r = O.retarded(A, w.hopx)
    g = O.gnp(w, 1.0, 0.0, 0.01, r["GRLef"], r["GRRgt"], c["GammaL"], c["GammaR"])
    O.observables(w, 1.0, 0.0, 0.01, g, r["GRRgt"], c["GammaL"], c["GammaR"])
    t_st = time.perf_counter() - t0
    t_full = 2.0 * (fixed + per_it * k_mean) + t_st / 2.0 * cfg["npx"]
    sample = (f"oracle pieces at Npy=Npz={cfg['npy']} (n={n}), E={energy:.4f}, MKL {threads} threads: "
              f"one lead's Sancho-Rubio at 1 and 3 iterations ({ts[1]:.1f}s, {ts[3]:.1f}s -> "
              f"{per_it:.2f}s/iteration), RGF+Gnp+observables at Npx=2 ({t_st:.1f}s); "
              f"{ts[1] + ts[3] + t_st:.1f}s of CPU work; one energy = 2 leads x ({fixed:.2f}s + "
              f"{k_mean:.1f} GPU-measured iterations) + {cfg['npx']} slices = {t_full:.0f}s")
    return dict(value=1.0 / t_full, unit="energy-points/s", cores=threads, kind="port", sample=sample)


# Committed summaries per (workload, kernel kind) at the shape the bench launches
# the kernel with: `alone_tflops` (the kernel alone, bench.py --gemm-only) and,
# where counters were collected, `traffic_bytes_per_launch` (rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter's reading for
# a copy kernel of known size).
TRAFFIC_PROFILES = {("c4", "fp32_gemm_3m"): "traffic_c64_gemm3m_n2304_L200.md",
                    ("c4", "fp64_gemm_3m"): "alone_fp64_gemm3m_n2304_b2.md"}


def profile_value(config, kind, key):
    """A `key: value` line of the dominant kernel's committed counter summary, or
    None when no summary covers this workload and form."""
    name = TRAFFIC_PROFILES.get((config, kind))
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    return None


def traffic_from_profiles(config, kind):
    """HBM bytes per launch of the dominant kernel (FETCH_SIZE + WRITE_SIZE, scaled)."""
    return profile_value(config, kind, "traffic_bytes_per_launch")


def gemm_only(args):
    """The dominant kernel alone: the 3M complex product (complex64 in the mixed form,
    complex128 in the plain form) with the block size of the workload and the batch
    one lane launches it with (the energies in flight per lane)."""
    cfg = CONFIGS[args.config]
    n = cfg["npy"] * cfg["npz"] * 4
    per = args.per_step or cfg["per_step"]
    lanes = args.lanes or 4
    batch = max(1, per // lanes)
    reps = 8
    ms = capi.bench_gemm(n, batch, complex64=(args.precision == "mixed"), m3=True, reps=reps)
    flops = 8.0 * n ** 3 * batch
    print(json.dumps({"kernel": "k_gemm<%s,0,0,0,2,2,0,1>" % ("float" if args.precision == "mixed" else "double"),
                      "n": n, "batch": batch, "launches": reps + 1, "ms_per_launch": ms,
                      "algorithmic_flops_per_launch": flops, "tflops": flops / ms / 1e9,
                      "algorithmic_bytes_per_launch": 3 * batch * n * n * (8 if args.precision == "mixed" else 16)}))


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


def run_form(w, cfg, precision, flags, per, grid, dE, step_indices, steps, warmup, profile_every,
             label):
    """Warm up, then time `steps` steps of one form; returns timings, the
    accumulators of the last timed step and the GEMM profile."""
    L = cfg["npx"]
    prec = capi.NEGF_PREC_MIXED if precision == "mixed" else capi.NEGF_PREC_FP64
    ctx = capi.Context(L, cfg["npy"], cfg["npz"], 4, eta=ETA, sr_tol=SR_TOL, sr_max_iter=SR_MAX,
                       max_batch=per, flags=flags, precision=prec, hbm_budget_gb=cfg["budget_gb"],
                       lanes=LANES[0])
    ctx.set_hamiltonian(w)

    def run_step(k, acc):
        idx = step_indices(k)
        ctx.submit(grid[idx], idx, np.full(len(idx), dE))
        return ctx.flush(MU, MUT, acc)

    for k in range(warmup):
        t0 = time.perf_counter()
        run_step(k, capi.Accumulators(L, cfg["npy"], cfg["npz"], cfg["ne"]))
        print(f"[bench] {label} warmup step {k} {time.perf_counter() - t0:.2f}s", file=sys.stderr,
              flush=True)
    capi.profile_begin(profile_every)
    times, iters, status = [], [], []
    acc = None
    t_start = time.perf_counter()
    for k in range(warmup, warmup + steps):
        acc = capi.Accumulators(L, cfg["npy"], cfg["npz"], cfg["ne"])
        t0 = time.perf_counter()
        st, it = run_step(k, acc)  # flush() returns after the device streams have drained
        times.append(time.perf_counter() - t0)
        iters.append(it)
        status.append(st)
        print(f"[bench] {label} step {k} {times[-1]:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    prof = capi.profile_end()
    ctx.close()
    return dict(times=times, elapsed=elapsed, iters=np.concatenate(iters), status=status, acc=acc,
                prof=prof, last_step=warmup + steps - 1)


def errors_vs(acc, ref):
    """Max error of every output relative to its natural scale (T12 and the modes:
    the open-channel count, at least 1; spin: the density; others: their max)."""
    out = {}
    nm = max(1.0, float(np.abs(ref.Modec1).max()))
    rho = float(np.abs(ref.RhoN).max())
    for k in OUTPUTS:
        a, b = np.asarray(getattr(acc, k)), np.asarray(getattr(ref, k))
        scale = max(float(np.abs(b).max()), 1e-300)
        if k in ("T12", "Modec1", "Modec2"):
            scale = nm
        if k.startswith("Sig"):
            scale = max(scale, rho)
        out[k] = float(np.abs(a - b).max()) / scale
    return out


def server_workload(name, args):
    """Run a many-wire workload through capi.Server; returns the JSON line."""
    import threading
    wl = SERVER_WORKLOADS[name]
    rng = np.random.default_rng(wl["seed"])
    reqs = []  # (arrival_s, wire, energies)
    if name == "c3":
        grid = energy_grid(-0.4, 0.4, 32)
        for r in range(wl["instances"]):
            ny = int(rng.choice(wl["ns"]))
            L = int(rng.integers(wl["L"][0], wl["L"][1] + 1))
            w = build_wire(L, ny, ny, disorder=float(rng.uniform(0, 1)), seed=wl["seed"] + r,
                           flux=float(rng.uniform(0, 2)), zeeman=float(rng.uniform(0, 0.05)), dense=False)
            sel = grid[(np.arange(wl["energies_per"]) * (32 // wl["energies_per"]) + r) % 32]
            reqs.append((0.0, w, sel))
        desc = (f"c3 sample: {wl['instances']} of the 256 seeded realisations x {wl['energies_per']} of "
                f"their 32 energies in [-0.4,0.4]; Npy=Npz in {wl['ns']}, Npx in {wl['L']}, W in [0,1], "
                f"g*B in [0,0.05], flux in [0,2]; all submitted at t=0 as compact Hamiltonians "
                f"(slice template + site potentials)")
    else:
        t = 0.0
        for r in range(wl["requests"]):
            t += float(rng.exponential(1.0 / wl["rate"]))
            ny = int(rng.choice(wl["ns"]))
            L = int(rng.integers(wl["L"][0], wl["L"][1] + 1))
            ne = int(rng.integers(16, wl["ne_max"] + 1))
            w = build_wire(L, ny, ny, disorder=float(rng.uniform(0, 1)), seed=wl["seed"] + r,
                           flux=float(rng.uniform(0, 2)), dense=False)
            reqs.append((t, w, energy_grid(-0.4, 0.4, ne)))
        desc = (f"c5 stream: {wl['requests']} requests, Poisson arrivals at {wl['rate']}/s; per request "
                f"Npy=Npz in {wl['ns']}, Npx in {wl['L']}, 16..{wl['ne_max']} energies in [-0.4,0.4], "
                f"flux in [0,2], W in [0,1], own seed (scaled down from Npy=Npz 8..24, Npx 50..300, "
                f"16..256 energies); compact Hamiltonians (slice template + site potentials)")
    prec = capi.NEGF_PREC_MIXED if args.precision == "mixed" else capi.NEGF_PREC_FP64
    npts = sum(len(e) for _, _, e in reqs)
    print(f"[bench] {name}: {len(reqs)} requests, {npts} energy points", file=sys.stderr, flush=True)
    srv = capi.Server(eta=ETA, sr_tol=SR_TOL, sr_max_iter=SR_MAX, precision=prec, flags=args.flags,
                      lanes=args.lanes)
    ids = [None] * len(reqs)
    capi.profile_begin(args.profile_every)
    t0 = time.perf_counter()

    def submitter():
        for i, (at, w, e) in enumerate(reqs):
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit(w, e, e[1] - e[0] if len(e) > 1 else 0.0, MU, MUT, compact=True)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lat, stat = [], []
    for i, (rid, acc) in enumerate(ids):
        st, l = srv.wait(rid)
        lat.append(l)
        stat.append(st)
        print(f"[bench] request {i} n={reqs[i][1].n} Npx={reqs[i][1].npx} energies={len(reqs[i][2])} "
              f"latency {l / 1e3:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    prof = capi.profile_end()
    srv.close()
    dominant = "fp32_gemm_3m" if args.precision == "mixed" else "fp64_gemm_3m"
    peak = PEAK_FP32_MFMA if args.precision == "mixed" else PEAK_FP64_MFMA
    g_flops, g_ms, g_launches = prof[dominant]
    gemm_tflops = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    return {
        "metric": "energy-points/s on 1 GPU vs plain FP64 form at stated error; p50/p99 solve latency",
        "value": npts / elapsed,
        "unit": "energy-points/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * elapsed,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+c64 (mixed)" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded 4-band TI wires, host-built Hamiltonians)",
        "config": {"workload": desc, "requests": len(reqs), "energy_points": npts,
                   "precision": args.precision},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "note": "per request, submission to completion (queueing included)"},
        "status_nonzero": int(sum(1 for s in stat if s)),
        "roofline": {"bound": "mfma", "kernel": f"k_gemm 3M complex product ({dominant})",
                     "achieved": gemm_tflops, "peak": peak, "unit": "TFLOP/s", "frac": gemm_tflops / peak,
                     "traffic": None, "launches": g_launches, "avg_launch_ms": g_ms / max(1, g_launches)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS) + sorted(SERVER_WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=("fp64", "mixed"))
    ap.add_argument("--lanes", type=int, default=0, help="concurrent device lanes (0 = library default)")
    ap.add_argument("--per-step", type=int, default=0, help="energies per step (default per config)")
    ap.add_argument("--budget-gb", type=float, default=-1.0,
                    help="HBM for energies in flight (default per config; 0 = library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain-FP64 comparison step of the mixed form")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--flags", type=int, default=0, help="NEGF_FLAG_* algorithm selection")
    ap.add_argument("--profile-every", type=int, default=32,
                    help="time one MFMA GEMM launch in this many with HIP events")
    ap.add_argument("--gemm-only", action="store_true",
                    help="run only the dominant kernel (the 3M complex product) at the shape the "
                         "sweep launches it with, for rocprofv3 --pmc counter runs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB 8 times (the known-size "
                         "reference for rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gemm_only:
        gemm_only(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: this benchmark runs on exactly one GPU "
                                   "(energies shard across GPUs with no communication until the "
                                   "final sums; only --gpus 1 is measured)"}))
        sys.exit(2)

    LANES[0] = args.lanes
    if args.config in SERVER_WORKLOADS:
        print(json.dumps(server_workload(args.config, args)), flush=True)
        return
    cfg = dict(CONFIGS[args.config])
    if args.budget_gb >= 0:
        cfg["budget_gb"] = args.budget_gb
    per = args.per_step or cfg["per_step"]
    n = cfg["npy"] * cfg["npz"] * 4
    L = cfg["npx"]
    grid = energy_grid(cfg["emin"], cfg["emax"], cfg["ne"])
    dE = grid[1] - grid[0]
    nsteps = args.warmup + args.steps
    stride = max(1, cfg["ne"] // per)
    shift = max(1, stride // max(1, nsteps))

    def step_indices(k):
        return (np.arange(per) * stride + k * shift) % cfg["ne"]

    print(f"[bench] building wire {args.config}: Npx={L} Npy=Npz={cfg['npy']} n={n}", file=sys.stderr,
          flush=True)
    w = build_wire(L, cfg["npy"], cfg["npz"], disorder=cfg["W"], seed=cfg["seed"])
    run = run_form(w, cfg, args.precision, args.flags, per, grid, dE, step_indices, args.steps,
                   args.warmup, args.profile_every, args.precision)
    npts = per * args.steps
    elapsed = run["elapsed"]
    value = npts / elapsed
    compare = None
    if args.precision == "mixed" and not args.no_compare:
        # the plain FP64 form on the same energies as the last timed step
        last = run["last_step"]
        base = run_form(w, cfg, "fp64", args.flags, per, grid, dE, lambda k: step_indices(last), 1, 0,
                        args.profile_every, "fp64")
        plain_rate = per / base["elapsed"]
        compare = {
            "plain_fp64_energy_points_per_s": plain_rate,
            "speedup_vs_plain_fp64": value / plain_rate,
            "error_vs_plain_fp64": errors_vs(run["acc"], base["acc"]),
            "note": ("errors: max |mixed - fp64| over the last timed step's energies / the output's "
                     "natural scale; the plain FP64 form agrees with the complex128 CPU oracle to "
                     "1e-10 in the test suite"),
        }

    iters = run["iters"]
    useful = sum(f_useful(n, L, a, b) for a, b in iters)
    prof = run["prof"]
    dominant = "fp32_gemm_3m" if args.precision == "mixed" else "fp64_gemm_3m"
    peak = PEAK_FP32_MFMA if args.precision == "mixed" else PEAK_FP64_MFMA
    g_flops, g_ms, g_launches = prof[dominant]
    gemm_tflops = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    kernels = {k: {"tflops": (f / (m * 1e-3) / 1e12 if m > 0 else 0.0), "launches": c,
                   "avg_launch_ms": m / max(1, c)}
               for k, (f, m, c) in prof.items() if c}
    for k, v in kernels.items():  # the kernel alone at the bench's shape, where measured
        a = profile_value(args.config, k, "alone_tflops")
        if a is not None:
            v["alone_tflops"] = a
    kname = ("k_gemm<float,0,0,0,2,2,0,1> (complex64 MFMA product, 3M form, n x n x n per matrix)"
             if args.precision == "mixed" else
             "k_gemm<double,0,0,0,2,2,0,1> (FP64 MFMA complex product, 3M form, n x n x n per matrix)")
    result = {
        "metric": "energy-points/s on 1 GPU vs plain FP64 form at stated error; p50/p99 solve latency",
        "value": value,
        "unit": "energy-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+c64 (mixed)" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded 4-band TI wire, host-built Hamiltonian)",
        "config": {"workload": (f"{args.config}: Npx={L}, Npy=Npz={cfg['npy']} (block {n}), "
                                f"{cfg['ne']} energies in [{cfg['emin']},{cfg['emax']}], "
                                f"Anderson W={cfg['W']} seed {cfg['seed']}, eta={ETA}, "
                                f"SR tol {SR_TOL} max {SR_MAX}; {args.precision} form"),
                   "energies_per_step": per, "precision": args.precision, "flags": args.flags},
        "latency_ms": {"p50": 1e3 * float(np.percentile(run["times"], 50)),
                       "p99": 1e3 * float(np.percentile(run["times"], 99)),
                       "note": ("a step's energies complete together: per-energy solve latency = "
                                "step time")},
        "sr_iters_mean": float(iters.mean()),
        "status_nonzero": int(sum(int(np.count_nonzero(s)) for s in run["status"])),
        "path_tflops_useful": useful / elapsed / 1e12,
        "path_frac_fp64_peak": useful / elapsed / 1e12 / PEAK_FP64_MFMA,
        "roofline": {"bound": "mfma", "kernel": kname,
                     "achieved": gemm_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": gemm_tflops / peak, "traffic": traffic_from_profiles(args.config, dominant),
                     "launches": g_launches, "avg_launch_ms": g_ms / max(1, g_launches),
                     "note": ("achieved = algorithmic 8 n^3 flops per complex product x batch / HIP-event "
                              "duration on the launching stream (one launch in --profile-every sampled); the "
                              "3M form executes 6 n^3 on the matrix cores; the lanes run concurrently, so "
                              "durations include the other lanes' kernels sharing the CUs; traffic: HBM-side "
                              "bytes per launch from FETCH_SIZE/WRITE_SIZE counter runs of the kernel alone "
                              "(bench.py --gemm-only), scaled by a copy-kernel calibration")},
        "kernels": kernels,
    }
    alone = profile_value(args.config, dominant, "alone_tflops")
    if alone is not None:
        # the same kernel and shape with the GPU to itself (bench.py --gemm-only)
        result["roofline"]["alone"] = {"achieved": alone, "frac": alone / peak,
                                       "source": "profiles/" + TRAFFIC_PROFILES[(args.config, dominant)]}
    if compare is not None:
        result["compare"] = compare
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(cfg, float(grid[cfg["ne"] // 3]), threads,
                                              float(iters.mean()))
        print(f"[bench] cpu baseline {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
