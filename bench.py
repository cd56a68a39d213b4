"""Throughput benchmark of the per-energy NEGF solve on one MI355X.

A step is one batch of energy points of one wire solved end to end on the GPU
(Sancho-Rubio contacts, RGF, Gn/Gp site blocks, T12, modes, density and spin).

Default workload (the configuration the energy-points/s target is stated on):
the 4-band TI wire with Npx=200, Npy=Npz=24 (block n=2304), 1024 energies in
[-0.3, 0.3], Anderson disorder W=0.3 seed 1, eta=1e-6, Sancho-Rubio tol 1e-10
/ max 200; each step takes a strided slice of that grid so every step samples
the whole range.  The timed form is the complex64 block-LU form (lu32: FP64
contacts; block inverses, both column chains and site blocks in complex64; T12
from an FP64-refined solve of the block-tridiagonal system on the leading
columns of the right lead's broadening factor), the fastest form meeting the
<= 1e-8 relative transmission error; the plain FP64 form is timed on the same
energies right after, and the timed form's error against it is reported for
every output (the plain form itself is pinned to the complex128 CPU oracle by
the test suite).

Prints one JSON line.  `roofline` is for the dominant kernel -- the MFMA kernel
kind with the largest summed duration among the sampled launches (the FP64 3M
complex product: last-column chain and decimation) -- timed with HIP events on
the stream it runs on (a uniform sample of its launches inside the timed
region); its `traffic` (and the kernel's stand-alone rate, `alone`) come from
the committed rocprofv3 counter summary of that kernel when one covers the
workload.  `kernels` lists every sampled kind, the Gauss-Jordan panel with its
achieved bytes/s.  `cpu_baseline` runs the complex128 CPU oracle in the
reference's layout -- one process per host core on MKL sequential, energies
round-robin -- on a bounded sample of the same wire (oracle/cpu_baseline.py).

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
# Truncation of the contact broadening in the mixed forms' density chains (T12 keeps
# the whole factors, so transmission is untouched).  Measured at block 2304 against
# the plain form: 1e-6 of tr(Gamma) leaves the densities' error at the complex64
# chain's own (3e-6 of their scale) for 7% more energy points per second; 1e-5
# moves them by 1.3e-5 for 16%.  The default is 1e-5: the target bounds the
# transmission error, and the densities' error is reported beside it.
GAMMA_TRUNC_MIXED = 1e-5

CONFIGS = {
    # small, CPU-runnable: 64 energies of a clean 6x6 wire
    "c1": dict(npx=20, npy=6, npz=6, ne=64, emin=-0.3, emax=0.3, W=0.0, seed=1, per_step=64,
               budget_gb=0.0),
    # plain FP64 mid-size: block 1024, Npx=100, W=0.5
    "c2": dict(npx=100, npy=16, npz=16, ne=512, emin=-0.4, emax=0.4, W=0.5, seed=1, per_step=64,
               budget_gb=0.0),
    # precision target size: block 2304, Npx=200, W=0.3.  262 GB of energies in flight
    # (the library caps it at 90% of free HBM, ~290 GB once Htotd is resident) holds
    # three energies per lane on four lanes in the lu32 form (20.8 GB each); the plain
    # FP64 form (36.8 GB each) fits one per lane
    "c4": dict(npx=200, npy=24, npz=24, ne=1024, emin=-0.3, emax=0.3, W=0.3, seed=1, per_step=12,
               budget_gb=262.0),
}
# Many-wire workloads run through the serving / batching layer (negf_server):
#  c3 -- a seeded fraction of the batching configuration (256 disorder realisations x
#        32 energies, Npy=Npz in {8,12,16,20}, Npx in [40,120], W in [0,1], g*B in
#        [0,0.05], flux in [0,2]): --c3-instances realisations x --c3-energies of
#        their 32 energies, all submitted at once; the fraction is in config.workload;
#  c5 -- a serving stream drawn from the configured distribution: Poisson arrivals
#        at --c5-rate requests/s, per request Npy=Npz in {8,...,24}, Npx in
#        [50,300], 16..256 energies in [-0.4,0.4], flux in [0,2], W in [0,1], own
#        seed.  One such request is ~100 s of GPU work, so --c5-requests bounds the
#        stream and --c5-energy-stride thins every request's energy list (every
#        k-th energy of its grid) to keep the run to minutes; the line reports the
#        offered load rho = rate x mean service time for the thinned and the full
#        lists, and for each configured rate.
SERVER_WORKLOADS = {
    "c3": dict(instances=16, energies_per=8, ns=(8, 12, 16, 20), L=(40, 120), seed=3000),
    "c5": dict(requests=8, rate=10.0, ns=(8, 12, 16, 20, 24), L=(50, 300), ne=(16, 256), stride=16,
               seed=5000, rates=(10, 20, 50, 100, 200, 500)),
}
ETA, SR_TOL, SR_MAX = 1e-6, 1e-10, 200
PRECISIONS = {"fp64": capi.NEGF_PREC_FP64, "mixed": capi.NEGF_PREC_MIXED,
              "refined": capi.NEGF_PREC_REFINED, "lu32": capi.NEGF_PREC_LU32}
DTYPES = {"fp64": "f64", "mixed": "f64+c64 (mixed)",
          "refined": "f64+c64 (complex64 block inverses with FP64 refinement)",
          "lu32": "c64+f64 (complex64 block LU, FP64-refined transmission solve)"}
MU, MUT = (-0.1, 0.0), 0.0
LANES = [0]  # set from --lanes
OUTPUTS = ("T12", "Modec1", "Modec2", "DOSc1", "DOSc2", "EDensity", "HDensity", "RhoN", "RhoP",
           "SigXExp", "SigYExp", "SigZExp")


def cpu_baseline(mp_ctx, cfg, energies, k_mean, ranks, mode):
    """The complex128 CPU oracle in the reference's run layout (oracle/cpu_baseline.py):
    `ranks` processes on MKL sequential, energies round-robin.  mode "auto" solves
    whole energies end to end when one takes at most ~30 s per core (blocks up to
    256), and otherwise times the pieces of one energy at the workload's block
    size and composes the energy from them."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_baseline as CB  # noqa: E402  (checker only)
    n = cfg["npy"] * cfg["npz"] * 4
    if mode == "auto":
        mode = "full" if n <= 256 else "pieces"
    spec = dict(npx=cfg["npx"], npy=cfg["npy"], npz=cfg["npz"], W=cfg["W"], seed=cfg["seed"])
    if mode == "full":
        energies = energies[:ranks * max(1, len(energies) // ranks)]
    return CB.run(mp_ctx, spec, ranks, mode, energies, k_mean, ETA, SR_TOL, SR_MAX, MU, MUT)


# Committed summaries per (workload, kernel kind) at the shape the bench launches
# the kernel with: `alone_tflops` (the kernel alone: bench.py --gemm-only for the
# FP64 product, --inverse-only for the complex64 trailing update) and, where
# counters were collected, `traffic_bytes_per_launch` (rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE passes, each scaled by the same counter's reading for a copy
# kernel of known size).
# (config, kind) -> counter summary of that kind's launches at the config's shapes; the
# 3M rank-128 update is measured with one block per update (the sweep groups four)
TRAFFIC_PROFILES = {("c4", "fp64_gemm_3m"): "traffic_fp64_gemm3m_n2304_b3.md",
                    ("c4", "fp32_gj_update_3m"): "traffic_c64_gj_update3m_n2304_b3.md",
                    ("c4", "fp32x3_gj_update"): "traffic_c64x3_gj_update_g4_n2304_b3.md"}
GEMM_KINDS = ("fp64_gemm", "fp64_gj_update", "fp32_gemm", "fp32_gj_update", "fp64_gemm_3m",
              "fp64_gj_update_3m", "fp32_gemm_3m", "fp32_gj_update_3m", "fp32x3_gemm", "fp32x3_gj_update")
KERNEL_NAMES = {
    "fp64_gemm_3m": "k_gemm<double,0,0,0,2,2,0,1> (FP64 MFMA complex product, 3M form)",
    "fp64_gemm": "k_gemm<double,0,0,0,2,2,0,0> (FP64 MFMA complex product, 4-product form)",
    "fp64_gj_update_3m": "k_gemm<double,0,1,1,2,2,0,1> (FP64 Gauss-Jordan trailing update, 3M form)",
    "fp64_gj_update": "k_gemm<double,0,1,1,2,2,*,0> (FP64 Gauss-Jordan trailing update)",
    "fp32_gemm_3m": "k_gemm<float,0,0,0,2,2,0,1> (complex64 MFMA product, 3M form)",
    "fp32_gj_update_3m": "k_gemm<float,0,1,1,2,2,0,1> (complex64 Gauss-Jordan trailing update, 3M form)",
    "fp32x3_gemm": "k_gemm_x3<0,0,0> (complex64 product on the bf16 matrix cores, operands split three ways)",
    "fp32x3_gj_update": "k_gemm_x3<0,1,1> (complex64 Gauss-Jordan trailing update on the bf16 matrix cores)",
}
PEAK_BF16_MFMA = 2500.0  # TFLOP/s, MI355X dense bf16 matrix


def kind_peak(kind):
    """(peak TFLOP/s of the matrix cores a kind runs on, executed flops per
    algorithmic flop): the 3M forms run 6 of 8 m n k; the bf16 split form runs four
    real products of six bf16 products each, 48 m n k."""
    if kind and kind.startswith("fp32x3"):
        return PEAK_BF16_MFMA, 6.0
    peak = PEAK_FP32_MFMA if kind and kind.startswith("fp32") else PEAK_FP64_MFMA
    return peak, 0.75 if kind and kind.endswith("_3m") else 1.0


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
    """A product kernel alone: the 3M complex product (FP64, or complex64 with
    --gemm-c64) with the block size of the workload and the batch one lane
    launches it with (the energies in flight per lane)."""
    cfg = CONFIGS[args.config]
    n = cfg["npy"] * cfg["npz"] * 4
    per = args.per_step or cfg["per_step"]
    lanes = args.lanes or 4
    batch = max(1, per // lanes)
    reps = 8
    m3 = 2 if args.gemm_bf16x3 else True
    esize = 8 if (args.gemm_c64 or args.gemm_bf16x3) else 16
    ms = capi.bench_gemm(n, batch, complex64=args.gemm_c64 or args.gemm_bf16x3, m3=m3, reps=reps)
    flops = 8.0 * n ** 3 * batch
    kname = ("k_gemm_x3<0,0,0>" if args.gemm_bf16x3 else
             "k_gemm<%s,0,0,0,2,2,0,1>" % ("float" if args.gemm_c64 else "double"))
    print(json.dumps({"kernel": kname,
                      "n": n, "batch": batch, "launches": reps + 1, "ms_per_launch": ms,
                      "algorithmic_flops_per_launch": flops, "tflops": flops / ms / 1e9,
                      "algorithmic_bytes_per_launch": 3 * batch * n * n * esize}))


def inverse_only(args):
    """The complex64 block inverse of the LU32 sweep alone (negf_invert_batch_local:
    block Gauss-Jordan pivoting inside 128-column diagonal blocks, 3M form) at the
    block size of the workload and the batch one lane inverts, on diagonally
    weighted random matrices: its rank-128 trailing updates (k_gemm<float,0,1,1,...>
    over all n columns) and diagonal-block inverses are the sweep's launches, for
    rocprofv3 --pmc counter runs."""
    cfg = CONFIGS[args.config]
    n = cfg["npy"] * cfg["npz"] * 4
    per = args.per_step or cfg["per_step"]
    batch = max(1, per // (args.lanes or 4))
    rng = np.random.default_rng(7)
    G = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))) / np.sqrt(n)
    G += 4.0 * np.eye(n)
    G64 = G.astype(np.complex64)
    for _ in range(2):
        t0 = time.perf_counter()
        capi.profile_begin(1)
        _, info = capi.invert_batch(G64, complex64=True, m3=True if args.fp32_mfma else 2, local=True,
                                    group=args.gj_group)
        prof = capi.profile_end()
    upd = {k: {"tflops": v[0] / v[1] / 1e9 if v[1] > 0 else None, "ms": v[1], "launches": v[2]}
           for k, v in prof.items() if v[2] and k in ("fp32_gj_update_3m", "fp32x3_gj_update")}
    print(json.dumps({"kernel": "k_gemm<float,0,1,1,2,2,0,1> inside negf_invert_batch_local", "n": n, "batch": batch,
                      "outer_blocks": (n + 127) // 128, "group": args.gj_group,
                      "seconds_incl_host_copies": time.perf_counter() - t0, "trailing_updates": upd,
                      "info": info.tolist()}))


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
             label, keep=1, gamma_trunc=0.0):
    """Warm up, then time `steps` steps of one form; returns timings, the
    accumulators of the last `keep` timed steps (by step index) and the launch
    profile."""
    L = cfg["npx"]
    prec = PRECISIONS[precision]
    ctx = capi.Context(L, cfg["npy"], cfg["npz"], 4, eta=ETA, sr_tol=SR_TOL, sr_max_iter=SR_MAX,
                       max_batch=per, flags=flags, precision=prec, hbm_budget_gb=cfg["budget_gb"],
                       lanes=LANES[0], gamma_trunc=gamma_trunc)
    t_up = time.perf_counter()
    ctx.set_hamiltonian(w)  # blocking host-to-device copies of Htotd, HopX and the spin matrices
    upload_s = time.perf_counter() - t_up

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
    times, iters, status, accs = [], [], [], {}
    t_start = time.perf_counter()
    for k in range(warmup, warmup + steps):
        acc = capi.Accumulators(L, cfg["npy"], cfg["npz"], cfg["ne"])
        t0 = time.perf_counter()
        st, it = run_step(k, acc)  # flush() returns after the device streams have drained
        times.append(time.perf_counter() - t0)
        iters.append(it)
        status.append(st)
        accs[k] = acc
        if len(accs) > keep:
            del accs[min(accs)]
        print(f"[bench] {label} step {k} {times[-1]:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    prof = capi.profile_end()
    refine = ctx.refine_stats()
    ctx.close()
    return dict(times=times, elapsed=elapsed, iters=np.concatenate(iters), status=status, accs=accs,
                prof=prof, refine=refine, upload_s=upload_s)


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
        wl = dict(wl, instances=args.c3_instances or wl["instances"],
                  energies_per=args.c3_energies or wl["energies_per"])
        grid = energy_grid(-0.4, 0.4, 32)
        for r in range(wl["instances"]):
            ny = int(rng.choice(wl["ns"]))
            L = int(rng.integers(wl["L"][0], wl["L"][1] + 1))
            w = build_wire(L, ny, ny, disorder=float(rng.uniform(0, 1)), seed=wl["seed"] + r,
                           flux=float(rng.uniform(0, 2)), zeeman=float(rng.uniform(0, 0.05)), dense=False)
            sel = grid[(np.arange(wl["energies_per"]) * (32 // wl["energies_per"]) + r) % 32]
            reqs.append((0.0, w, sel))
        frac = wl["instances"] * wl["energies_per"] / (256 * 32)
        desc = (f"c3 fraction {frac:.4f} ({wl['instances']} of the 256 seeded realisations x "
                f"{wl['energies_per']} of their 32 energies in [-0.4,0.4]); Npy=Npz in {wl['ns']}, Npx in "
                f"{wl['L']}, W in [0,1], g*B in [0,0.05], flux in [0,2]; all submitted at t=0 as compact "
                f"Hamiltonians (slice template + site potentials)")
    else:
        wl = dict(wl, requests=args.c5_requests or wl["requests"], rate=args.c5_rate or wl["rate"],
                  stride=args.c5_energy_stride or wl["stride"])
        t = 0.0
        full_points = 0
        for r in range(wl["requests"]):
            t += float(rng.exponential(1.0 / wl["rate"]))
            ny = int(rng.choice(wl["ns"]))
            L = int(rng.integers(wl["L"][0], wl["L"][1] + 1))
            ne = int(rng.integers(wl["ne"][0], wl["ne"][1] + 1))
            w = build_wire(L, ny, ny, disorder=float(rng.uniform(0, 1)), seed=wl["seed"] + r,
                           flux=float(rng.uniform(0, 2)), dense=False)
            grid = energy_grid(-0.4, 0.4, ne)
            full_points += ne
            reqs.append((t, w, grid[::wl["stride"]]))
        desc = (f"c5 stream: {wl['requests']} requests drawn from the configured distribution, Poisson "
                f"arrivals at {wl['rate']}/s; per request Npy=Npz in {wl['ns']}, Npx in {wl['L']}, "
                f"{wl['ne'][0]}..{wl['ne'][1]} energies in [-0.4,0.4] (every {wl['stride']}-th energy of "
                f"its grid solved), flux in [0,2], W in [0,1], own seed; compact Hamiltonians")
    prec = PRECISIONS[args.precision]
    npts = sum(len(e) for _, _, e in reqs)
    print(f"[bench] {name}: {len(reqs)} requests, {npts} energy points", file=sys.stderr, flush=True)
    # lane pools that hold one energy of the largest request (block 2304, Npx 300:
    # ~42 GB in the mixed forms)
    srv = capi.Server(eta=ETA, sr_tol=SR_TOL, sr_max_iter=SR_MAX, precision=prec, flags=args.flags,
                      lanes=args.lanes, lane_pool_gb=args.lane_pool_gb, gamma_trunc=args.gamma_trunc)
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
    dominant = dominant_kind(prof)
    peak, _ = kind_peak(dominant)
    g_flops, g_ms, g_launches = prof[dominant] if dominant else (0.0, 0.0, 0)
    gemm_tflops = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    out = {
        "metric": "energy-points/s on 1 GPU vs plain FP64 form at stated error; p50/p99 solve latency",
        "value": npts / elapsed,
        "unit": "energy-points/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * elapsed,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (seeded 4-band TI wires, host-built Hamiltonians)",
        "config": {"workload": desc, "requests": len(reqs), "energy_points": npts,
                   "gamma_trunc": args.gamma_trunc,
                   "precision": args.precision},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "note": "per request, submission to completion (queueing included)"},
        "status_nonzero": int(sum(1 for s in stat if s)),
        "roofline": {"bound": "mfma", "kernel": KERNEL_NAMES.get(dominant, dominant),
                     "achieved": gemm_tflops, "peak": peak, "unit": "TFLOP/s", "frac": gemm_tflops / peak,
                     "traffic": None, "launches": g_launches, "avg_launch_ms": g_ms / max(1, g_launches)},
        "kernels": kernel_table(prof, name),
    }
    if name == "c5":
        # the GPU is busy throughout (the stream arrives within ~requests/rate s), so the
        # mean service time of a request is elapsed / requests
        s_thin = elapsed / len(reqs)
        s_full = s_thin * full_points / max(1, npts)
        out["offered_load"] = {
            "mean_service_s_thinned": s_thin, "mean_service_s_full_lists": s_full,
            "rho_at_rate": {str(r): r * s_full for r in wl["rates"]},
            "rho_this_run": wl["rate"] * s_thin,
            "note": ("rho = arrival rate x mean service time of one request (full energy lists: the "
                     "measured service time scaled by the energy points the thinning dropped); rho >> 1: "
                     "the queue grows by ~(rate - 1/mean service) requests per second, so latencies are "
                     "set by the backlog; p50/p99 above are for this finite stream")}
    return out


def dominant_kind(prof):
    """The MFMA kernel kind with the largest summed duration among the sampled launches."""
    best = max(GEMM_KINDS, key=lambda k: prof[k][1])
    return best if prof[best][2] else None


def kernel_table(prof, config):
    """Per sampled kind: achieved rate (TFLOP/s algorithmic for GEMMs, GB/s for the
    Gauss-Jordan panels), launches and mean launch time."""
    out = {}
    for k, (work, ms, cnt) in prof.items():
        if not cnt:
            continue
        if k.endswith("gj_panel"):
            out[k] = {"gbytes_per_s": work / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
                      "hbm_frac": work / (ms * 1e-3) / 8e12 if ms > 0 else 0.0,
                      "launches": cnt, "avg_launch_ms": ms / cnt,
                      "note": "bytes: the panel's columns and pivot rows read and written once"}
            continue
        out[k] = {"tflops": work / (ms * 1e-3) / 1e12 if ms > 0 else 0.0, "launches": cnt,
                  "avg_launch_ms": ms / cnt, "summed_sampled_ms": ms}
        a = profile_value(config, k, "alone_tflops")
        if a is not None:
            out[k]["alone_tflops"] = a
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS) + sorted(SERVER_WORKLOADS))
    ap.add_argument("--precision", default="lu32", choices=tuple(PRECISIONS))
    ap.add_argument("--lanes", type=int, default=0, help="concurrent device lanes (0 = library default)")
    ap.add_argument("--per-step", type=int, default=0, help="energies per step (default per config)")
    ap.add_argument("--budget-gb", type=float, default=-1.0,
                    help="HBM for energies in flight (default per config; 0 = library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain-FP64 comparison leg of a reduced-precision form")
    ap.add_argument("--compare-steps", type=int, default=3,
                    help="timed steps of the plain-FP64 comparison leg (after the same warmup, "
                         "capped at one step, on the energies of the last timed steps)")
    ap.add_argument("--cpu-ranks", type=int, default=0,
                    help="CPU baseline ranks (0 = one per host core, at most 16)")
    ap.add_argument("--cpu-baseline-mode", default="auto", choices=("auto", "full", "pieces"))
    ap.add_argument("--c3-instances", type=int, default=0, help="c3: realisations (default 16 of 256)")
    ap.add_argument("--c3-energies", type=int, default=0, help="c3: energies per realisation (default 8 of 32)")
    ap.add_argument("--c5-requests", type=int, default=0, help="c5: requests in the stream (default 8)")
    ap.add_argument("--c5-rate", type=float, default=0.0, help="c5: arrivals per second (default 10)")
    ap.add_argument("--c5-energy-stride", type=int, default=0,
                    help="c5: solve every k-th energy of each request's grid (default 16)")
    ap.add_argument("--lane-pool-gb", type=float, default=45.0,
                    help="server workloads: working pool per lane (GB)")
    ap.add_argument("--flags", type=int, default=0, help="NEGF_FLAG_* algorithm selection")
    ap.add_argument("--gamma-trunc", type=float, default=None,
                    help="negf_config.gamma_trunc of the measured form (default: %g in the reduced-precision "
                         "forms, 0 in the plain one; the plain FP64 compare leg always uses 0)"
                         % GAMMA_TRUNC_MIXED)
    ap.add_argument("--profile-every", type=int, default=32,
                    help="time one launch in this many of each kind with HIP events")
    ap.add_argument("--gemm-only", action="store_true",
                    help="run only the 3M complex product kernel at the shape the sweep launches "
                         "it with, for rocprofv3 --pmc counter runs")
    ap.add_argument("--gemm-c64", action="store_true",
                    help="with --gemm-only: the complex64 product (default FP64)")
    ap.add_argument("--gemm-bf16x3", action="store_true",
                    help="with --gemm-only: the complex64 product on the bf16 matrix cores (three-way split)")
    ap.add_argument("--inverse-only", action="store_true",
                    help="run only the complex64 Gauss-Jordan inverse at the block size and batch of one "
                         "lane, for rocprofv3 --pmc counter runs of its trailing updates")
    ap.add_argument("--gj-group", type=int, default=1,
                    help="--inverse-only: diagonal blocks per delayed trailing update")
    ap.add_argument("--fp32-mfma", action="store_true",
                    help="lu32 form: complex64 products on the FP32 matrix cores (NEGF_FLAG_LU32_FP32_MFMA) "
                         "instead of the bf16 split form; with --inverse-only, the same for the inverse")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB 8 times (the known-size "
                         "reference for rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.fp32_mfma:
        args.flags |= capi.FLAG_LU32_FP32_MFMA
    if args.gemm_only:
        gemm_only(args)
        return
    if args.inverse_only:
        inverse_only(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: this benchmark runs on exactly one GPU "
                                   "(energies shard across GPUs with no communication until the "
                                   "final sums; only --gpus 1 is measured)"}))
        sys.exit(2)

    LANES[0] = args.lanes
    if args.gamma_trunc is None:
        args.gamma_trunc = GAMMA_TRUNC_MIXED if args.precision != "fp64" else 0.0
    if args.config in SERVER_WORKLOADS:
        print(json.dumps(server_workload(args.config, args)), flush=True)
        return
    mp_ctx = None
    if not args.no_cpu_baseline:
        # CPU-baseline ranks are forked from a server process started now, before
        # this process touches the GPU
        import multiprocessing as mp
        mp_ctx = mp.get_context("forkserver")
        mp_ctx.set_forkserver_preload(["numpy"])
        from multiprocessing import forkserver
        forkserver.ensure_running()
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
    ncmp = max(1, min(args.compare_steps, args.steps))
    run = run_form(w, cfg, args.precision, args.flags, per, grid, dE, step_indices, args.steps,
                   args.warmup, args.profile_every, args.precision, keep=ncmp,
                   gamma_trunc=args.gamma_trunc)
    npts = per * args.steps
    elapsed = run["elapsed"]
    value = npts / elapsed
    timed_idx = np.unique(np.concatenate([step_indices(k) for k in range(args.warmup, nsteps)]))
    compare = None
    if args.precision != "fp64" and not args.no_compare:
        # the plain FP64 form, warmed up the same way (one step at most), on the
        # energies of the last `ncmp` timed steps
        last = sorted(run["accs"])
        base = run_form(w, cfg, "fp64", args.flags, per, grid, dE, lambda k: step_indices(last[k - 1]),
                        ncmp, min(args.warmup, 1), args.profile_every, "fp64", keep=ncmp)
        plain_rate = per * ncmp / base["elapsed"]
        err = {}
        for j, k in enumerate(last):
            for o, e in errors_vs(run["accs"][k], base["accs"][min(base["accs"]) + j]).items():
                err[o] = max(err.get(o, 0.0), e)
        compare = {
            "plain_fp64_energy_points_per_s": plain_rate,
            "plain_fp64_steps": ncmp,
            "speedup_vs_plain_fp64": value / plain_rate,
            "error_vs_plain_fp64": err,
            "note": (f"plain FP64 form after {min(args.warmup, 1)} warmup step, {ncmp} timed steps on "
                     "the energies of the last timed steps; errors: max over those steps of "
                     "|timed form - fp64| / the output's natural scale; the plain FP64 form agrees with "
                     "the complex128 CPU oracle to 1e-10 in the test suite"),
        }

    iters = run["iters"]
    prof = run["prof"]
    dom = dominant_kind(prof)
    peak, executed = kind_peak(dom)
    g_flops, g_ms, g_launches = prof[dom] if dom else (0.0, 0.0, 0)
    gemm_tflops = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
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
        "dtype": DTYPES[args.precision],
        "data": "synthetic (seeded 4-band TI wire, host-built Hamiltonian)",
        "config": {"workload": (f"{args.config}: Npx={L}, Npy=Npz={cfg['npy']} (block {n}), "
                                f"{cfg['ne']} energies in [{cfg['emin']},{cfg['emax']}], "
                                f"Anderson W={cfg['W']} seed {cfg['seed']}, eta={ETA}, "
                                f"SR tol {SR_TOL} max {SR_MAX}; {args.precision} form"
                                + (f" with the contact broadening truncated to trace {args.gamma_trunc:g} "
                                   "(densities and spin; T12 on the whole factors)"
                                   if args.gamma_trunc > 0 else "") + "; timed: "
                                f"{len(timed_idx)} of the {cfg['ne']} grid energies, {per} per step "
                                f"(strided slices of the grid, shifted each step)"),
                   "energies_per_step": per, "energies_timed": int(len(timed_idx)),
                   "precision": args.precision, "flags": args.flags, "lanes": args.lanes or 4,
                   "gamma_trunc": args.gamma_trunc},
        "latency_ms": {"p50": 1e3 * float(np.percentile(run["times"], 50)),
                       "p99": 1e3 * float(np.percentile(run["times"], 99)),
                       "note": ("a step's energies complete together: per-energy solve latency = "
                                "step time")},
        "sr_iters_mean": float(iters.mean()),
        "status_nonzero": int(sum(int(np.count_nonzero(s)) for s in run["status"])),
        "host_upload": {
            "seconds": run["upload_s"],
            "bytes": int(w.htotd.nbytes + w.hopx.nbytes),
            "energy_points_per_s_incl_upload": cfg["ne"] / (cfg["ne"] / value + run["upload_s"]),
            "note": ("the Hamiltonian handed over from host memory once (Fortran-order copy on the "
                     "host, then blocking copies over PCIe), outside the timed region; the rate "
                     "including it is for the whole grid of this Hamiltonian at the timed rate"),
        },
        "roofline": {"bound": "mfma", "kernel": KERNEL_NAMES.get(dom, dom),
                     "achieved": gemm_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": gemm_tflops / peak, "executed_frac": executed * gemm_tflops / peak,
                     "traffic": traffic_from_profiles(args.config, dom),
                     "launches": g_launches, "avg_launch_ms": g_ms / max(1, g_launches),
                     "note": ("kernel: the MFMA kind with the largest summed sampled duration; achieved "
                              "= algorithmic 8 n^3 flops per complex product x batch / HIP-event duration "
                              "on the launching stream (one launch in --profile-every of each kind); "
                              "peak: the matrix cores the kind runs on (FP64, FP32, or bf16 for the "
                              "split form); executed_frac counts what they run: 6 of 8 n^3 in the 3M "
                              "form, 48 in the bf16 split form (four real products of six); the "
                              "lanes run concurrently, so durations include other lanes' kernels sharing "
                              "the CUs; traffic: HBM-side bytes per launch from FETCH_SIZE/WRITE_SIZE "
                              "counter runs of the kernel alone at this shape (bench.py --gemm-only or "
                              "--inverse-only), scaled by a copy-kernel calibration")},
        "kernels": kernel_table(prof, args.config),
    }
    alone = profile_value(args.config, dom, "alone_tflops") if dom else None
    if alone is not None:
        # the same kernel and shape with the GPU to itself (bench.py --gemm-only or --inverse-only)
        result["roofline"]["alone"] = {"achieved": alone, "frac": alone / peak,
                                       "executed_frac": executed * alone / peak,
                                       "source": "profiles/" + TRAFFIC_PROFILES[(args.config, dom)]}
    if compare is not None:
        result["compare"] = compare
    if args.precision == "refined":
        b, st, e = run["refine"]
        result["refine"] = {"block_inverses": b, "fp64_newton_steps": st, "escalated_to_fp64": e,
                            "note": "counts since the context was created (warmup steps included)"}
    if args.precision == "lu32":
        b, st, e = run["refine"]
        result["refine"] = {"energies": b, "fp64_refinement_sweeps": st, "escalated_to_fp64": e,
                            "note": "counts since the context was created (warmup steps included)"}
    if mp_ctx is not None:
        t0 = time.perf_counter()
        ranks = args.cpu_ranks or min(16, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(mp_ctx, cfg, grid[timed_idx], float(iters.mean()), ranks,
                                              args.cpu_baseline_mode)
        print(f"[bench] cpu baseline {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
