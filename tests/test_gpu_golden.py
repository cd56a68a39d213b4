"""The GPU path against the committed golden fixtures (tests/golden/negf_golden.npz):
the reference's own compiled Get_Retarded_Greens / Calculate_Gnp outputs and the
oracle's Energy_Sweep outputs, which travel to the GPU machine in place of the
reference.  Staged entry points: relative 1e-10 of each array's max-norm; the fused
sweep: the plain-FP64 criterion of tests/parity.py with the oracle's
summation-order spread measured live."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

import parity as P  # noqa: E402

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "negf_golden.npz")
CASES = {"clean": dict(), "dfz": dict(disorder=0.4, seed=6, flux=0.2, zeeman=0.02),
         "strong": dict(disorder=0.9, seed=1)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_stages_match_reference_golden(name):
    gold = np.load(PATH)
    p = name + "/"
    d = gold[p + "dims"]
    w = build_wire(int(d[0]), int(d[1]), int(d[2]), **CASES[name])
    print(f"golden stages {name}", flush=True)
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6) as ctx:
        ctx.set_hamiltonian(w)
        r = ctx.retarded(gold[p + "invgrd"])
        g = ctx.gnp(1.0, 0.0, 0.01, r["GRLef"], r["GRRgt"], gold[p + "GammaL"], gold[p + "GammaR"])
    for k in ("GRLef", "GRRgt", "GRd"):
        ref = gold[p + "ref_" + k]
        assert np.abs(r[k] - ref).max() <= 1e-10 * np.abs(ref).max(), k
    for k in ("Gnd", "RhoN", "RhoP"):
        ref = gold[p + "ref_" + k]
        assert np.abs(g[k] - ref).max() <= 1e-10 * np.abs(ref).max(), k


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_sweep_matches_golden(name):
    gold = np.load(PATH)
    p = name + "/"
    d = gold[p + "dims"]
    w = build_wire(int(d[0]), int(d[1]), int(d[2]), **CASES[name])
    E = float(gold[p + "energy"][0])
    mu = (-2.0, 0.0)
    print(f"golden sweep {name} E={E}", flush=True)
    acc = capi.Accumulators(w.npx, w.npy, w.npz, 1)
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200) as ctx:
        ctx.set_hamiltonian(w)
        ctx.submit([E], [0], [0.01])
        st, it = ctx.flush(mu, 0.0, acc)
    assert np.all(np.asarray(st) == 0)
    ref = {k: gold[p + "sweep_" + k] for k in P.PER_E}
    for k in ("RhoN", "RhoP", "SigX", "SigY", "SigZ", "ContactCurrent"):
        ref[k] = gold[p + "sweep_" + k].ravel(order="F") if k != "ContactCurrent" else gold[p + "sweep_" + k]
    _, noise = P.reference(w, [E], 0.01, mu=mu, muT=0.0)
    got = {k: getattr(acc, k) for k in P.PER_E + ("ContactCurrent",)}
    for k, a in (("RhoN", "RhoN"), ("RhoP", "RhoP"), ("SigX", "SigXExp"), ("SigY", "SigYExp"),
                 ("SigZ", "SigZExp")):
        got[k] = np.asarray(getattr(acc, a)).ravel(order="F")
    P.check_all(got, ref, noise)
