"""The complex64 block-LU form (NEGF_PREC_LU32) against the complex128 oracle.

Every RGF block inverse and both column chains run in complex64; the contacts stay
FP64.  The leading factor columns (at most 64) of both block columns are solved
again to FP64 accuracy by iterative refinement of the block-tridiagonal system;
every output sums over factor columns, taking those columns from the refined
solve and the others from the complex64 chains.  T12 = |W_L' G_1L W_R|^2 adds the
complex64 parts past W_R's refined columns (the first-row chain and the tail
sketch, or the exact tail when the chain runs wide); an energy that does not pass
the form's checks is solved again in the plain FP64 form.  Tolerances: T12 within
1e-8 of its natural scale (the open-channel count) or the oracle's own spread,
DOSc within 1e-7, the other outputs (densities, spin, modes, currents) within 2e-5
of their scales; Sancho-Rubio iteration counts equal to the oracle's.  The
measured errors are printed."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

import parity as P  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_bench_size as MB  # noqa: E402

MU, MUT = (-0.1, 0.0), 0.0
DE = 0.01
T_RTOL = 1e-8
DOS_RTOL = 1e-7
C64_RTOL = 2e-5


def _collect(acc):
    d = {k: np.array(getattr(acc, k)) for k in P.PER_E + ("ContactCurrent",)}
    for k, a in (("RhoN", "RhoN"), ("RhoP", "RhoP"), ("SigX", "SigXExp"), ("SigY", "SigYExp"),
                 ("SigZ", "SigZExp")):
        d[k] = np.asarray(getattr(acc, a)).ravel(order="F")
    return d


def _run(w, energies, dE, precision=None, flags=0, batches=None, gamma_trunc=0.0, compact=False):
    precision = capi.NEGF_PREC_LU32 if precision is None else precision
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(energies), norb=w.norb)
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                      precision=precision, flags=flags, gamma_trunc=gamma_trunc) as ctx:
        if compact:
            ctx.set_hamiltonian_compact(w)
        else:
            ctx.set_hamiltonian(w)
        its = []
        for part in (batches or [list(range(len(energies)))]):
            ctx.submit([energies[i] for i in part], part, np.full(len(part), dE))
            st, it = ctx.flush(MU, MUT, acc)
            assert np.all(np.asarray(st) == 0), st
            its.append(np.asarray(it))
        stats = ctx.refine_stats()
    return acc, np.concatenate(its), stats


def check(got, ref, noise, label):
    sc = P.scales(ref)
    lim = np.maximum(T_RTOL * sc["T12"], P.FLOOR * noise["T12"])
    err = np.abs(got["T12"] - ref["T12"])
    print(f"{label}: T12 rel err {(err / sc['T12']).max():.2e}", flush=True)
    assert np.all(err <= lim), ("T12", err / sc["T12"])
    for k in ("DOSc1", "DOSc2"):
        e = float(np.abs(got[k] - ref[k]).max()) / sc[k]
        assert e <= DOS_RTOL, (k, e)
    errs = {}
    for k in ("Modec1", "Modec2", "EDensity", "HDensity", "RhoN", "RhoP", "SigX", "SigY", "SigZ",
              "ContactCurrent"):
        errs[k] = float(np.abs(np.asarray(got[k]) - np.asarray(ref[k])).max()) / sc[k]
    print(f"{label}: relative errors", {k: f"{v:.1e}" for k, v in errs.items()},
          flush=True)
    for k, v in errs.items():
        assert v <= C64_RTOL, (k, v)


@pytest.mark.parametrize("case", [
    ((8, 3, 2), dict(disorder=0.5, seed=7, flux=0.4, zeeman=0.03), [-1.7, -0.3, 0.05, 0.25, 1.1]),
    ((12, 4, 4), dict(disorder=0.5, seed=1), [-0.4, -0.1, 0.2, 0.4]),
    ((1, 2, 3), dict(zeeman=0.02), [0.15, 0.9]),
    ((6, 1, 1), dict(model="metal", params=dict(t=1.0, eps=0.0), disorder=0.3, seed=2), [-0.7, 0.4]),
], ids=["dfz", "disorder", "one_slice", "chain"])
def test_lu32_matches_oracle(case):
    dims, kw, energies = case
    w = build_wire(*dims, **kw)
    print(f"lu32 case dims={dims}", flush=True)
    ref, noise = P.reference(w, energies, DE, mu=MU, muT=MUT)
    acc, it, (ne, sweeps, esc) = _run(w, energies, DE)
    assert np.array_equal(it, ref["iters"])
    check(_collect(acc), ref, noise, f"dims {dims}")
    assert ne >= len(energies) and sweeps == 2 * ne
    print(f"energies {ne}, sweeps {sweeps}, escalated {esc}", flush=True)


def test_lu32_slice_couplings_off_the_stencil_match_oracle():
    # The refinement's products H_x V_x take a stencil kernel when every slice keeps to
    # the site blocks of a site and its y, z neighbours, and the dense product otherwise:
    # a hopping between opposite corners of the cross-section in the interior slices
    # (the leads keep the stencil) takes the second path, to the same bounds
    w = build_wire(7, 3, 3, disorder=0.4, seed=5)
    no = w.norb
    a, b = 0, (w.npy * w.npz - 1) * no  # site (0, 0) and site (2, 2)
    t = 0.15 * np.eye(no) + 0.05j * np.roll(np.eye(no), 1, axis=1)
    for x in range(1, w.npx - 1):
        w.htotd[a:a + no, b:b + no, x] += t
        w.htotd[b:b + no, a:a + no, x] += t.conj().T
    energies = [-0.9, 0.2, 1.2]
    ref, noise = P.reference(w, energies, DE, mu=MU, muT=MUT)
    acc, it, (ne, sweeps, esc) = _run(w, energies, DE)
    assert np.array_equal(it, ref["iters"])
    check(_collect(acc), ref, noise, "off-stencil slice couplings")


def test_lu32_fp32_matrix_cores_match_oracle():
    # NEGF_FLAG_LU32_FP32_MFMA: the complex64 products on the FP32 matrix cores (3M)
    # instead of the bf16 split form, to the same bounds
    w = build_wire(8, 3, 2, disorder=0.5, seed=7, flux=0.4, zeeman=0.03)
    energies = [-1.7, -0.3, 0.05, 0.25, 1.1]
    ref, noise = P.reference(w, energies, DE, mu=MU, muT=MUT)
    acc, it, (ne, sweeps, esc) = _run(w, energies, DE, flags=capi.FLAG_LU32_FP32_MFMA)
    assert np.array_equal(it, ref["iters"])
    check(_collect(acc), ref, noise, "FP32 matrix cores")


CASES = [(name, tag, dims, kw) for name, lst in sorted(MB.CASES.items())
         for tag, dims, kw, _, _ in lst]


@pytest.mark.parametrize("name,tag,dims,kw", CASES, ids=[c[1] for c in CASES])
def test_lu32_bench_size_matches_oracle(name, tag, dims, kw):
    """The benchmark workloads' block sizes (committed oracle fixtures): block 1024
    over 100 slices, the block-2304 cross-section, the batching cross-sections."""
    print(f"lu32 bench-size fixture {tag}: dims {dims} {kw}", flush=True)
    w = build_wire(*dims, **kw)
    energies, dE, digest, ref, noise = MB.load(name, tag)
    assert digest == MB.wire_digest(w), "fixture was made for another wire"
    acc, it, (ne, sweeps, esc) = _run(w, [float(e) for e in energies], dE)
    assert np.array_equal(it, ref["iters"]), (it.tolist(), ref["iters"].tolist())
    check(_collect(acc), ref, noise, tag)
    print(f"{tag}: energies {ne}, escalated {esc}", flush=True)
    assert esc == 0


@pytest.mark.parametrize("tag", ["c2", "c4x"])
def test_lu32_truncated_broadening_keeps_t12(tag):
    """With the broadening truncated for the densities (gamma_trunc = 1e-5, the bench
    setting), the last-column chain runs narrow and T12's part past W_R's refined
    columns comes from the first-row chain and the tail sketch: T12 still matches the
    oracle (whole Gamma) within 1e-8 of its scale, and no energy is escalated."""
    name = "c2" if tag == "c2" else "c4"
    (t, dims, kw, _, _), = [c for c in MB.CASES[name] if c[0] == tag]
    print(f"lu32 truncated broadening, fixture {tag}", flush=True)
    w = build_wire(*dims, **kw)
    energies, dE, digest, ref, noise = MB.load(name, tag)
    assert digest == MB.wire_digest(w)
    acc, it, (ne, sweeps, esc) = _run(w, [float(e) for e in energies], dE, gamma_trunc=1e-5)
    got = _collect(acc)
    sc = P.scales(ref)
    err = np.abs(got["T12"] - ref["T12"]) / sc["T12"]
    print(f"{tag}: T12 rel err {err.max():.2e}, escalated {esc}", flush=True)
    assert np.all(err <= T_RTOL), err
    assert esc == 0


def test_lu32_compact_equals_dense():
    # The complex64 chains see the same assembled slices either way; the FP64
    # residual adds the site potentials of a compact slice apart from the template
    # product, so the refined outputs agree to rounding: 1e-12 of each output's scale
    # (max(1, |value|) for T12 and the modes, the array's largest entry otherwise).
    w = build_wire(7, 3, 3, disorder=0.6, seed=5, flux=0.2)
    energies = [-1.4, -0.35, 0.1, 0.3, 1.2]
    dense = _run(w, energies, DE)[0]
    comp = _run(w, energies, DE, compact=True)[0]
    for k in ("T12", "Modec1", "Modec2", "RhoN", "RhoP", "SigZExp", "Jx"):
        a, b = np.asarray(getattr(dense, k)), np.asarray(getattr(comp, k))
        scale = np.maximum(1.0, np.abs(a)) if k in ("T12", "Modec1", "Modec2") else np.abs(a).max()
        assert np.all(np.abs(a - b) <= 1e-12 * scale), (k, np.abs(a - b).max())


def test_lu32_is_batch_and_graph_invariant():
    w = build_wire(9, 3, 2, disorder=0.5, seed=11, zeeman=0.02)
    energies = [-1.6, -0.25, 0.05, 1.3]
    one = _run(w, energies, DE, batches=[[i] for i in range(4)], gamma_trunc=1e-3)[0]
    mix = _run(w, energies, DE, batches=[[0, 1], [2, 3]], gamma_trunc=1e-3)[0]
    gra = _run(w, energies, DE, batches=[[0, 1], [2, 3]], flags=capi.FLAG_GRAPH, gamma_trunc=1e-3)[0]
    for k in ("RhoN", "RhoP", "SigXExp", "T12", "Modec1", "Modec2", "Jx", "Jy"):
        a = np.asarray(getattr(one, k))
        assert np.array_equal(a, np.asarray(getattr(mix, k))), k
        assert np.array_equal(a, np.asarray(getattr(gra, k))), k


def test_lu32_escalates_near_singular_wire():
    # a weakly coupled wire at an eigenvalue of an interior slice: the block LU's
    # complex64 factors cannot carry the ~1/eta condition number, the refinement
    # does not contract, and the energy is solved again in the plain FP64 form
    w = build_wire(5, 2, 2, disorder=0.6, seed=3)
    w.hopx = np.asfortranarray(w.hopx * 1e-3)
    lam = np.linalg.eigvalsh(w.htotd[:, :, 2])
    E = float(lam[len(lam) // 2])
    energies = [E, E + 0.3]
    print(f"lu32 escalation at E={E}", flush=True)
    ref, noise = P.reference(w, energies, DE, mu=MU, muT=MUT)
    acc, it, (ne, sweeps, esc) = _run(w, energies, DE)
    assert esc >= 1, (ne, sweeps, esc)
    assert np.array_equal(it, ref["iters"])
    check(_collect(acc), ref, noise, "escalation")


def test_lu32_dense_hopx_takes_the_plain_form():
    # a HopX with a coupling between different sites is not site-block-diagonal:
    # every energy is solved in the plain FP64 form, to the plain form's bound
    w = build_wire(6, 2, 2, disorder=0.4, seed=8)
    h = np.array(w.hopx)
    h[0, 5] = 0.05
    w.hopx = np.asfortranarray(h)
    energies = [-0.3, 0.2]
    ref, noise = P.reference(w, energies, DE, mu=MU, muT=MUT)
    acc, it, (ne, sweeps, esc) = _run(w, energies, DE)
    assert esc == len(energies)
    assert np.array_equal(it, ref["iters"])
    check(_collect(acc), ref, noise, "dense HopX")
