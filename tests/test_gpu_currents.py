"""Bond currents and the energy-/orbital-resolved outputs on the fused path
(negf_flush, negf_energy_sweep, the server), against the complex128 oracle.

The device never forms Gnd/Gnu/Gnl: the bond terms of Calculate_Observables.F90
:111-168 come from the neighbour site blocks of V V' (V = G W, Gamma = W W') --
y and z bonds inside a slice, x bonds between the column blocks of consecutive
slices.  Plain FP64 form: the criterion of tests/parity.py on the current scale
q^2/hbar * Norb^2 * max|H| * max RhoN (a current term is an H x Gn product);
mixed form: 2e-5 of that scale (the first-column chain is complex64).  The
resolved arrays are checked against oracle.resolved (a plain-loop restatement of
the erCD / orCD branches, pinned on the CPU against the integrated outputs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

import oracle as O  # noqa: E402
import parity as P  # noqa: E402

MU, MUT = (-0.1, 0.0), 0.0
ENERGIES = [-1.6, -0.25, 0.05, 1.3]
DE = 0.02
QQ_HBAR = 1.602e-19 ** 2 / (6.626e-34 / (2 * np.pi))


def _wire(which):
    if which == "dfz":
        return build_wire(5, 2, 3, disorder=0.5, seed=12, flux=0.3, zeeman=0.03)
    w = build_wire(6, 3, 2, disorder=0.4, seed=5)
    w.onsite[-1] += 0.15  # a gate on the last slice: separate decimations
    a = np.arange(w.n)
    w.htotd[a, a, -1] += 0.15
    return w


def jscale(w, ref):
    return QQ_HBAR * w.norb ** 2 * np.abs(w.htotd).max() * float(np.abs(ref["RhoN"]).max())


def _fused(w, precision, flags=0, compact=False, resolved=False):
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(ENERGIES), resolved=resolved, norb=w.norb)
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                      precision=precision, flags=flags) as ctx:
        if compact:
            ctx.set_hamiltonian_compact(w)
        else:
            ctx.set_hamiltonian(w)
        ctx.submit(ENERGIES, np.arange(len(ENERGIES)), np.full(len(ENERGIES), DE))
        st, _ = ctx.flush(MU, MUT, acc)
    assert np.all(np.asarray(st) == 0)
    return acc


@pytest.mark.parametrize("which,flags,compact", [("dfz", 0, False), ("dfz", 0, True),
                                                 ("dfz", capi.FLAG_RGF_THREE_PASS, False),
                                                 ("gated", 0, False)])
def test_fused_bond_currents_match_oracle(which, flags, compact):
    w = _wire(which)
    print(f"fused currents: {which} flags={flags} compact={compact}", flush=True)
    ref, noise = P.reference(w, ENERGIES, DE, mu=MU, muT=MUT)
    sc = jscale(w, ref)
    acc = _fused(w, capi.NEGF_PREC_FP64, flags, compact)
    for k in ("Jx", "Jy", "Jz"):
        P.check(k, np.asarray(getattr(acc, k)).ravel(order="F"), ref[k], noise[k], sc)
        assert np.abs(ref[k]).max() > 0  # the currents are not trivially zero


def test_fused_bond_currents_mixed_form():
    w = _wire("dfz")
    print("fused currents: mixed form", flush=True)
    ref = P.oracle_sweep(w, ENERGIES, DE, MU, MUT)
    sc = jscale(w, ref)
    acc = _fused(w, capi.NEGF_PREC_MIXED)
    for k in ("Jx", "Jy", "Jz"):
        err = np.abs(np.asarray(getattr(acc, k)).ravel(order="F") - ref[k]).max() / sc
        assert err <= 2e-5, (k, err)


@pytest.mark.parametrize("precision", [capi.NEGF_PREC_FP64, capi.NEGF_PREC_MIXED],
                         ids=["fp64", "mixed"])
def test_resolved_outputs_match_oracle(precision):
    w = _wire("dfz")
    print(f"resolved outputs: precision={precision}", flush=True)
    ne = len(ENERGIES)
    ref = None
    for e, E in enumerate(ENERGIES):
        r = O.energy_resolved(w, E, DE, MU, MUT, 1e-6, 1e-10, 200, eCtr=e, estpnum=ne)
        ref = r if ref is None else {k: ref[k] + r[k] for k in r}
    acc = _fused(w, precision, resolved=True)
    rho = float(np.abs(acc.RhoN).max())
    tol = 1e-10 if precision == capi.NEGF_PREC_FP64 else 2e-5
    wgt = DE / (2 * np.pi)
    sj = jscale(w, {"RhoN": acc.RhoN})
    scales = {"ERJx": sj / wgt, "ERJy": sj / wgt, "ERJz": sj / wgt, "ORJx": sj, "ORJy": sj,
              "ORJz": sj, "ERRhon": rho / wgt, "ORRhon": rho}
    for k, s in scales.items():
        got = getattr(acc, k)
        assert got.shape == ref[k].shape, (k, got.shape, ref[k].shape)
        # 100x the plain bound: the sums run over 4 energies of band-edge decimations,
        # whose oracle spread is ~1e-9 (tests/parity.py)
        lim = (100 * tol if precision == capi.NEGF_PREC_FP64 else tol) * s
        err = np.abs(got - ref[k]).max()
        assert err <= lim, (k, err / s)
    # and they are consistent with the integrated outputs of the same run
    assert np.allclose(acc.ORJy.sum(axis=3), acc.Jy, rtol=1e-12, atol=1e-12 * sj)
    assert np.allclose(acc.ORRhon.sum(axis=3), acc.RhoN, rtol=1e-12, atol=1e-14 * rho)


def test_server_currents_equal_flush():
    w = _wire("dfz")
    acc = _fused(w, capi.NEGF_PREC_FP64, compact=True)
    with capi.Server(eta=1e-6, sr_tol=1e-10, sr_max_iter=200, lanes=2) as srv:
        rid, sacc = srv.submit(w, ENERGIES, DE, MU, MUT, compact=True)
        st, _ = srv.wait(rid)
    assert st == 0
    for k in ("Jx", "Jy", "Jz", "RhoN", "T12"):
        assert np.array_equal(getattr(sacc, k), getattr(acc, k)), k
