"""GPU tests of the device-side Hamiltonian assembly (negf_set_hamiltonian_compact):
a wire given as its slice template H0 plus site potentials must give bitwise the
same outputs as the same wire given as the dense Htotd -- in the fused sweep (both
precision forms, unequal end slices included), in the staged observables with the
bond currents, and as a server request."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

import oracle as O  # noqa: E402

MU, MUT, DE = (-0.1, 0.0), 0.0, 0.01
FIELDS = capi.Accumulators.PER_ENERGY + capi.Accumulators.SPATIAL + ("ContactCurrent",)


def _wire(which):
    if which == "disorder_flux_zeeman":
        return build_wire(7, 3, 2, disorder=0.7, seed=4, flux=0.3, zeeman=0.03)
    if which == "clean":
        return build_wire(5, 2, 2)
    w = build_wire(6, 2, 3, disorder=0.5, seed=8)
    # a gate on the last slice: end slices differ (separate Sancho-Rubio decimations)
    w.onsite[-1] += 0.2
    a = np.arange(w.n)
    w.htotd[a, a, -1] += 0.2
    return w


def _sweep(w, energies, compact, precision):
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(energies))
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                      precision=precision) as ctx:
        if compact:
            ctx.set_hamiltonian_compact(w)
        else:
            ctx.set_hamiltonian(w)
        ctx.submit(energies, np.arange(len(energies)), np.full(len(energies), DE))
        st, it = ctx.flush(MU, MUT, acc)
    assert np.all(np.asarray(st) == 0)
    return acc


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("which", ["disorder_flux_zeeman", "clean", "gated_end"])
def test_compact_sweep_equals_dense(which, precision):
    w = _wire(which)
    # the compact description reproduces the dense slices exactly
    a = np.arange(w.n)
    rec = np.repeat(w.h0[:, :, None], w.npx, axis=2)
    site = a // w.norb
    rec[a, a, :] += w.onsite[:, site % w.npy, site // w.npy].T
    assert np.array_equal(rec, w.htotd)
    energies = [-1.7, -0.2, 0.15, 1.2]
    print(f"compact vs dense: {which} precision={precision}", flush=True)
    dense = _sweep(w, energies, False, precision)
    comp = _sweep(w, energies, True, precision)
    for k in FIELDS:
        assert np.array_equal(getattr(dense, k), getattr(comp, k)), k


def test_compact_staged_observables_equal_dense():
    w = _wire("disorder_flux_zeeman")
    E = 0.2
    c = O.contact_self_energy(w, E, 1e-6)
    A = O.assemble_invg(w, c["SigmaL"], c["SigmaR"], E, 1e-6)
    r = O.retarded(A, w.hopx)
    out = []
    for compact in (False, True):
        with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6) as ctx:
            if compact:
                ctx.set_hamiltonian_compact(w)
            else:
                ctx.set_hamiltonian(w)
            g = ctx.gnp(1.0, 0.0, 0.0123, r["GRLef"], r["GRRgt"], c["GammaL"], c["GammaR"])
            out.append(ctx.observables(1.0, 0.0, 0.0123, g, r["GRRgt"], c["GammaL"], c["GammaR"]))
    for k in out[0]:
        assert np.array_equal(np.asarray(out[0][k]), np.asarray(out[1][k])), k
    assert np.abs(out[1]["Jy"]).max() > 0


def test_compact_server_request_equals_dense():
    w = _wire("gated_end")
    energies = [-0.3, 0.1, 0.9]
    with capi.Server() as srv:
        rd, dense = srv.submit(w, energies, DE, MU, MUT)
        rc, comp = srv.submit(w, energies, DE, MU, MUT, compact=True)
        assert srv.wait(rd)[0] == 0
        assert srv.wait(rc)[0] == 0
    for k in FIELDS:
        assert np.array_equal(getattr(dense, k), getattr(comp, k)), k
