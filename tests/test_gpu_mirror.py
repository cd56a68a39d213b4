"""GPU parity of the paired decimation taken on the two sectors of a lead mirror
(the default when the end slices are equal and the end slice and HopX keep a mirror
of the cross-section) against the complex128 CPU oracle, whose Sancho-Rubio works
on the whole n x n blocks (Contact_Self_Energy.F90:141-209).

Tolerance as the other plain FP64 sweeps (tests/parity.py): relative 1e-10 of each
output's natural scale, and the Sancho-Rubio iteration counts equal to the oracle's
(the stop test reads the site-basis max|change|).  DOSc, the trace of the O(eta)
broadening inside the gap, is the output most sensitive to Sigma's error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

import parity as P  # noqa: E402

MU, MUT = (-0.1, 0.0), 0.0


def _sweep(w, energies, dE, flags=0, precision=capi.NEGF_PREC_FP64, compact=False, onsite=None):
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(energies))
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200, flags=flags,
                      precision=precision) as ctx:
        if compact:
            ctx.set_hamiltonian_compact(w)
            if onsite is not None:
                ctx.set_onsite(onsite)
        else:
            ctx.set_hamiltonian(w)
        mirror = ctx.contact_mirror()
        ctx.submit(energies, np.arange(len(energies)), np.full(len(energies), dE))
        st, it = ctx.flush(MU, MUT, acc)
    return acc, st, it, mirror


def _as_dict(acc):
    d = {k: getattr(acc, k) for k in P.PER_E + ("ContactCurrent",)}
    for k, a in (("RhoN", "RhoN"), ("RhoP", "RhoP"), ("SigX", "SigXExp"), ("SigY", "SigYExp"),
                 ("SigZ", "SigZExp")):
        d[k] = np.asarray(getattr(acc, a)).ravel(order="F")
    return d


CASES = [
    # (dims, build_wire kwargs, energies (in and out of the gap), mirror axis, paired sectors)
    ((6, 4, 4), dict(), [-0.9, -0.05, 0.2, 1.1], 2, True),
    ((5, 4, 3), dict(disorder=0.5, seed=2), [-0.3, 0.1, 0.6], 1, False),
    ((6, 2, 1), dict(model="metal", params=dict(t=1.0, eps=0.0), disorder=0.3, seed=2), [-1.5, 0.4, 2.3], 1,
     False),
    ((4, 6, 6), dict(disorder=0.4, seed=5), [-0.25, 0.0, 0.3], 2, True),
    # two commuting mirrors: the sectors are decimated separately.  Inside this lead's
    # bands (e.g. -1.2, 1.6) the decimation, which converges only through eta, is
    # sensitive to rounding: the oracle's own Sigma moves by 4e-9 - 6e-9 relative when
    # its decimation is taken in the mirror basis, and with one orbital there is no
    # orbital permutation to give a noise floor, so those energies test neither path
    # to 1e-10; these three move it by at most 6e-13
    ((5, 2, 2), dict(model="metal", params=dict(t=1.0, eps=0.0)), [-2.5, 0.3, 2.3], 2, False),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_mirror_sweep_matches_oracle(case):
    dims, kw, energies, axis, paired = CASES[case]
    w = build_wire(*dims, **kw)
    print(f"mirror sweep {dims} {kw}", flush=True)
    ref, noise = P.reference(w, energies, 0.01, mu=MU, muT=MUT)
    acc, st, it, mirror = _sweep(w, energies, 0.01)
    assert mirror[0] == axis and mirror[3] == paired, mirror
    assert np.all(st == 0), st
    P.check_all(_as_dict(acc), ref, noise)
    assert np.array_equal(np.asarray(it), ref["iters"]), (it, ref["iters"])
    # the n x n decimation: no mirror, the same results to the same bar
    acc2, st2, it2, mirror2 = _sweep(w, energies, 0.01, flags=capi.FLAG_SR_NO_MIRROR)
    assert mirror2[0] == 0
    assert np.array_equal(np.asarray(it2), np.asarray(it))
    P.check_all(_as_dict(acc2), ref, noise)


def test_mirror_follows_the_site_potentials():
    """Compact Hamiltonian: the mirror is found again for each set of site
    potentials (the end slices carry them) and the sweep stays on the oracle."""
    w = build_wire(5, 4, 4, disorder=0.3, seed=8)
    energies = [-0.4, 0.15]
    site_y = np.arange(w.npy)[None, :, None] * np.ones((w.npx, 1, w.npz))
    site_z = np.arange(w.npz)[None, None, :] * np.ones((w.npx, w.npy, 1))
    for V, axis in ((None, 2), (w.onsite + 0.04 * site_z, 1), (w.onsite + 0.04 * site_z + 0.03 * site_y, 0)):
        wv = build_wire(5, 4, 4, disorder=0.3, seed=8)
        if V is not None:
            wv.onsite[...] = V
            site = np.arange(wv.n) // wv.norb
            for x in range(wv.npx):
                wv.htotd[:, :, x] = wv.h0 + np.diag(V[x, site % wv.npy, site // wv.npy])
        print(f"potentials -> axis {axis}", flush=True)
        ref, noise = P.reference(wv, energies, 0.01, mu=MU, muT=MUT)
        acc, st, it, mirror = _sweep(w, energies, 0.01, compact=True, onsite=V)
        assert mirror[0] == axis, mirror
        assert np.all(st == 0), st
        P.check_all(_as_dict(acc), ref, noise)
        assert np.array_equal(np.asarray(it), ref["iters"])


def test_mirror_off_where_it_does_not_apply():
    w = build_wire(4, 4, 4)
    for flags in (capi.FLAG_SR_NO_MIRROR, capi.FLAG_SR_SEPARATE_LEADS):
        with capi.Context(4, 4, 4, 4, flags=flags) as ctx:
            ctx.set_hamiltonian(w)
            assert ctx.contact_mirror()[0] == 0
    w.htotd[:, :, -1] += 0.1 * np.eye(w.n)  # unequal end slices: separate decimations
    with capi.Context(4, 4, 4, 4) as ctx:
        ctx.set_hamiltonian(w)
        assert ctx.contact_mirror()[0] == 0
    wf = build_wire(4, 4, 4, flux=0.25)  # no mirror kept
    with capi.Context(4, 4, 4, 4) as ctx:
        ctx.set_hamiltonian(wf)
        assert ctx.contact_mirror()[0] == 0


@pytest.mark.parametrize("precision", [capi.NEGF_PREC_MIXED, capi.NEGF_PREC_LU32])
def test_mirror_reduced_forms_match_no_mirror(precision):
    """The reduced forms take the sectors too (3M products in the decimation): same
    iteration counts, and T12 within the forms' 1e-8 of the channel count."""
    w = build_wire(6, 8, 8, disorder=0.4, seed=12)
    energies = [-0.35, -0.1, 0.2]
    a1, s1, i1, m1 = _sweep(w, energies, 0.01, precision=precision)
    a2, s2, i2, m2 = _sweep(w, energies, 0.01, precision=precision, flags=capi.FLAG_SR_NO_MIRROR)
    assert m1[0] == 2 and m1[3] and m2[0] == 0
    assert np.all(s1 == 0) and np.all(s2 == 0)
    assert np.array_equal(np.asarray(i1), np.asarray(i2))
    scale = max(1.0, float(np.abs(a2.T12).max()))
    assert np.abs(a1.T12 - a2.T12).max() <= 1e-8 * scale, (a1.T12, a2.T12)
    for k in ("DOSc1", "DOSc2"):
        g, r = getattr(a1, k), getattr(a2, k)
        assert np.abs(g - r).max() <= 1e-9 * np.abs(r).max(), k
