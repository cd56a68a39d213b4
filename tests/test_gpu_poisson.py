"""GPU tests of the Poisson-iteration support: negf_set_onsite (new site
potentials on a compact Hamiltonian) and NEGF_FLAG_CACHE_CONTACTS (each energy's
contact self-energies reused while HopX and the end slices are unchanged).
Every sweep must equal, bit for bit, a fresh context built with the same
Hamiltonian and no cache; the cache statistics show when it was used."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

MU, MUT, DE = (-0.1, 0.0), 0.0, 0.01
FIELDS = capi.Accumulators.PER_ENERGY + capi.Accumulators.SPATIAL + ("ContactCurrent",)
ENERGIES = [-1.7, -0.2, 0.15, 1.2]
RTOL_LOOP = 3e-8


def _fresh(w, onsite):
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(ENERGIES))
    v = w.onsite
    w.onsite = onsite
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200) as ctx:
        ctx.set_hamiltonian_compact(w)
        ctx.submit(ENERGIES, np.arange(len(ENERGIES)), np.full(len(ENERGIES), DE))
        st, it = ctx.flush(MU, MUT, acc)
    w.onsite = v
    return acc, np.asarray(it)


def _sweep(ctx, w):
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(ENERGIES))
    ctx.submit(ENERGIES, np.arange(len(ENERGIES)), np.full(len(ENERGIES), DE))
    st, it = ctx.flush(MU, MUT, acc)
    assert np.all(np.asarray(st) == 0)
    return acc, np.asarray(it)


def _equal(a, b):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_poisson_iterations_reuse_contacts_and_match_fresh():
    w = build_wire(7, 3, 2, disorder=0.5, seed=11, flux=0.2, zeeman=0.02)
    rng = np.random.default_rng(5)
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                      flags=capi.FLAG_CACHE_CONTACTS) as ctx:
        ctx.set_hamiltonian_compact(w)
        a0, it0 = _sweep(ctx, w)
        ref, rit = _fresh(w, w.onsite)
        _equal(a0, ref)
        assert ctx.contact_cache_stats() == (len(ENERGIES), 0)
        # "Poisson update": new interior potential, end slices untouched
        for k in range(2):
            v = w.onsite.copy()
            v[1:-1] += 0.05 * rng.standard_normal(v[1:-1].shape)
            ctx.set_onsite(v)
            a, it = _sweep(ctx, w)
            ref, rit = _fresh(w, v)
            _equal(a, ref)
            assert np.array_equal(it, rit)
            assert ctx.contact_cache_stats() == (len(ENERGIES), (k + 1) * len(ENERGIES))
        # a potential on the first slice changes the left lead: the cache is dropped
        v = v.copy()
        v[0] += 0.1
        ctx.set_onsite(v)
        assert ctx.contact_cache_stats() == (0, 0)
        a, it = _sweep(ctx, w)
        ref, rit = _fresh(w, v)
        _equal(a, ref)
        assert ctx.contact_cache_stats() == (len(ENERGIES), 0)


def test_dense_reload_with_same_leads_keeps_contacts():
    w = build_wire(6, 2, 3, disorder=0.6, seed=2)
    with capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                      flags=capi.FLAG_CACHE_CONTACTS) as ctx:
        ctx.set_hamiltonian(w)
        _sweep(ctx, w)
        w2 = build_wire(6, 2, 3, disorder=0.6, seed=3)  # new interior disorder, same clean ends
        ctx.set_hamiltonian(w2)
        a, it = _sweep(ctx, w2)
        assert ctx.contact_cache_stats() == (len(ENERGIES), len(ENERGIES))
    ref, rit = _fresh(w2, w2.onsite)
    _equal(a, ref)


def test_self_consistent_loop_matches_oracle():
    """The host loop of NEGF_Main.F90:7-30 -- potential update, Hamiltonian with the
    potential on every site, RhoN = RhoP = 0, energy sweep -- run with the GPU
    sweep (compact Hamiltonian, new site potentials per iteration, contact
    self-energies reused) and with the complex128 oracle's sweep on the dense
    Htotd built from the same potentials.  Every round's densities and potentials
    agree to 3e-8 relative: the energies sit in the bands, where ~23 decimation
    steps amplify rounding and the oracle itself moves by ~3e-9 under a change of
    summation order (tests/parity.py); the update carries that difference
    forward linearly.  The contacts were solved once."""
    import parity as P
    from nonequilibriumgreensfunction_amd import poisson

    w = build_wire(6, 2, 3, disorder=0.6, seed=21, flux=0.25, zeeman=0.02)
    energies = np.linspace(-1.9, 1.5, 6)
    dE = float(energies[1] - energies[0])
    iters = 3
    update = poisson.linear_screening(strength=10.0, mix=0.6)
    print("self-consistent loop: GPU sweeps", flush=True)
    phi_g, rn_g, rp_g, hist_g, stats = poisson.self_consistent_gpu(
        w, energies, dE, update, iters, MU, MUT, eta=1e-6, sr_tol=1e-10, sr_max_iter=200)
    assert stats == (len(energies), (iters - 1) * len(energies))

    print("self-consistent loop: oracle sweeps", flush=True)
    phi = np.zeros((w.npx, w.npy, w.npz), order="F")
    a = np.arange(w.n)
    site = a // w.norb
    for k in range(iters):
        v = w.onsite + phi
        wo = build_wire(6, 2, 3, disorder=0.6, seed=21, flux=0.25, zeeman=0.02)
        wo.htotd = np.asfortranarray(np.repeat(w.h0[:, :, None], w.npx, axis=2))
        wo.htotd[a, a, :] += v[:, site % w.npy, site // w.npy].T
        r = P.oracle_sweep(wo, list(energies), dE, MU, MUT)
        rn = r["RhoN"].reshape((w.npx, w.npy, w.npz), order="F")
        rp = r["RhoP"].reshape((w.npx, w.npy, w.npz), order="F")
        # the same iteration on both sides: the GPU's densities of round k match
        for got, ref in ((hist_g[k][1], rn), (hist_g[k][2], rp)):
            assert np.abs(got - ref).max() <= RTOL_LOOP * np.abs(ref).max(), (k, np.abs(got - ref).max())
        assert np.abs(hist_g[k][0] - phi).max() <= RTOL_LOOP * max(np.abs(phi).max(), 1e-12), k
        phi = update(phi, rn, rp)
    assert np.abs(phi_g - phi).max() <= RTOL_LOOP * np.abs(phi).max()
    assert np.abs(phi).max() > 1e-3  # the potential did move the wire
