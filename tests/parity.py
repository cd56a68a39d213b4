"""Shared parity criterion for the GPU tests.

The plain FP64 form must agree with the complex128 oracle to a relative 1e-10
of each output's natural scale -- unless the problem itself is less stable than
that.  At energies where Sancho-Rubio needs many decimation steps (~2^K
amplification of rounding at eta = 1e-6) the oracle disagrees with itself at
the 1e-9 level when only the order of its floating-point sums changes.  That
spread is measured by solving the physically identical problem with the
orbitals permuted inside every site (H' = Q H Q^T, sigma' = P sigma P^T), with
two different permutations: every output is invariant, every GEMM and LU
summation order changes.  All outputs of one energy inherit the conditioning of
that energy's contact self-energies, so the spread is taken per energy as the
largest relative spread over that energy's outputs (each relative to its
natural scale), and for each energy-summed site array as its largest relative
spread over all sites.  The criterion is, per output element,

    |GPU - oracle| <= max(RTOL, FLOOR * spread) * scale

with RTOL = 1e-10 and FLOOR = 10: the GPU must be as close to the oracle as the
oracle is to itself under a different summation order.
"""
import numpy as np

import oracle as O

RTOL = 1e-10
FLOOR = 10.0
PER_E = ("T12", "Modec1", "Modec2", "DOSc1", "DOSc2", "EDensity", "HDensity")
SPATIAL = ("RhoN", "RhoP", "SigX", "SigY", "SigZ", "Jx", "Jy", "Jz")


def oracle_sweep(w, energies, dE, mu=(-0.1, 0.0), muT=0.0, eta=1e-6, tol=1e-10, maxit=200,
                 threads=None):
    """Oracle Energy_Sweep over `energies`: per-energy arrays, summed spatial arrays
    (flattened in Fortran order), ContactCurrent and SR iteration counts."""
    if threads is not None:
        O.set_threads(threads)
    acc = {k: [] for k in PER_E}
    sp = {k: 0.0 for k in SPATIAL}
    cc = np.zeros(2)
    iters = []
    for E in energies:
        r = O.energy_sweep(w, E, dE, mu, muT, eta, tol, maxit, True)
        for k in PER_E:
            acc[k].append(r[k])
        for k in SPATIAL:
            sp[k] = sp[k] + np.asarray(r[k]).ravel(order="F")
        cc += r["ContactCurrent"]
        iters.append(r["iters"])
    out = {k: np.array(v) for k, v in acc.items()}
    out.update(sp)
    out["ContactCurrent"] = cc
    out["iters"] = np.array(iters)
    if threads is not None:
        O.set_threads(8)
    return out


PERMS = {1: ([0], [0]), 2: ([1, 0], [1, 0]), 4: ([2, 0, 3, 1], [3, 2, 1, 0])}


def permuted(w, which=0):
    """The same wire with the orbitals of every site reordered."""
    import copy
    perm = PERMS[w.norb][which]
    nsite = w.n // w.norb
    idx = np.concatenate([s * w.norb + np.array(perm) for s in range(nsite)])
    v = copy.copy(w)
    v.htotd = np.asfortranarray(w.htotd[np.ix_(idx, idx, np.arange(w.npx))])
    v.hopx = np.asfortranarray(w.hopx[np.ix_(idx, idx)])
    v.gm = np.stack([m[np.ix_(perm, perm)] for m in w.gm])
    return v


def scales(ref):
    """Natural scale of each output: the open-channel count (at least 1) for T12 and
    the modes, the density for spin, the output's own maximum otherwise."""
    nm = max(1.0, float(np.abs(ref["Modec1"]).max()))
    rho = float(np.abs(ref["RhoN"]).max())
    sc = {}
    for k in PER_E + SPATIAL + ("ContactCurrent",):
        sc[k] = max(float(np.abs(np.asarray(ref[k])).max(initial=0.0)), 1e-300)
    for k in ("T12", "Modec1", "Modec2"):
        sc[k] = nm
    for k in ("SigX", "SigY", "SigZ"):
        sc[k] = max(sc[k], rho)
    return sc


def reference(w, energies, dE, nperm=2, **kw):
    """(oracle result, absolute noise floor per output element); the floor is the
    spread over `nperm` (1 or 2) orbital permutations of the wire."""
    a = oracle_sweep(w, energies, dE, **kw)
    others = [oracle_sweep(permuted(w, i), energies, dE, **kw) for i in range(nperm)]
    sc = scales(a)
    ne = len(energies)
    rel_e = np.zeros(ne)
    rel_sp = {k: 0.0 for k in SPATIAL + ("ContactCurrent",)}
    for b in others:
        for k in PER_E:
            rel_e = np.maximum(rel_e, np.abs(np.asarray(a[k]) - np.asarray(b[k])) / sc[k])
        for k in rel_sp:
            rel_sp[k] = max(rel_sp[k], float(np.abs(np.asarray(a[k]) - np.asarray(b[k])).max()) / sc[k])
    noise = {k: rel_e * sc[k] for k in PER_E}
    for k in rel_sp:
        noise[k] = np.full(np.asarray(a[k]).shape, rel_sp[k] * sc[k])
    return a, noise


def check(name, got, ref, noise, scale=0.0):
    got, ref = np.asarray(got).ravel(), np.asarray(ref).ravel()  # real or complex
    noise = np.broadcast_to(np.asarray(noise, float).ravel(), ref.shape)
    lim = np.maximum(RTOL * max(np.abs(ref).max(initial=0.0), scale), FLOOR * noise)
    err = np.abs(got - ref)
    i = int(np.argmax(err - lim))
    assert err[i] <= lim[i], (f"{name}[{i}]: err {err[i]:.3e} > limit {lim[i]:.3e} "
                              f"(ref {abs(ref[i]):.6e}, oracle noise {noise[i]:.3e})")


def check_all(got, ref, noise, jx=False):
    """got: dict with the PER_E keys, RhoN..SigZ (flattened F order) and ContactCurrent."""
    nm = max(1.0, float(np.abs(ref["Modec1"]).max()))
    for k in ("T12", "Modec1", "Modec2"):
        check(k, got[k], ref[k], noise[k], nm)
    for k in ("DOSc1", "DOSc2", "EDensity", "HDensity", "RhoN", "RhoP", "ContactCurrent"):
        check(k, got[k], ref[k], noise[k])
    # |Tr(Gn_site sigma)| <= Tr(Gn_site): spin is checked on the density's scale
    rho = float(np.abs(ref["RhoN"]).max())
    for k in ("SigX", "SigY", "SigZ"):
        check(k, got[k], ref[k], noise[k], rho)
