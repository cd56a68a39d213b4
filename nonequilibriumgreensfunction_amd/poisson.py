"""Host-side self-consistent loop around the per-energy solve.

The reference's outer loop (NEGF_Main.F90:7-30) alternates, on the host, a
potential update (SIP_Main, the Poisson solver) with a full energy sweep:

    DO ite = 1, iteMax
        CALL SIP_Main              ! Phi from RhoN, RhoP          (NEGF_Main.F90:12)
        CALL Build_Hamiltonian     ! Phi inserted on every site   (Build_Hamiltonian.F90:104-112)
        RhoN = 0; RhoP = 0
        DO eCtr: CALL Energy_Sweep ! accumulates RhoN, RhoP
    END DO

Only the diagonal site potential changes between iterations, so here the
device keeps the Hamiltonian as its compact description (slice template plus
site potentials) and each iteration uploads just the new potentials
(negf_set_onsite); with NEGF_FLAG_CACHE_CONTACTS the contact self-energies of
every energy are computed once and reused while the end slices are unchanged.

The Poisson solver itself (SIP_*.F90) stays with the host program: `update` is
any callable mapping (Phi, RhoN, RhoP) to the next Phi.  `linear_screening` is
a small stand-in used by the tests and the example: a damped linear response
to the net charge, zero on the two end slices (the gates that hold the
contacts' potential fixed, SIP_Routines.F90:60-65).
"""
from __future__ import annotations

import numpy as np

from . import capi


def linear_screening(strength: float = 20.0, mix: float = 0.5):
    """Phi <- (1 - mix) Phi + mix * (-strength) * (net - <net>), net = RhoN - RhoP,
    <.> the mean over the interior slices; the end slices stay at 0."""

    def update(phi, rho_n, rho_p):
        net = np.asarray(rho_n) - np.asarray(rho_p)
        inner = net[1:-1]
        target = np.zeros_like(phi)
        target[1:-1] = -strength * (inner - inner.mean())
        return (1.0 - mix) * phi + mix * target

    return update


def self_consistent_gpu(wire, energies, deltaE, update, iterations, mu=(-0.1, 0.0), muT=0.0,
                        precision=capi.NEGF_PREC_FP64, **ctx_kw):
    """Runs `iterations` rounds of sweep + potential update on the GPU.

    Returns (Phi, RhoN, RhoP, history, cache_stats): Phi the potential after the
    last update, RhoN/RhoP of the last sweep, history the per-iteration
    (Phi, RhoN, RhoP) and cache_stats the contact cache's (entries, hits)."""
    e = np.asarray(energies, float)
    phi = np.zeros((wire.npx, wire.npy, wire.npz), order="F")
    history = []
    flags = ctx_kw.pop("flags", 0) | capi.FLAG_CACHE_CONTACTS
    with capi.Context(wire.npx, wire.npy, wire.npz, wire.norb, precision=precision, flags=flags,
                      **ctx_kw) as ctx:
        ctx.set_hamiltonian_compact(wire)
        for _ in range(iterations):
            ctx.set_onsite(wire.onsite + phi)  # Build_Hamiltonian with Phi on every site
            acc = capi.Accumulators(wire.npx, wire.npy, wire.npz, len(e))  # RhoN = RhoP = 0
            ctx.submit(e, np.arange(len(e)), np.full(len(e), deltaE))
            st, _ = ctx.flush(mu, muT, acc)
            if np.any(np.asarray(st) != 0):
                raise capi.NegfError(int(np.asarray(st)[np.nonzero(st)[0][0]]),
                                     "an energy of the sweep reported a nonzero status")
            rho_n, rho_p = acc.RhoN.copy(), acc.RhoP.copy()
            history.append((phi.copy(), rho_n, rho_p))
            phi = update(phi, rho_n, rho_p)
        stats = ctx.contact_cache_stats()
    return phi, rho_n, rho_p, history, stats
