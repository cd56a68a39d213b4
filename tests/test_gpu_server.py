"""GPU tests of the serving / batching layer (negf_server_*): whole solves of
different wires in flight together, admitted by HBM footprint.

* a request's outputs are bitwise identical alone, in a mix of other requests,
  and through the single-wire Context path (per-energy results do not depend on
  batching; each request accumulates in energy order);
* every request in the mix matches the complex128 oracle under the plain-FP64
  criterion of tests/parity.py;
* a budget that holds one resident Hamiltonian at a time queues the rest and
  still completes them with the same results;
* a request too large for a lane pool is rejected with NEGF_ERR_MEMORY."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

import parity as P  # noqa: E402

MU, MUT, DE = (-0.1, 0.0), 0.0, 0.01
FIELDS = capi.Accumulators.PER_ENERGY + capi.Accumulators.SPATIAL + ("ContactCurrent",)


def _requests():
    return [
        ("A", build_wire(6, 3, 2, disorder=0.6, seed=9, flux=0.2, zeeman=0.02), [-1.7, -0.2, 0.15, 1.2]),
        ("B", build_wire(9, 4, 4, disorder=0.5, seed=3), [-0.4, -0.1, 0.2, 0.4, 0.9]),
        ("C", build_wire(4, 2, 2, zeeman=0.03), [0.15, 0.9, 1.6]),
    ]


def _equal(a, b):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def _as_dict(acc):
    d = {k: getattr(acc, k) for k in P.PER_E + ("ContactCurrent",)}
    for k, a in (("RhoN", "RhoN"), ("RhoP", "RhoP"), ("SigX", "SigXExp"), ("SigY", "SigYExp"),
                 ("SigZ", "SigZExp")):
        d[k] = np.asarray(getattr(acc, a)).ravel(order="F")
    return d


def test_server_alone_equals_mix_and_matches_oracle():
    reqs = _requests()
    name, wA, eA = reqs[0]
    print("server: A alone", flush=True)
    with capi.Server() as srv:
        rid, alone = srv.submit(wA, eA, DE, MU, MUT)
        st, lat = srv.wait(rid)
        assert st == 0 and lat > 0
    print("server: B, A, C together", flush=True)
    with capi.Server() as srv:
        order = [reqs[1], reqs[0], reqs[2]]
        subs = [(nm, w, e) + srv.submit(w, e, DE, MU, MUT) for nm, w, e in order]
        out = {}
        for nm, w, e, rid, acc in subs:
            st, lat = srv.wait(rid)
            assert st == 0, (nm, st)
            out[nm] = (w, e, acc)
    _equal(alone, out["A"][2])
    # the single-wire path gives the same bits
    ctx_acc = capi.Accumulators(wA.npx, wA.npy, wA.npz, len(eA))
    with capi.Context(wA.npx, wA.npy, wA.npz, wA.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200) as ctx:
        ctx.set_hamiltonian(wA)
        ctx.submit(eA, np.arange(len(eA)), np.full(len(eA), DE))
        ctx.flush(MU, MUT, ctx_acc)
    _equal(alone, ctx_acc)
    for nm, (w, e, acc) in out.items():
        ref, noise = P.reference(w, e, DE, mu=MU, muT=MUT)
        P.check_all(_as_dict(acc), ref, noise)


def test_server_queues_by_hbm_footprint():
    # three block-256 wires with ~25 MB resident each (Htotd + couplings), lane pools
    # of 160 MB (two energies of ~76 MB) and 40 MB left for residents: they are
    # admitted one after another
    wires = [build_wire(20, 8, 8, disorder=0.5, seed=s) for s in (1, 2, 3)]
    energies = [-0.2, 0.1, 0.3]
    print("server: queued admissions", flush=True)
    with capi.Server(hbm_budget_gb=0.36, lane_pool_gb=0.16, lanes=2) as srv:
        subs = [srv.submit(w, energies, DE, MU, MUT) for w in wires]
        res = [srv.wait(rid) for rid, _ in subs]
    assert all(st == 0 for st, _ in res)
    # one at a time, first come first served: completion order follows submission
    lat = [l for _, l in res]
    assert lat[0] < lat[1] < lat[2]
    with capi.Server() as srv:
        rid, alone = srv.submit(wires[2], energies, DE, MU, MUT)
        srv.wait(rid)
    _equal(alone, subs[2][1])


def test_server_rejects_request_larger_than_a_lane_pool():
    w = build_wire(30, 12, 12)  # n = 576: one energy needs ~0.4 GB of lane pool
    with capi.Server(hbm_budget_gb=1.0, lane_pool_gb=0.05) as srv:
        with pytest.raises(capi.NegfError) as e:
            srv.submit(w, [0.1], DE, MU, MUT)
        assert e.value.code == capi.NEGF_ERR_MEMORY
