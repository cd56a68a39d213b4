"""NEGF_FLAG_GRAPH: each batch's launches after the contacts are captured into a
HIP graph once per (lane, context, Hamiltonian upload, batch size) and replayed.
Every replay -- with the eager Sancho-Rubio work, host synchronisations and new
site potentials between replays -- must give bitwise the results of the same
batches launched one by one, in both precision forms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

MU, MUT, DE = (-0.1, 0.0), 0.0, 0.01
FIELDS = (capi.Accumulators.PER_ENERGY + capi.Accumulators.SPATIAL + capi.Accumulators.CURRENTS
          + ("ContactCurrent",))
ENERGIES = [-1.7, -1.1, -0.3, 0.05, 0.4, 0.9, 1.3, 1.8, -0.6, 0.2]  # 10: chunks of 3, 3, 3, 1


def _ctx(w, precision, flags):
    ctx = capi.Context(w.npx, w.npy, w.npz, w.norb, eta=1e-6, sr_tol=1e-10, sr_max_iter=200,
                       precision=precision, flags=flags, max_batch=3, lanes=1)
    ctx.set_hamiltonian_compact(w)
    return ctx


def _sweep(ctx, w):
    acc = capi.Accumulators(w.npx, w.npy, w.npz, len(ENERGIES))
    ctx.submit(ENERGIES, np.arange(len(ENERGIES)), np.full(len(ENERGIES), DE))
    st, it = ctx.flush(MU, MUT, acc)
    assert np.all(np.asarray(st) == 0)
    return acc


@pytest.mark.parametrize("precision", [capi.NEGF_PREC_FP64, capi.NEGF_PREC_MIXED], ids=["fp64", "mixed"])
def test_graph_replays_equal_eager(precision):
    w = build_wire(7, 2, 3, disorder=0.6, seed=17, flux=0.3, zeeman=0.02)
    rng = np.random.default_rng(3)
    gctx = _ctx(w, precision, capi.FLAG_GRAPH)
    try:
        v = w.onsite.copy()
        for rep in range(3):
            print(f"graph replay round {rep} precision={precision}", flush=True)
            if rep:
                v[1:-1] += 0.05 * rng.standard_normal(v[1:-1].shape)  # eager upload between replays
                gctx.set_onsite(v)
            got = _sweep(gctx, w)
            w2 = build_wire(7, 2, 3, disorder=0.6, seed=17, flux=0.3, zeeman=0.02)
            w2.onsite = v.copy()
            with _ctx(w2, precision, 0) as ectx:
                ref = _sweep(ectx, w2)
            for k in FIELDS:
                assert np.array_equal(getattr(got, k), getattr(ref, k)), (rep, k)
        built, replays = gctx.graph_stats()
        # chunks of 3 and a last one of 1 per sweep: two graphs, every other chunk replayed
        assert built == 2, built
        assert replays == 3 * 4, replays
    finally:
        gctx.close()
