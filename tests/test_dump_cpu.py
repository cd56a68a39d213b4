"""Output files in the reference's formats (nonequilibriumgreensfunction_amd/dump.py):
ES11.4 fields as Dump_Data.F90:152 writes them, the loop orders and line counts of
Dump_Data.F90:6-139, a round trip through the NEGF_ReadData.m order, and the
full-precision archive."""
import os

import numpy as np

from nonequilibriumgreensfunction_amd.capi import Accumulators
from nonequilibriumgreensfunction_amd.dump import _es, dump_data, read_data


def test_es11_4_fields():
    assert _es(1.23456) == " 1.2346E+00"
    assert _es(-0.000123) == "-1.2300E-04"
    assert _es(0.0) == " 0.0000E+00"
    assert len(_es(-9.87654e12)) == 11


def _filled(dims, ne, rng):
    npx, npy, npz, _ = dims
    acc = Accumulators(npx, npy, npz, ne)
    for k in acc.PER_ENERGY:
        setattr(acc, k, rng.standard_normal(ne))
    for k in acc.SPATIAL:
        setattr(acc, k, np.asfortranarray(rng.standard_normal((npx, npy, npz))))
    acc.ContactCurrent = rng.standard_normal(2) * 1e14
    return acc


def test_dump_round_trip(tmp_path, rng):
    dims = (3, 2, 4, 4)
    ne = 5
    e = np.linspace(-0.3, 0.3, ne)
    acc = _filled(dims, ne, rng)
    J = [rng.standard_normal(dims[:3]) * 1e-6 for _ in range(3)]
    out = str(tmp_path / "run")
    dump_data(out, dims, e, acc, mu=(-0.1, 0.0), muT=0.0, currents=J, disorder=0.5)
    nsites = 3 * 2 * 4
    lines = lambda f: sum(1 for _ in open(os.path.join(out, f)))  # noqa: E731
    assert lines("scalars.txt") == 15
    assert lines("Greens.txt") == 3 * nsites
    assert lines("Transmissions.txt") == 2 * ne + ne + 2 * ne + ne + 2 + 3 * nsites + 2 * ne
    assert lines("SigExp.txt") == 3 * nsites
    r = read_data(out)
    assert r["dims"] == dims and r["eStpnum"] == ne
    close = lambda a, b: np.all(np.abs(np.asarray(a) - np.asarray(b)) <= 5e-5 * np.abs(b) + 1e-300)  # noqa: E731
    for k in ("T12", "Modec1", "Modec2", "DOSc1", "DOSc2", "EDensity", "HDensity"):
        assert close(r[k], getattr(acc, k)), k
    for k in ("RhoN", "RhoP", "SigXExp", "SigYExp", "SigZExp"):
        assert close(r[k], getattr(acc, k)), k
    assert np.allclose(r["EAxis"], e, rtol=0, atol=1e-15)
    assert np.allclose(r["ContactCurrent"], acc.ContactCurrent, rtol=1e-14)
    for i, k in enumerate(("Jx", "Jy", "Jz")):
        assert np.allclose(r[k], J[i], rtol=1e-14)
    assert np.all(r["DeltaR"] == 0)
    full = np.load(os.path.join(out, "results.npz"))
    for k in ("T12", "RhoN", "SigZExp", "ContactCurrent"):
        assert np.array_equal(full[k], np.asarray(getattr(acc, k))), k
