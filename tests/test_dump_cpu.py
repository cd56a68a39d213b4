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


def test_es11_4_matches_the_fortran_runtime(tmp_path, rng):
    # the text fields against a Fortran program's own `FORMAT(ES11.4)` output (the
    # format of Dump_Data.F90:152), written by amdflang's runtime
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        import pytest
        pytest.skip("no Fortran compiler")
    vals = np.concatenate([rng.standard_normal(20) * 10.0 ** rng.integers(-30, 30, 20),
                           [0.0, 1.0, -1.0, 9.99995e-3, 1.23455e-7, -7.65435e12, 5e-310, -2.5e-300,
                            1.5e300, np.inf, -np.inf, np.nan]])
    src = tmp_path / "w.f90"
    src.write_text("program w\n  real(8) :: v\n  integer :: i\n"
                   "  open(10, file='in.txt'); open(13, file='out.txt')\n"
                   f"  do i = 1, {len(vals)}\n    read(10, *) v\n    write(13, 43) v\n  end do\n"
                   "  close(13)\n43 format(ES11.4)\nend program w\n")
    (tmp_path / "in.txt").write_text("\n".join(repr(float(v)) for v in vals) + "\n")
    subprocess.run([fc, str(src), "-o", str(tmp_path / "w")], check=True, capture_output=True)
    subprocess.run([str(tmp_path / "w")], check=True, cwd=tmp_path, capture_output=True)
    fortran = (tmp_path / "out.txt").read_text().splitlines()
    assert fortran == [_es(v) for v in vals]


def test_resolved_files_round_trip(tmp_path, rng):
    dims = (2, 3, 2, 4)
    ne = 3
    npx, npy, npz, no = dims
    acc = _filled(dims, ne, rng)
    acc2 = Accumulators(npx, npy, npz, ne, resolved=True, norb=no)
    for k in acc2.RESOLVED + acc2.CURRENTS:
        a = getattr(acc2, k)
        a[...] = rng.standard_normal(a.shape)
        setattr(acc, k, a)
    out = str(tmp_path / "res")
    phi = rng.standard_normal((npx, npy, npz))
    dump_data(out, dims, np.linspace(-0.1, 0.1, ne), acc, phi=phi, ro=phi * 2)
    ns = npx * npy * npz
    assert sum(1 for _ in open(os.path.join(out, "ERCD.txt"))) == 4 * ne * ns
    assert sum(1 for _ in open(os.path.join(out, "ORCD.txt"))) == (3 * no * no + no) * ns
    assert sum(1 for _ in open(os.path.join(out, "Poissons.txt"))) == 2 * ns
    r = read_data(out)
    for k in acc2.RESOLVED + acc2.CURRENTS:
        b = getattr(acc, k)
        assert np.all(np.abs(r[k] - b) <= 5e-5 * np.abs(b) + 1e-300), k
    assert np.all(np.abs(r["phi"] - phi) <= 5e-5 * np.abs(phi))
    # the accumulators' currents were used (no explicit currents argument)
    assert np.all(r["Jy"] != 0)
