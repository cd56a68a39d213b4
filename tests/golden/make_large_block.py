"""Writes tests/golden/large_block_n1088.npz: the complex128 oracle's Energy_Sweep
outputs, and their noise floor, for the block-1088 wire of test_gpu_large_block.py.

Wire: build_wire(3, 16, 17, disorder=0.4, seed=5, flux=0.3), one energy -0.2,
deltaE 0.01, mu (-0.1, 0.0), muT 0, eta 1e-6, Sancho-Rubio tol 1e-10.  The noise
floor is the oracle's spread over the same wire with the orbitals of every site
permuted (tests/parity.py).  Three oracle sweeps at this size take about a
minute of CPU; the fixture spares the GPU run that.  Run from the repository root
after `make -C oracle`:
    python tests/golden/make_large_block.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import parity as P  # noqa: E402
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "large_block_n1088.npz")
WIRE = dict(npx=3, npy=16, npz=17, disorder=0.4, seed=5, flux=0.3)
ENERGIES = [-0.2]
DE = 0.01
MU, MUT = (-0.1, 0.0), 0.0


def wire_digest(w):
    """sha256 of the wire's Htotd and HopX bytes (column-major)."""
    h = hashlib.sha256()
    h.update(np.asfortranarray(w.htotd).tobytes(order="F"))
    h.update(np.asfortranarray(w.hopx).tobytes(order="F"))
    return h.hexdigest()


def main():
    w = build_wire(WIRE["npx"], WIRE["npy"], WIRE["npz"], disorder=WIRE["disorder"],
                   seed=WIRE["seed"], flux=WIRE["flux"])
    ref, noise = P.reference(w, ENERGIES, DE, mu=MU, muT=MUT)
    out = {"energies": np.array(ENERGIES), "wire_sha256": np.array(wire_digest(w))}
    for k, v in ref.items():
        out["ref_" + k] = np.asarray(v)
    for k, v in noise.items():
        out["noise_" + k] = np.asarray(v)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
