"""Writes tests/golden/negf_golden.npz: small fixtures of the per-energy solve.

For three seeded wires (clean; disorder + flux + Zeeman; strong disorder) at one
energy each, the file holds the inputs (Htotd, HopX, gM, InvGRd, the contact
Sigma/Gamma) and
  * GRLef, GRRgt, GRd from the reference's own Get_Retarded_Greens and
    Gnd, RhoN, RhoP from its Calculate_Gnp (f1 = 1, f2 = 0), both compiled
    unmodified from the reference sources into oracle/_ref/libnegf_ref.so;
  * the complex128 oracle's Energy_Sweep outputs for the same energy (T12,
    Modec1/2, DOSc1/2, E/HDensity, RhoN, RhoP, spin, bond currents,
    ContactCurrent).
The reference cannot run on the GPU machine; these vectors travel instead.
Run from the repository root after `make -C oracle`:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from nonequilibriumgreensfunction_amd.hamiltonian import build_wire  # noqa: E402

ETA = 1e-6
CASES = [
    ("clean", (4, 2, 2), dict(), 1.6),
    ("dfz", (5, 2, 3), dict(disorder=0.4, seed=6, flux=0.2, zeeman=0.02), 0.2),
    ("strong", (3, 3, 2), dict(disorder=0.9, seed=1), -1.7),
]


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref/libnegf_ref.so is not built (make -C oracle where the reference exists)")
    out = {}
    for name, dims, kw, E in CASES:
        w = build_wire(*dims, **kw)
        c = O.contact_self_energy(w, E, ETA)
        A = O.assemble_invg(w, c["SigmaL"], c["SigmaR"], E, ETA)
        r = O.retarded(A, w.hopx, use_ref=True)
        g = O.gnp(w, 1.0, 0.0, 0.01, r["GRLef"], r["GRRgt"], c["GammaL"], c["GammaR"], use_ref=True)
        s = O.energy_sweep(w, E, 0.01, mu=(-2.0, 0.0), muT=0.0, eta=ETA)
        p = name + "/"
        out[p + "dims"] = np.array(dims + (w.norb,))
        out[p + "energy"] = np.array([E])
        out[p + "htotd"] = w.htotd
        out[p + "hopx"] = w.hopx
        out[p + "gm"] = w.gm
        out[p + "invgrd"] = A
        for k in ("SigmaL", "SigmaR", "GammaL", "GammaR"):
            out[p + k] = c[k]
        for k in ("GRLef", "GRRgt", "GRd"):
            out[p + "ref_" + k] = r[k]
        for k in ("Gnd", "RhoN", "RhoP"):
            out[p + "ref_" + k] = g[k]
        for k in ("T12", "Modec1", "Modec2", "DOSc1", "DOSc2", "EDensity", "HDensity"):
            out[p + "sweep_" + k] = np.array([s[k]])
        for k in ("RhoN", "RhoP", "SigX", "SigY", "SigZ", "Jx", "Jy", "Jz", "ContactCurrent"):
            out[p + "sweep_" + k] = np.asarray(s[k])
    path = os.path.join(ROOT, "tests", "golden", "negf_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
