"""Host-side tight-binding nanowire Hamiltonians (the inputs of the per-energy solve).

Hamiltonian construction stays on the host, as in the reference
(Build_Hamiltonian.F90:1-150 is called on the root rank only, NEGF_Main.F90:15).
This module produces the same two arrays the reference's solve reads from
NEGF_Module (NEGF_Module.F90:56-57):

* ``Htotd`` -- complex128, shape (n, n, Npx), Fortran order: the on-slice blocks;
* ``HopX``  -- complex128, shape (n, n), Fortran order: the slice-to-slice block
  H_{x,x+1}; H_{x+1,x} = HopX^H (Recursive_Greens.F90:6).

with n = Npy*Npz*Norb and the basis index ``orb + Norb*(y + Npy*z)`` (0-based),
the ordering of Build_Hamiltonian.F90:66-95 and Calculate_Gnp.F90:122.

Models:

* ``ti4``: the 4-band cubic topological-insulator lattice model
  H(k) = M(k) G0 + A sum_i sin(k_i) G_i + (C + D sum_i 2(1-cos k_i)) I,
  M(k) = M0 + B sum_i 2(1 - cos k_i), in the Dirac-matrix basis of
  Read_In_Allocate.F90:161-184.  Its real-space form is an on-site block
  (M0 + 6B/a^2) G0 + (C + 6D/a^2) I and hops T_i = -(B/a^2) G0 - (D/a^2) I
  - (iA/2a) G_i along +i.  The reference's channelType 0
  (Build_Hamiltonian.F90:13-18, MM = 1.5) is the special case A=1, B=-1/2, M0=1.5.
* ``metal``: the single-orbital cubic metal of Build_Hamiltonian.F90:34-38
  (on-site 6t, hops -t).

Disorder, flux and Zeeman terms follow Build_Hamiltonian.F90: Anderson on-site
disorder (u-1/2)W on slices 2..Npx-1 only (:278-296, z fastest in the draw
order), the Aharonov-Bohm Peierls phase on y-hops (:186-207) and a Zeeman term
g*B*gMx on every site (the Insert_Magnetic_Impurity primitive, :348-360).  The
random numbers come from a seeded SplitMix64 stream, not the compiler's
RANDOM_NUMBER, so they are reproducible across compilers and machines.
"""
from __future__ import annotations

import dataclasses

import numpy as np

_I = 1j
# Dirac and spin matrices of Read_In_Allocate.F90:161-196 (rows as written there).
G1 = np.array([[0, 0, 0, -_I], [0, 0, _I, 0], [0, -_I, 0, 0], [_I, 0, 0, 0]], complex)
G2 = np.array([[0, 0, 0, -1], [0, 0, -1, 0], [0, -1, 0, 0], [-1, 0, 0, 0]], complex)
G3 = np.array([[0, 0, -_I, 0], [0, 0, 0, -_I], [_I, 0, 0, 0], [0, _I, 0, 0]], complex)
G0 = np.diag([1, 1, -1, -1]).astype(complex)
GMX = np.array([[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], complex)
GMY = np.array([[0, -_I, 0, 0], [_I, 0, 0, 0], [0, 0, 0, -_I], [0, 0, _I, 0]], complex)
GMZ = np.diag([1, -1, 1, -1]).astype(complex)


@dataclasses.dataclass
class Wire:
    """One transport instance: the arrays the solve reads plus its metadata."""

    npx: int
    npy: int
    npz: int
    norb: int
    htotd: np.ndarray  # (n, n, npx) complex128, Fortran order
    hopx: np.ndarray   # (n, n) complex128, Fortran order
    gm: np.ndarray     # (3, norb, norb) complex128: spin matrices gMx, gMy, gMz
    # compact description: htotd[:, :, x] = h0 + diag(onsite[x] on every orbital of its site)
    h0: np.ndarray | None = None      # (n, n) complex128 slice template (flux, Zeeman included)
    onsite: np.ndarray | None = None  # (npx, npy, npz) float64 site potentials, Fortran order

    @property
    def n(self) -> int:
        return self.npy * self.npz * self.norb

    def gm_fortran(self) -> np.ndarray:
        """gMx, gMy, gMz as three consecutive column-major norb x norb blocks."""
        return np.ascontiguousarray(np.stack([m.T for m in self.gm]))


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """``count`` doubles in [0, 1) from a SplitMix64 stream started at ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _model_blocks(model: str, params: dict):
    if model == "ti4":
        A = params.get("A", 1.0)
        B = params.get("B", 1.0)
        M0 = params.get("M0", -0.5)
        C = params.get("C", 0.0)
        D = params.get("D", 0.0)
        a = params.get("a", 1.0)
        eye = np.eye(4, dtype=complex)
        onsite = (M0 + 6.0 * B / a**2) * G0 + (C + 6.0 * D / a**2) * eye
        hops = [-(B / a**2) * G0 - (D / a**2) * eye - (1j * A / (2.0 * a)) * g for g in (G1, G2, G3)]
        return 4, onsite, hops[0], hops[1], hops[2], np.stack([GMX, GMY, GMZ])
    if model == "metal":
        t = params.get("t", 1.0)
        eps = params.get("eps", 6.0 * t)
        one = np.ones((1, 1), complex)
        return 1, eps * one, -t * one, -t * one, -t * one, np.stack([one, one * 0, one * 0])
    raise ValueError(f"unknown model {model!r}")


def build_wire(npx: int, npy: int, npz: int, model: str = "ti4", params: dict | None = None,
               bcy: int = 0, bcz: int = 0, disorder: float = 0.0, seed: int = 1,
               flux: float = 0.0, zeeman: float = 0.0, dense: bool = True) -> Wire:
    """Build ``Htotd`` and ``HopX`` for an Npx x Npy x Npz wire.

    ``disorder`` is the Anderson width W, ``flux`` the Aharonov-Bohm phase
    phi/phi0, ``zeeman`` the g*B prefactor of gMx.  The compact description
    (``h0``, ``onsite``) is always built; ``dense=False`` skips the n x n x Npx
    ``Htotd`` (None) for callers that upload the compact form.
    """
    norb, onsite, hx, hy, hz, gm = _model_blocks(model, params or {})
    NN = npy * norb
    n = npz * NN
    # slice block: y chains (Build_Hamiltonian.F90:66-79), then z layers (:81-95)
    hh1 = np.zeros((NN, NN), complex)
    for y in range(npy):
        hh1[y * norb:(y + 1) * norb, y * norb:(y + 1) * norb] = onsite
    for y in range(npy - 1):
        hh1[y * norb:(y + 1) * norb, (y + 1) * norb:(y + 2) * norb] = hy
        hh1[(y + 1) * norb:(y + 2) * norb, y * norb:(y + 1) * norb] = hy.conj().T
    if bcy and npy > 1:
        hh1[0:norb, NN - norb:NN] = hy.conj().T
        hh1[NN - norb:NN, 0:norb] = hy
    hopz = np.kron(np.eye(npy), hz)
    hh2 = np.zeros((n, n), complex)
    for z in range(npz):
        hh2[z * NN:(z + 1) * NN, z * NN:(z + 1) * NN] = hh1
    for z in range(npz - 1):
        hh2[z * NN:(z + 1) * NN, (z + 1) * NN:(z + 2) * NN] = hopz
        hh2[(z + 1) * NN:(z + 2) * NN, z * NN:(z + 1) * NN] = hopz.conj().T
    if bcz and npz > 1:
        hh2[0:NN, NN * (npz - 1):NN * npz] = hopz.conj().T
        hh2[NN * (npz - 1):NN * npz, 0:NN] = hopz
    hopx = np.kron(np.eye(npy * npz), hx)  # site-block-diagonal (Build_Hamiltonian.F90:62-64)
    h0 = hh2
    # Aharonov-Bohm phase on y-hops (Build_Hamiltonian.F90:186-207): the same in every slice
    if abs(flux) > 1e-10 and npy > 1 and npz > 1:
        for z in range(npz):
            ph = np.exp(1j * 2.0 * np.pi * z * flux / (npy - 1) / (npz - 1))
            for y in range(npy - 1):
                i0 = y * norb + z * NN
                blk = h0[i0:i0 + norb, i0 + norb:i0 + 2 * norb] * ph
                h0[i0:i0 + norb, i0 + norb:i0 + 2 * norb] = blk
                h0[i0 + norb:i0 + 2 * norb, i0:i0 + norb] = np.conj(blk.T)
    # Zeeman g*B*gMx on every site (Insert_Magnetic_Impurity, Build_Hamiltonian.F90:348-360);
    # gMx has a zero diagonal, so it commutes bitwise with the potentials added below
    if zeeman != 0.0 and norb == 4:
        h0 = h0 + np.kron(np.eye(npy * npz), zeeman * GMX)
    # Anderson disorder on slices 2..Npx-1 (Build_Hamiltonian.F90:278-296), z fastest
    onsite = np.zeros((npx, npy, npz), order="F")
    if abs(disorder) > 1e-20 and npx > 2:
        u = splitmix64_uniform(seed, (npx - 2) * npy * npz).reshape(npx - 2, npy, npz)
        onsite[1:npx - 1] = (u - 0.5) * disorder
    H = None
    if dense:
        H = np.empty((n, n, npx), complex, order="F")
        H[:, :, :] = h0[:, :, None]
        a = np.arange(n)
        site = a // norb
        H[a, a, :] += onsite[:, site % npy, site // npy].T
    return Wire(npx, npy, npz, norb, H, np.asfortranarray(hopx), gm, np.asfortranarray(h0), onsite)


def energy_grid(emin: float, emax: float, count: int) -> np.ndarray:
    """Uniform grid with inclusive endpoints (the rule of Read_In_Allocate.F90:255-259)."""
    if count == 1:
        return np.array([emin], float)
    return emin + (emax - emin) / (count - 1) * np.arange(count, dtype=float)
