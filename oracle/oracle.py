"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/_build/libnegf_oracle.so`` (the C++ complex128 restatement of
the reference's Energy_Sweep path, negf_oracle.cpp) and, where it was built,
``oracle/_ref/libnegf_ref.so`` (the reference's own Get_Retarded_Greens,
Calculate_Gnp and Invert_Matrix compiled from its sources).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this
module; the product library never does.

Arrays are numpy complex128 in Fortran order, matching NEGF_Module's layouts.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

SR_NOT_CONVERGED = 1 << 20
SR_NAN = 1 << 21


def _p(a):
    return a.ctypes.data_as(_dp)


def lib():
    global _LIB
    if _LIB is None:
        # NEGF_ORACLE_MKL=sequential: the build linked with MKL sequential, as the
        # reference links it (Implement/Makefile:3), for one-rank-per-core runs
        seq = os.environ.get("NEGF_ORACLE_MKL", "") == "sequential"
        path = os.path.join(_HERE, "_build", "libnegf_oracle_seq.so" if seq else "libnegf_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        _LIB = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        _LIB.oracle_zgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp]
        _LIB.oracle_sancho_rubio.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip]
        _LIB.oracle_contact_self_energy.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip]
        _LIB.oracle_assemble_invg.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                              ctypes.c_double, ctypes.c_double, _dp]
        _LIB.oracle_retarded.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _LIB.oracle_gnp.argtypes = ([ctypes.c_int] * 5 + [ctypes.c_double] * 3 + [_dp] * 12)
        _LIB.oracle_gnp.restype = None
        _LIB.oracle_observables.argtypes = ([ctypes.c_int] * 5 + [ctypes.c_double] * 3 + [_dp] * 20)
        _LIB.oracle_observables.restype = None
        _LIB.oracle_energy_sweep.argtypes = ([ctypes.c_int] * 4 + [_dp] * 3 + [ctypes.c_double] * 7
                                             + [ctypes.c_int] * 2 + [_dp] * 10 + [_ip])
        _LIB.oracle_invert_matrix.argtypes = [ctypes.c_int, _dp, _dp]
        _LIB.oracle_set_threads.argtypes = [ctypes.c_int]
        _LIB.oracle_set_threads.restype = None
    return _LIB


def ref():
    """The reference-compiled library, or None where it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libnegf_ref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        _REF.ref_retarded.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 7
        _REF.ref_gnp.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double] * 3 + [_dp] * 12
        _REF.ref_gnp.restype = None
        _REF.ref_invert.argtypes = [ctypes.c_int, _dp, _dp]
    return _REF


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def _c(a):
    return np.asfortranarray(a, dtype=np.complex128)


def invert(G):
    G = _c(G)
    out = np.zeros_like(G)
    info = lib().oracle_invert_matrix(G.shape[0], _p(G), _p(out))
    return out, info


def zgemm(ta, tb, A, B, alpha=1.0, beta=0.0, C=None):
    A, B = _c(A), _c(B)
    n = A.shape[0]
    C = np.zeros((n, n), complex, order="F") if C is None else _c(C).copy(order="F")
    alpha, beta = complex(alpha), complex(beta)
    lib().oracle_zgemm(ta.encode(), tb.encode(), n, alpha.real, alpha.imag, _p(A), _p(B),
                       beta.real, beta.imag, _p(C))
    return C


def sancho_rubio(alpha, beta, eta, tol=1e-10, max_iter=200, extra_eta=True):
    alpha, beta = _c(alpha), _c(beta)
    n = alpha.shape[0]
    sig = np.zeros((n, n), complex, order="F")
    it = ctypes.c_int(0)
    st = lib().oracle_sancho_rubio(n, _p(alpha), _p(beta), _p(sig), eta, tol, max_iter,
                                   int(bool(extra_eta)), ctypes.byref(it))
    return sig, st, it.value


def contact_self_energy(wire, energy, eta, tol=1e-10, max_iter=200, extra_eta=True):
    n = wire.n
    H = wire.htotd
    h1 = np.asfortranarray(H[:, :, 0])
    hl = np.asfortranarray(H[:, :, -1])
    out = [np.zeros((n, n), complex, order="F") for _ in range(4)]
    dos = np.zeros(2)
    iters = (ctypes.c_int * 2)()
    st = lib().oracle_contact_self_energy(n, _p(h1), _p(hl), _p(wire.hopx), energy, eta, tol,
                                          max_iter, int(bool(extra_eta)), *[_p(o) for o in out],
                                          _p(dos), iters)
    return dict(SigmaL=out[0], SigmaR=out[1], GammaL=out[2], GammaR=out[3], DOSc1=dos[0],
                DOSc2=dos[1], iters=(iters[0], iters[1]), status=st)


def assemble_invg(wire, sigL, sigR, energy, eta):
    n, L = wire.n, wire.npx
    out = np.zeros((n, n, L), complex, order="F")
    lib().oracle_assemble_invg(n, L, _p(wire.htotd), _p(_c(sigL)), _p(_c(sigR)), energy, eta,
                               _p(out))
    return out


def _rgf_outputs(n, L):
    Lm = max(L - 1, 1)
    return dict(GRd=np.zeros((n, n, L), complex, order="F"),
                GRu=np.zeros((n, n, Lm), complex, order="F"),
                GRl=np.zeros((n, n, Lm), complex, order="F"),
                GRLef=np.zeros((n * L, n), complex, order="F"),
                GRRgt=np.zeros((n * L, n), complex, order="F"))


def retarded(invgrd, hopx, use_ref=False):
    invgrd, hopx = _c(invgrd), _c(hopx)
    n, L = invgrd.shape[0], invgrd.shape[2]
    o = _rgf_outputs(n, L)
    f = ref().ref_retarded if use_ref else lib().oracle_retarded
    st = f(n, L, _p(invgrd), _p(hopx), _p(o["GRd"]), _p(o["GRu"]), _p(o["GRl"]), _p(o["GRLef"]),
           _p(o["GRRgt"]))
    o["status"] = st
    return o


def gnp(wire, f1, f2, deltaE, GRLef, GRRgt, GammaL, GammaR, use_ref=False):
    n, L = wire.n, wire.npx
    Lm = max(L - 1, 1)
    o = {k: np.zeros((n, n, L if k.endswith("d") else Lm), complex, order="F")
         for k in ("Gnd", "Gnu", "Gnl", "Gpd", "Gpu", "Gpl")}
    o["RhoN"] = np.zeros((L, wire.npy, wire.npz), order="F")
    o["RhoP"] = np.zeros((L, wire.npy, wire.npz), order="F")
    args = [_p(_c(GRLef)), _p(_c(GRRgt)), _p(_c(GammaL)), _p(_c(GammaR))]
    if use_ref:
        ref().ref_gnp(n, L, wire.npy, wire.npz, wire.norb, f1, f2, deltaE, *args,
                      _p(o["Gnd"]), _p(o["Gnu"]), _p(o["Gnl"]), _p(o["Gpd"]), _p(o["Gpu"]),
                      _p(o["Gpl"]), _p(o["RhoN"]), _p(o["RhoP"]))
    else:
        lib().oracle_gnp(n, L, wire.npy, wire.npz, wire.norb, f1, f2, deltaE, *args,
                         _p(o["Gnd"]), _p(o["Gnu"]), _p(o["Gnl"]), _p(o["Gpd"]), _p(o["Gpu"]),
                         _p(o["Gpl"]), _p(o["RhoN"]), _p(o["RhoP"]))
    return o


def observables(wire, f1, f2, deltaE, g, GRRgt, GammaL, GammaR):
    """g: dict with Gnd..Gpl (from gnp)."""
    L = wire.npx
    sp = (L, wire.npy, wire.npz)
    o = {k: np.zeros(sp, order="F") for k in ("SigX", "SigY", "SigZ", "Jx", "Jy", "Jz")}
    perE = np.zeros(7)
    cc = np.zeros(2)
    gm = wire.gm_fortran()
    lib().oracle_observables(wire.n, L, wire.npy, wire.npz, wire.norb, f1, f2, deltaE,
                             _p(wire.htotd), _p(wire.hopx), _p(g["Gnd"]), _p(g["Gpd"]),
                             _p(g["Gnu"]), _p(g["Gnl"]), _p(g["Gpu"]), _p(g["Gpl"]),
                             _p(_c(GRRgt)), _p(_c(GammaL)), _p(_c(GammaR)), _p(gm), _p(perE),
                             _p(o["SigX"]), _p(o["SigY"]), _p(o["SigZ"]), _p(o["Jx"]),
                             _p(o["Jy"]), _p(o["Jz"]), _p(cc))
    o.update(T12=perE[0], Modec1=perE[1], Modec2=perE[2], EDensity=perE[3], HDensity=perE[4],
             ContactCurrent=cc)
    return o


PER_ENERGY = ("T12", "Modec1", "Modec2", "EDensity", "HDensity", "DOSc1", "DOSc2")
SPATIAL = ("RhoN", "RhoP", "SigX", "SigY", "SigZ", "Jx", "Jy", "Jz")


def energy_sweep(wire, energy, deltaE, mu=(-0.1, 0.0), muT=0.0, eta=1e-6, sr_tol=1e-10,
                 sr_max_iter=200, sr_extra_eta=True):
    """One Energy_Sweep call; returns per-energy values and this energy's
    contributions to the energy-integrated spatial arrays."""
    L = wire.npx
    sp = (L, wire.npy, wire.npz)
    acc = {k: np.zeros(sp, order="F") for k in SPATIAL}
    perE = np.zeros(7)
    cc = np.zeros(2)
    iters = (ctypes.c_int * 2)()
    gm = wire.gm_fortran()
    st = lib().oracle_energy_sweep(wire.npx, wire.npy, wire.npz, wire.norb, _p(wire.htotd),
                                   _p(wire.hopx), _p(gm), energy, deltaE, mu[0], mu[1], muT, eta,
                                   sr_tol, sr_max_iter, int(bool(sr_extra_eta)), _p(perE),
                                   *[_p(acc[k]) for k in SPATIAL], _p(cc), iters)
    out = dict(zip(PER_ENERGY, perE))
    out.update(acc)
    out.update(ContactCurrent=cc, status=st, iters=(iters[0], iters[1]))
    return out


def resolved(wire, f1, f2, deltaE, g, eCtr=0, estpnum=1):
    """Energy- and orbital-resolved outputs of one energy from its Gn/Gp blocks (g:
    dict with Gnd..Gpl, as gnp returns).  Restates, in plain loops (small wires only):
      * Calculate_Gnp.F90:118-130 -- ERRhon(x,y,z,eCtr) += sum_ii Re Gnd(i,i) - Re Gpd(i,i),
        ORRhon(x,y,z,ii) += Re Gnd(i,i) dE/2pi;
      * Calculate_Observables.F90:111-168 -- per bond and orbital pair (ii, jj)
        dumVar = Re i q^2/hbar (H_ab Gn_ba - Gn_ab H_ba - H_ab Gp_ba + Gp_ab H_ba)
        (x bonds: HopX with Gnl/Gnu/Gpl/Gpu), ERJ(x,y,z,eCtr) += dumVar,
        ORJ(x,y,z,(ii-1)*Norb+jj) += dumVar dE/2pi, kept in REAL(8) (the reference
        rounds dumVar to REAL(4), Calculate_Observables.F90:7).
    Returns a dict of those arrays for this energy alone."""
    L, npy, npz, no = wire.npx, wire.npy, wire.npz, wire.norb
    n = wire.n
    NN = npy * no
    H, X = wire.htotd, wire.hopx
    Gnd, Gpd, Gnu, Gnl, Gpu, Gpl = (g[k] for k in ("Gnd", "Gpd", "Gnu", "Gnl", "Gpu", "Gpl"))
    pref = 1j * 1.602e-19 ** 2 / (6.626e-34 / (2 * np.pi))
    w = deltaE / (2 * np.pi)
    out = {k: np.zeros((L, npy, npz, estpnum)) for k in ("ERJx", "ERJy", "ERJz", "ERRhon")}
    out.update({k: np.zeros((L, npy, npz, no * no)) for k in ("ORJx", "ORJy", "ORJz")})
    out["ORRhon"] = np.zeros((L, npy, npz, no))
    for x in range(L):
        for y in range(npy):
            for z in range(npz):
                idx = y * no + z * NN
                for ii in range(no):
                    i = idx + ii
                    out["ERRhon"][x, y, z, eCtr] += Gnd[i, i, x].real - Gpd[i, i, x].real
                    out["ORRhon"][x, y, z, ii] += Gnd[i, i, x].real * w
                    for jj in range(no):
                        p = ii * no + jj
                        terms = []
                        if y < npy - 1:
                            a, b = idx + ii, idx + no + jj
                            terms.append(("y", H[a, b, x] * Gnd[b, a, x] - Gnd[a, b, x] * H[b, a, x]
                                          - H[a, b, x] * Gpd[b, a, x] + Gpd[a, b, x] * H[b, a, x]))
                        if z < npz - 1:
                            a, b = idx + ii, idx + NN + jj
                            terms.append(("z", H[a, b, x] * Gnd[b, a, x] - Gnd[a, b, x] * H[b, a, x]
                                          - H[a, b, x] * Gpd[b, a, x] + Gpd[a, b, x] * H[b, a, x]))
                        if x < L - 1:
                            a, b = idx + ii, idx + jj
                            hx = X[a, b]
                            terms.append(("x", hx * Gnl[b, a, x] - Gnu[a, b, x] * np.conj(hx)
                                          - hx * Gpl[b, a, x] + Gpu[a, b, x] * np.conj(hx)))
                        for d, t in terms:
                            v = (pref * t).real
                            out["ERJ" + d][x, y, z, eCtr] += v
                            out["ORJ" + d][x, y, z, p] += v * w
    return out


def energy_resolved(wire, energy, deltaE, mu=(-0.1, 0.0), muT=0.0, eta=1e-6, sr_tol=1e-10,
                    sr_max_iter=200, eCtr=0, estpnum=1):
    """One Energy_Sweep's resolved outputs: contacts, InvGRd, RGF, Gn/Gp (the oracle's
    stages), then `resolved`."""
    f1 = 1.0 if energy <= muT - mu[0] else 0.0  # Energy_Sweep.F90:11-23
    f2 = 1.0 if energy <= muT - mu[1] else 0.0
    c = contact_self_energy(wire, energy, eta, sr_tol, sr_max_iter, True)
    A = assemble_invg(wire, c["SigmaL"], c["SigmaR"], energy, eta)
    r = retarded(A, wire.hopx)
    g = gnp(wire, f1, f2, deltaE, r["GRLef"], r["GRRgt"], c["GammaL"], c["GammaR"])
    return resolved(wire, f1, f2, deltaE, g, eCtr, estpnum)
