"""Output files of a sweep in the reference's formats.

``dump_data`` writes the text files of Dump_Data (Dump_Data.F90:6-147) -- the
files the reference's Analyze/NEGF_ReadData.m reads -- from the accumulators of
a solve, plus ``results.npz`` with the same arrays at full precision (the text
files keep ~5 significant digits: ES11.4, Dump_Data.F90:152).  ``read_data``
reads the text files back in NEGF_ReadData.m's order.

Files (one value per line, in the loop orders of Dump_Data):
  scalars.txt        Npx Npy Npz Norb muT eMin eMax eStpnum deltaE mu(1:2)
                     channelType [a0 if channelType == 1] abPhi corrLength
                     disorderStrength                              (:6-25)
  Greens.txt         RhoN, RhoP per site (x, then y, then z), then DeltaR per
                     site (zero: no superconducting pairing here)     (:27-45)
  Transmissions.txt  EDensity, HDensity per energy; T12; Modec1, Modec2 per
                     energy; EAxis; ContactCurrent(1:2); Jx, Jy, Jz per site;
                     DOSc1, DOSc2 per energy                          (:47-83)
  SigExp.txt         SigXExp, SigYExp, SigZExp per site               (:127-139)
  Poissons.txt       Phi, ro per site, when a potential is given     (:84-97)
  ERCD.txt           per site, per energy: ERJx ERJy ERJz ERRhon, when the
                     accumulators hold the energy-resolved arrays     (:100-113)
  ORCD.txt           per site: ORJx ORJy ORJz per orbital pair, then ORRhon
                     per orbital, when they hold the orbital-resolved
                     arrays                                           (:116-132)
"""
from __future__ import annotations

import os

import numpy as np


def _es(x: float) -> str:
    """Fortran ES11.4 (Dump_Data.F90:152): a 5-digit mantissa and a signed 2-digit
    exponent 'E+dd'; a 3-digit exponent replaces the 'E' (Fortran's Ew.d rule for
    exponents beyond 99); right-justified to 11 characters."""
    x = float(x)
    if not np.isfinite(x):  # as the LLVM Fortran runtime writes them
        return ("NaN" if x != x else ("Inf" if x > 0 else "-Inf")).rjust(11)
    mant, exp = f"{x:.4E}".split("E")
    e = int(exp)
    field = f"{mant}E{e:+03d}" if abs(e) <= 99 else f"{mant}{e:+04d}"
    return field.rjust(11)


def _ld(x) -> str:
    """List-directed output of a real or integer (WRITE(13, *))."""
    if isinstance(x, (int, np.integer)):
        return f"{int(x):12d}"
    return f"{float(x): .15E}"


def _sites(npx, npy, npz):
    for x in range(npx):
        for y in range(npy):
            for z in range(npz):
                yield x, y, z


def dump_data(outdir, dims, energies, acc, *, mu=(-0.1, 0.0), muT=0.0, channel_type=0, a0=1.0,
              ab_phi=0.0, corr_length=0.0, disorder=0.0, currents=None, phi=None, ro=None,
              full_precision=True):
    """Write the Dump_Data files for one sweep.

    dims = (Npx, Npy, Npz, Norb); energies = EAxis; acc: an Accumulators (T12 ..
    ContactCurrent, and Jx/Jy/Jz and the resolved arrays where it holds them);
    currents: (Jx, Jy, Jz) to use instead of the accumulators' (zero when neither
    has them); phi, ro: the Poisson potential and charge (Npx, Npy, Npz), written
    to Poissons.txt when given."""
    npx, npy, npz, norb = dims
    e = np.asarray(energies, float)
    ne = len(e)
    de = e[1] - e[0] if ne > 1 else 0.0
    os.makedirs(outdir, exist_ok=True)
    if currents is None and getattr(acc, "Jx", None) is not None:
        currents = (acc.Jx, acc.Jy, acc.Jz)
    J = currents if currents is not None else [np.zeros((npx, npy, npz)) for _ in range(3)]
    with open(os.path.join(outdir, "scalars.txt"), "w") as f:
        lines = [_ld(npx), _ld(npy), _ld(npz), _ld(norb), _es(muT), _es(e.min() if ne else 0.0),
                 _es(e.max() if ne else 0.0), _ld(ne), _es(de), _es(mu[0]), _es(mu[1]),
                 _ld(channel_type)]
        if channel_type == 1:
            lines.append(_ld(float(a0)))
        lines += [_ld(float(ab_phi)), _es(corr_length), _es(disorder)]
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(outdir, "Greens.txt"), "w") as f:
        out = []
        for x, y, z in _sites(npx, npy, npz):
            out += [_es(acc.RhoN[x, y, z]), _es(acc.RhoP[x, y, z])]
        out += [_es(0.0) for _ in _sites(npx, npy, npz)]
        f.write("\n".join(out) + "\n")
    with open(os.path.join(outdir, "Transmissions.txt"), "w") as f:
        out = []
        for i in range(ne):
            out += [_es(acc.EDensity[i]), _es(acc.HDensity[i])]
        out += [_es(acc.T12[i]) for i in range(ne)]
        for i in range(ne):
            out += [_es(acc.Modec1[i]), _es(acc.Modec2[i])]
        out += [_ld(float(x)) for x in e]
        out += [_ld(float(acc.ContactCurrent[k])) for k in range(2)]
        for x, y, z in _sites(npx, npy, npz):
            out += [_ld(float(J[0][x, y, z])), _ld(float(J[1][x, y, z])), _ld(float(J[2][x, y, z]))]
        for i in range(ne):
            out += [_es(acc.DOSc1[i]), _es(acc.DOSc2[i])]
        f.write("\n".join(out) + "\n")
    with open(os.path.join(outdir, "SigExp.txt"), "w") as f:
        out = []
        for x, y, z in _sites(npx, npy, npz):
            out += [_es(acc.SigXExp[x, y, z]), _es(acc.SigYExp[x, y, z]), _es(acc.SigZExp[x, y, z])]
        f.write("\n".join(out) + "\n")
    if phi is not None:
        with open(os.path.join(outdir, "Poissons.txt"), "w") as f:
            out = []
            for x, y, z in _sites(npx, npy, npz):
                out += [_es(phi[x, y, z]), _es(ro[x, y, z] if ro is not None else 0.0)]
            f.write("\n".join(out) + "\n")
    er = getattr(acc, "ERJx", None) is not None
    orb = getattr(acc, "ORJx", None) is not None
    if er:
        with open(os.path.join(outdir, "ERCD.txt"), "w") as f:
            out = []
            for x, y, z in _sites(npx, npy, npz):
                for i in range(ne):
                    out += [_es(acc.ERJx[x, y, z, i]), _es(acc.ERJy[x, y, z, i]),
                            _es(acc.ERJz[x, y, z, i]), _es(acc.ERRhon[x, y, z, i])]
            f.write("\n".join(out) + "\n")
    if orb:
        with open(os.path.join(outdir, "ORCD.txt"), "w") as f:
            out = []
            for x, y, z in _sites(npx, npy, npz):
                for p in range(norb * norb):
                    out += [_es(acc.ORJx[x, y, z, p]), _es(acc.ORJy[x, y, z, p]), _es(acc.ORJz[x, y, z, p])]
                out += [_es(acc.ORRhon[x, y, z, r]) for r in range(norb)]
            f.write("\n".join(out) + "\n")
    if full_precision:
        names = ["T12", "Modec1", "Modec2", "DOSc1", "DOSc2", "EDensity", "HDensity", "RhoN", "RhoP",
                 "SigXExp", "SigYExp", "SigZExp", "ContactCurrent"]
        if er:
            names += ["ERJx", "ERJy", "ERJz", "ERRhon"]
        if orb:
            names += ["ORJx", "ORJy", "ORJz", "ORRhon"]
        arrays = {k: np.asarray(getattr(acc, k)) for k in names}
        np.savez(os.path.join(outdir, "results.npz"), EAxis=e, Jx=J[0], Jy=J[1], Jz=J[2],
                 dims=np.array(dims), mu=np.array(mu), muT=np.array(muT), **arrays)


def read_data(outdir):
    """Read the text files back in NEGF_ReadData.m's cursor order (Analyze/NEGF_ReadData.m)."""
    s = np.loadtxt(os.path.join(outdir, "scalars.txt"))
    c = 0
    npx, npy, npz, norb = (int(s[c + k]) for k in range(4))
    c += 4
    muT, emin, emax = s[c], s[c + 1], s[c + 2]
    c += 3
    ne = int(s[c]); c += 1
    de = s[c]; c += 1
    mu = (s[c], s[c + 1]); c += 2
    ctype = int(s[c]); c += 1
    if ctype == 1:
        c += 1
    out = dict(dims=(npx, npy, npz, norb), muT=muT, eMin=emin, eMax=emax, eStpnum=ne, deltaE=de, mu=mu)
    g = np.loadtxt(os.path.join(outdir, "Greens.txt"))
    sites = list(_sites(npx, npy, npz))
    rho = {k: np.zeros((npx, npy, npz)) for k in ("RhoN", "RhoP", "DeltaR")}
    c = 0
    for x, y, z in sites:
        rho["RhoN"][x, y, z], rho["RhoP"][x, y, z] = g[c], g[c + 1]
        c += 2
    for x, y, z in sites:
        rho["DeltaR"][x, y, z] = g[c]
        c += 1
    out.update(rho)
    t = np.loadtxt(os.path.join(outdir, "Transmissions.txt"))
    c = 0
    out["EDensity"], out["HDensity"] = t[c:c + 2 * ne:2], t[c + 1:c + 2 * ne:2]
    c += 2 * ne
    out["T12"] = t[c:c + ne]; c += ne
    out["Modec1"], out["Modec2"] = t[c:c + 2 * ne:2], t[c + 1:c + 2 * ne:2]
    c += 2 * ne
    out["EAxis"] = t[c:c + ne]; c += ne
    out["ContactCurrent"] = t[c:c + 2]; c += 2
    J = {k: np.zeros((npx, npy, npz)) for k in ("Jx", "Jy", "Jz")}
    for x, y, z in sites:
        J["Jx"][x, y, z], J["Jy"][x, y, z], J["Jz"][x, y, z] = t[c], t[c + 1], t[c + 2]
        c += 3
    out.update(J)
    out["DOSc1"], out["DOSc2"] = t[c:c + 2 * ne:2], t[c + 1:c + 2 * ne:2]
    sg = np.loadtxt(os.path.join(outdir, "SigExp.txt"))
    S = {k: np.zeros((npx, npy, npz)) for k in ("SigXExp", "SigYExp", "SigZExp")}
    c = 0
    for x, y, z in sites:
        S["SigXExp"][x, y, z], S["SigYExp"][x, y, z], S["SigZExp"][x, y, z] = sg[c], sg[c + 1], sg[c + 2]
        c += 3
    out.update(S)
    p = os.path.join(outdir, "Poissons.txt")
    if os.path.exists(p):
        v = np.loadtxt(p)
        out["phi"], out["ro"] = np.zeros((npx, npy, npz)), np.zeros((npx, npy, npz))
        for k, (x, y, z) in enumerate(sites):
            out["phi"][x, y, z], out["ro"][x, y, z] = v[2 * k], v[2 * k + 1]
    p = os.path.join(outdir, "ERCD.txt")
    if os.path.exists(p):
        v = np.loadtxt(p).reshape(npx, npy, npz, ne, 4)
        for i, k in enumerate(("ERJx", "ERJy", "ERJz", "ERRhon")):
            out[k] = v[..., i]
    p = os.path.join(outdir, "ORCD.txt")
    if os.path.exists(p):
        v = np.loadtxt(p).reshape(npx, npy, npz, 3 * norb * norb + norb)
        j = v[..., :3 * norb * norb].reshape(npx, npy, npz, norb * norb, 3)
        for i, k in enumerate(("ORJx", "ORJy", "ORJz")):
            out[k] = j[..., i]
        out["ORRhon"] = v[..., 3 * norb * norb:]
    return out
