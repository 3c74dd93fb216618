"""Seeded synthetic inputs for the five benchmark workloads.

Line distributions (all workloads): intensity log-uniform 1e-26..1e-19
cm/molecule, air half-width U(0.04, 0.10) cm-1/atm, self half-width
U(2, 5) x air, temperature exponent U(0.5, 0.8), lower-state energy
U(0, 3000) cm-1, pressure shift U(-0.005, 0.005) cm-1/atm, isotope 1,
no line coupling.  Seeds: PCG64(20261203 + workload index).
"""
import math

import numpy as np

from .abi import LineTable

SEED0 = 20261203
P0 = 1013.25
T0 = 296.0
K_BOLTZ_CGS = 1.380649e-16
G0 = 9.80665
M_AIR = 28.964e-3 / 6.02214076e23      # kg per molecule
MASSES = [18, 44, 48, 44, 28, 16, 32]  # molecules 1..7: H2O CO2 O3 N2O CO CH4 O2


def synth_lines(n, v1, v2, mols, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    vnu = np.sort(rng.uniform(v1, v2, n))
    sp = 10.0 ** rng.uniform(-26.0, -19.0, n)
    alfa = rng.uniform(0.04, 0.10, n)
    hwhm = alfa * rng.uniform(2.0, 5.0, n)
    tmpalf = rng.uniform(0.5, 0.8, n)
    epp = rng.uniform(0.0, 3000.0, n)
    pshift = rng.uniform(-0.005, 0.005, n)
    mol = np.asarray(mols)[rng.integers(0, len(mols), n)] + 100
    return LineTable.from_lines(vnu, sp, alfa, epp, mol, hwhm, tmpalf, pshift)


def coupled_lines(n, v1, v2, mols, seed, frac1=0.05, frac3=0.02, frac2=0.01):
    """synth_lines with line coupling: a fraction frac1 of the lines with
    first-order coefficients (IFLG 1: Y within +-0.1, G within +-0.3 at
    200, 250, 296, 340 K), frac3 with reduced widths (IFLG 3: GAMMA1,
    GAMMA2 within +-0.05) and frac2 flagged IFLG 2 (no coefficients)."""
    t = synth_lines(n, v1, v2, mols, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 7))
    keep = np.concatenate([np.arange(b * 250, b * 250 + t.nrec[b]) for b in range(t.nblk)])
    u = rng.random(n)
    iflg = np.where(u < frac1, 1, np.where(u < frac1 + frac3, 3, np.where(u < frac1 + frac3 + frac2, 2, 0)))
    y = np.where((iflg == 1)[:, None], rng.uniform(-0.1, 0.1, (n, 4)), rng.uniform(-0.05, 0.05, (n, 4)))
    g = np.where((iflg == 1)[:, None], rng.uniform(-0.3, 0.3, (n, 4)), rng.uniform(-0.05, 0.05, (n, 4)))
    f = {k: getattr(t, k)[keep] for k in ("vnu", "sp", "alfa", "epp", "mol", "hwhm", "tmpalf", "pshift")}
    return LineTable.from_lines_coupled(f["vnu"], f["sp"], f["alfa"], f["epp"], f["mol"], f["hwhm"],
                                        f["tmpalf"], f["pshift"], iflg, y, g)


def config0_lines(n=20000):
    """20,000 H2O-mass lines in 600-1100 cm-1."""
    return synth_lines(n, 600.0, 1100.0, [1], SEED0 + 0)


def config0_layer():
    """One homogeneous layer: 500 hPa, 250 K, 1 km path, absorber column
    1e22 molecules/cm2.  Returns (pave, tave, wk list, wbroad)."""
    p, t, path_cm = 500.0, 250.0, 1.0e5
    n_air = (p * 1.0e3) / (K_BOLTZ_CGS * t) * path_cm  # dyn/cm2 / (erg/K K) * cm
    w = 1.0e22
    return p, t, [w], n_air - w


def config1_lines(n=500000):
    """500,000 lines over 500-2500 cm-1, molecules 1..7 drawn uniformly."""
    return synth_lines(n, 500.0, 2500.0, [1, 2, 3, 4, 5, 6, 7], SEED0 + 1)


# surface VMR and scale height (km; inf = constant) per molecule 1..7
VMR_PROFILE = [(1.0e-2, 2.0), (4.0e-4, math.inf), (3.0e-8, -8.0), (3.2e-7, 60.0),
               (1.5e-7, 20.0), (1.8e-6, 50.0), (0.2095, math.inf)]


def atmosphere(nlev=51, psfc=1013.0, ptop=0.1, tsfc=295.0, seed=SEED0 + 1, nmol=7):
    """Layered atmosphere: levels log-spaced psfc->ptop hPa, 6.5 K/km lapse
    to 217 K then isothermal, with seeded clipped N(0,1.5) K noise (|dT| <=
    3 K) on the layer temperatures.  Returns dict of per-layer pave (hPa),
    tave (K), wk [nlay, nmol] (molecules/cm2), wbroad, z levels (km)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nlay = nlev - 1
    p = psfc * (ptop / psfc) ** (np.arange(nlev) / nlay)
    z = np.zeros(nlev)
    tl = np.zeros(nlev)
    tl[0] = tsfc
    for i in range(1, nlev):
        # hypsometric step with the lower level's temperature, then lapse
        dz = 287.05 * tl[i - 1] / G0 * math.log(p[i - 1] / p[i]) / 1000.0
        z[i] = z[i - 1] + dz
        tl[i] = max(217.0, tsfc - 6.5 * z[i])
    pave = (p[:-1] - p[1:]) / np.log(p[:-1] / p[1:])
    noise = np.clip(rng.normal(0.0, 1.5, nlay), -3.0, 3.0)
    tave = 0.5 * (tl[:-1] + tl[1:]) + noise
    zmid = 0.5 * (z[:-1] + z[1:])
    col = (p[:-1] - p[1:]) * 100.0 / (G0 * M_AIR) / 1.0e4  # molecules/cm2
    wk = np.zeros((nlay, nmol))
    for m in range(nmol):
        v0, h = VMR_PROFILE[m]
        vmr = v0 * np.exp(-zmid / h) if math.isfinite(h) else np.full(nlay, v0)
        if m == 0:
            vmr = np.maximum(vmr, 3.0e-6)
        if m == 2:
            vmr = np.minimum(vmr, 8.0e-6)
        wk[:, m] = vmr * col
    wbroad = col - wk.sum(axis=1)
    return dict(pave=pave, tave=tave, wk=wk, wbroad=wbroad, z=z, plev=p, tlev=tl)


def batch_profiles(nprof=256, seed=SEED0 + 4, v1=500.0, v2=2500.0):
    """Ragged batch of independent profiles: layer counts uniform 20-80,
    surface pressure uniform 600-1050 hPa, band windows random 200-800 cm-1
    sub-ranges of [v1, v2].  Returns a list of dicts (atmosphere() fields
    plus 'v1', 'v2', 'nlay')."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(nprof):
        nlay = int(rng.integers(20, 81))
        psfc = float(rng.uniform(600.0, 1050.0))
        width = float(rng.uniform(200.0, 800.0))
        a = float(rng.uniform(v1, v2 - width))
        atm = atmosphere(nlev=nlay + 1, psfc=psfc, seed=seed * 1000 + k)
        atm.update(v1=round(a, 3), v2=round(a + width, 3), nlay=nlay)
        out.append(atm)
    return out
