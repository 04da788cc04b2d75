"""Water models, seeded lattices and parameter sets for the RPMD force-and-propagate path.

Units are the reference's atomic units (src/globals.inc:10-21).  Model parameter sets restate
the reference's namelist files (params/param_qspcfw, params/param_scott,
params/param_tip05_rigid); the lattice follows src/setup_positions.f90:94-150 (O on a simple
cubic lattice at i*L/n) with two additions the reference lacks: each molecule is rotated about
its oxygen by a seeded uniform random rotation, and body- and face-centred cubic lattices place
2 m^3 and 4 m^3 molecules in a cubic box.
"""
from __future__ import annotations

import math

import numpy as np

TOA = 0.5291772108          # bohr -> Angstrom (src/globals.inc:283)
TOFS = 41.341373            # a.u. time per fs (src/globals.inc:292)
KBOLTZ = 3.166829e-6        # Hartree/K (src/globals.inc:281, src/main.f90:168)
EMASS = 9.1093826e-31       # kg (src/globals.inc:285)
KCAL = 627.5094740631       # kcal/mol per Hartree

# ---- intramolecular model switch (mirrors RPMD_INTRA_* in include/rpmd_mi355x.h) -----------
INTRA_AUTO, INTRA_HARMONIC, INTRA_SPCFW, INTRA_MORSE_HARM, INTRA_MORSE = 0, 1, 2, 3, 4
THERM = {"NONE": 0, "AND": 0, "PRO": 1, "PRM": 2, "PRG": 3, "PRL": 4, "GLE": 5}
PREC_FP64, PREC_MIXED = 0, 1

# ---- model parameter sets (&param namelists) ---------------------------------------------
MODELS = {
    # params/param_qspcfw: q-SPC/Fw, qO=-0.84, LJ sigma 3.1655 A, eps 0.1554 kcal/mol,
    # r_OH = 1.0 A, kb = 1059.162 kcal/mol/A^2 (apot), theta 112 deg, ktheta = 75.90 kcal/mol/rad^2
    "qspcfw": dict(wmass=32831.2525, omass=29156.9471, hmass=1837.1527, qo=-0.84, alpha=1.0,
                   oo_sig=5.981913, oo_eps=2.47686e-4, oo_gam=0.0, thetad=112.0, reoh=1.889726,
                   apot=0.472655, bpot=0.120954, alp=0.0, alpb=0.0, wm=0.0, wh=0.0),
    # params/param_scott: q-TIP4P/F (the reference's regression-test model)
    "qtip4pf": dict(wmass=32831.2525, omass=29156.9471, hmass=1837.1527, qo=-1.1128, alpha=0.73612,
                    oo_sig=5.96946, oo_eps=2.95147e-4, oo_gam=0.0, thetad=107.4, reoh=1.78,
                    apot=0.185, bpot=0.07, alp=1.21, alpb=0.0, wm=0.0, wh=0.0),
    # params/param_tip05_rigid: TIP4P/2005 (rigid: apot = bpot = 0)
    "tip4p2005": dict(wmass=32831.2525, omass=29156.9471, hmass=1837.1527, qo=-1.1128, alpha=0.73612,
                      oo_sig=5.96946, oo_eps=2.95147e-4, oo_gam=0.0, thetad=104.52, reoh=1.8088465,
                      apot=0.0, bpot=0.0, alp=0.0, alpb=0.0, wm=0.0, wh=0.0),
    # params/param_smear: Gaussian (smeared) M-site and hydrogen charges, widths wm, wh; Morse
    # stretch and bend (alp, alpb > 0)
    "smear": dict(wmass=32831.2525, omass=29156.9471, hmass=1837.1527, qo=-1.084, alpha=0.797,
                  oo_sig=5.761, oo_eps=3.5211e-4, oo_gam=0.0, thetad=106.416, reoh=1.77956,
                  apot=0.17, bpot=0.52, alp=1.22, alpb=0.37, wm=0.482, wh=0.686),
}


def box_size(nm: int, rho: float, wmass: float, ncell=None, aspect=(1.0, 1.0, 1.0)) -> np.ndarray:
    """setup_box_size (src/setup_options.f90:191-243) for a cubic-cell lattice; ``aspect``
    stretches the box at fixed volume (an orthorhombic box the reference generator cannot make)."""
    vbox = (nm * EMASS * wmass) / (rho * 1000.0)
    vbox = vbox / (TOA * 1e-10) ** 3
    if ncell is None:
        n = round(nm ** (1.0 / 3.0))
        ncell = (n, n, n)
    totcell = float(ncell[0] * ncell[1] * ncell[2])
    boxl = (vbox / totcell) ** (1.0 / 3.0)
    box = np.array([boxl * ncell[0], boxl * ncell[1], boxl * ncell[2]])
    a = np.asarray(aspect, dtype=float)
    a = a / np.prod(a) ** (1.0 / 3.0)
    return box * a


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


# oxygen sites per lattice cell, in cell units: simple, body-centred and face-centred cubic
BASES = {
    "sc": np.zeros((1, 3)),
    "bcc": np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]),
    "fcc": np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]]),
}


def cubic_cells(nm: int):
    """(ncell, basis) of a cubic lattice of exactly nm molecules in a cubic box: m^3 simple cubic,
    2 m^3 body-centred or 4 m^3 face-centred, so that neighbouring oxygens are as far apart in
    every direction (128 = 2 * 4^3 and 1024 = 2 * 8^3 molecules on bcc, 256 = 4 * 4^3 on fcc: nearest
    oxygens at 3.4 and 3.5 angstrom at 1 g/cm^3, where a 4x4x8 or 8x8x16 simple lattice in a cube
    would put them 2.0 angstrom apart along its long axis)."""
    for basis in ("sc", "bcc", "fcc"):
        m = round((nm / len(BASES[basis])) ** (1.0 / 3.0))
        if len(BASES[basis]) * m ** 3 == nm:
            return (m, m, m), basis
    raise ValueError(f"{nm} molecules fill no cubic lattice (m^3, 2 m^3 or 4 m^3)")


def water_lattice(ncell, box, reoh: float, theta: float, seed: int, random_orient: bool = True,
                  basis: str = "sc") -> np.ndarray:
    """Positions (na, 3) of nm = len(BASES[basis]) * prod(ncell) molecules, (O,H,H) order, bohr."""
    ncell = tuple(int(c) for c in ncell)
    idx = np.stack(np.meshgrid(*[np.arange(c) for c in ncell], indexing="ij"), -1).reshape(-1, 3)
    # src/setup_positions.f90 loops x fastest; keep that molecule order (the basis sites of a cell
    # consecutive)
    idx = idx[np.lexsort((idx[:, 0], idx[:, 1], idx[:, 2]))]
    cell = np.asarray(box) / np.asarray(ncell)
    o = ((idx[:, None, :] + BASES[basis][None, :, :]) * cell).reshape(-1, 3)
    h1 = np.array([reoh, 0.0, 0.0])
    h2 = np.array([reoh * math.cos(theta), reoh * math.sin(theta), 0.0])
    nm = o.shape[0]
    if random_orient:
        rot = _random_rotations(np.random.default_rng(seed), nm)
        d1, d2 = rot @ h1, rot @ h2
    else:
        d1, d2 = np.tile(h1, (nm, 1)), np.tile(h2, (nm, 1))
    r = np.empty((nm, 3, 3))
    r[:, 0], r[:, 1], r[:, 2] = o, o + d1, o + d2
    return r.reshape(-1, 3)


def masses_charges(model: dict, na: int):
    """src/main.f90:526-586 (H2O isotope): mass O,H,H and charges qO, -qO/2, -qO/2."""
    mass = np.tile([model["omass"], model["hmass"], model["hmass"]], na // 3).astype(float)
    qh = -0.5 * model["qo"]
    z = np.tile([model["qo"], qh, qh], na // 3).astype(float)
    return mass, z


def make_params(model: dict, box, rcut_A: float = 9.0, ecut: float = 0.0, iamcub: bool = True,
                intra_model: int = INTRA_AUTO, sig_A: float = 0.0, nbdf1: int = 0, nbdf2: int = 0,
                mts: int = 1, therm: str = "NONE", ttaufs: float = 0.0, rigid: bool = False,
                oo_gam: float | None = None, rat_tol: float = 0.0) -> dict:
    """The COMMON-block values of src/main.f90 for a model, before setup_ewald."""
    theta = model["thetad"] * (math.pi / 180.0)
    rcut = rcut_A / TOA
    rcut = min(rcut, 0.5 * float(min(box)))      # src/main.f90:615-616
    # oo_gam != 0 selects the Buckingham exp-6 O-O term (src/md_forces.f90:226)
    gam = model["oo_gam"] if oo_gam is None else oo_gam
    return dict(oo_eps=model["oo_eps"], oo_sig=model["oo_sig"], oo_gam=gam, rcut=rcut,
                alpha=model["alpha"], ecut=ecut, wm=model["wm"], wh=model["wh"],
                wrcut=0.0, walpha=0.0, rkmax=0.0, kmax=0, iamcub=1 if iamcub else 0,
                apot=model["apot"], bpot=model["bpot"], alp=model["alp"], alpb=model["alpb"],
                theta=theta, reoh=model["reoh"], sig=sig_A / TOA if sig_A > 0 else 0.0,
                nbdf1=nbdf1, nbdf2=nbdf2, mts=mts, intra_model=intra_model,
                therm=THERM[therm], ttau=TOFS * ttaufs, rigid=1 if rigid else 0, rat_tol=rat_tol)


def setup_ewald(na: int, box, prm: dict) -> dict:
    """setup_ewald (src/setup_options.f90:104-189) through the library's C-ABI (rpmd_setup_ewald),
    the one implementation of the rule in this package."""
    from rpmd_mainz_amd import capi
    return capi.setup_ewald(na, [float(b) for b in box], prm)


def thermal_beads(r0: np.ndarray, nb: int, amp: float, seed: int, rigid: bool = False) -> np.ndarray:
    """(nb, na, 3) bead positions: r0 plus a seeded Gaussian displacement per bead (test inputs).
    rigid: one displacement per molecule and bead, so every bead keeps the molecular geometry."""
    rng = np.random.default_rng(seed)
    if rigid:
        d = amp * rng.normal(size=(nb, r0.shape[0] // 3, 1, 3))
        return r0[None, :, :] + np.repeat(d, 3, axis=2).reshape((nb,) + r0.shape)
    return r0[None, :, :] + amp * rng.normal(size=(nb,) + r0.shape)


def water_system(model_name: str, ncell, nb: int, rho: float = 0.997, seed: int = 1, amp: float = 0.05,
                 aspect=(1.0, 1.0, 1.0), basis: str = "sc", **kw):
    """A seeded bead configuration plus its parameter dict: (r(nb,na,3), box, z, mass, prm)."""
    model = MODELS[model_name]
    nm = len(BASES[basis]) * int(np.prod(ncell))
    box = box_size(nm, rho, model["wmass"], ncell, aspect)
    if np.allclose(box, box[0], rtol=1e-12, atol=0.0):
        box = np.full(3, box[0])          # e.g. an 8x8x16 lattice in a cubic box (aspect 2:2:1)
    theta = model["thetad"] * math.pi / 180.0
    r0 = water_lattice(ncell, box, model["reoh"], theta, seed, basis=basis)
    r = (thermal_beads(r0, nb, amp, seed + 1000, bool(kw.get("rigid", False))) if amp > 0
         else np.repeat(r0[None], nb, 0))
    na = r0.shape[0]
    mass, z = masses_charges(model, na)
    cub = bool(box[0] == box[1] == box[2])
    prm = make_params(model, box, iamcub=cub, **kw)
    prm = setup_ewald(na, box, prm)
    return r, box, z, mass, prm
