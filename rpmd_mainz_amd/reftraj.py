"""Restart and trajectory files in the reference's formats, and the force-only drivers that replay
stored frames through the GPU force path.

Formats (the reference writes them with Fortran list-directed output; any whitespace layout and
E or D exponents read back):

* restart (``print_structure``, src/io.f90:1-17; read back at src/main.f90:310-357):
  ``nm na nb`` / ``Lx Ly Lz`` (bohr) / ``r(3,na,nb)`` in column order (x, y, z of atom 1 of bead 1,
  then atom 2, ...).  A file with another bead count than the run's starts every bead at the file's
  first bead.
* VMD frame (``print_vmd_bead``, ``print_vmd_bead_forces``, ``print_vmd_bead_vels``,
  src/io.f90:549-770; ``print_vmd_full_forces``, force_matching/src/io.f90:409-477): ``na`` /
  ``BOX Lx Ly Lz`` (Angstrom) / one ``El x y z`` line per atom, O H H per molecule (D for
  deuterium).  Positions in Angstrom, forces -(dvdr + dvdr2) in Hartree/bohr, velocities p/m in
  atomic units.
* reference trajectory (reftraj > 0, src/main.f90:360-400): VMD position frames in Angstrom, all
  frames of bead 1, then all frames of bead 2, ...; each frame's box is its BOX line.

Drivers:

* ``replay_snapshots`` -- the reftraj > 0 branch of the main program (src/main.f90:891-898): every
  stored frame through the forces of one RPMD step (``full_forces``: ring-polymer contraction,
  Ewald and O-O on the contracted modes, intramolecular and TEM terms), per-bead forces and
  potential.
* ``force_match`` -- the force-matching program (force_matching/src/main.f90:336-366): per frame the
  forces of the whole potential (iopt 0) and of its four parts -- Coulomb (Ewald with the M-site
  chain rule), O-O (LJ or Buckingham), intramolecular stretch and bend
  (force_matching/src/md_forces.f90:1-7, intra_morse.f90:126-146) -- written as centroid forces to
  vmd.frc, vmd.frc1 .. vmd.frc4 and vmd.frcT (the sum of the four parts).
"""
import os

import numpy as np

from .configs import TOA

TOKJMOL = 2625.49963  # Hartree -> kJ/mol (src/globals.inc:18)
_LABELS = {"H2O": ("O", "H", "H"), "D2O": ("O", "D", "D"), "HDO": ("O", "D", "H")}


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def _fmt(x: float) -> str:
    return f"{x: .16E}"


def _nint(x):
    """Fortran nint: to the nearest integer, halves away from zero"""
    return np.trunc(x + np.copysign(0.5, x))


# ------------------------------------------------------------------------------------------------
# restart files
# ------------------------------------------------------------------------------------------------
def read_restart(path: str, nb: int = None) -> dict:
    """Read a restart file (src/main.f90:317-357).  nb: the run's bead count (default: the file's);
    when it differs from the file's, the file's first bead is read and copied to every bead.
    Returns dict(nm, na, nb_file, box (bohr), r (nb, na, 3) bohr)."""
    toks = open(path).read().split()
    if len(toks) < 6:
        raise ValueError(f"{path}: not a restart file (header and box missing)")
    nm, na, nbr = int(toks[0]), int(toks[1]), int(toks[2])
    box = np.array([_num(t) for t in toks[3:6]])
    nb = nbr if nb is None else int(nb)
    nread = nbr if nb == nbr else 1
    need = 3 * na * nread
    if len(toks) - 6 < need:
        raise ValueError(f"{path}: {len(toks) - 6} coordinates, {need} expected for {na} atoms x {nread} beads")
    r = np.array([_num(t) for t in toks[6:6 + need]]).reshape(nread, na, 3)
    if nread != nb:
        r = np.repeat(r, nb, axis=0)
    return dict(nm=nm, na=na, nb_file=nbr, box=box, r=np.ascontiguousarray(r))


def write_restart(path: str, r, box):
    """Write a restart file (print_structure, src/io.f90:1-17): r (nb, na, 3) and the box in bohr,
    17 significant digits, one atom per line."""
    r = np.asarray(r, np.float64)
    nb, na = r.shape[:2]
    with open(path, "w") as f:
        f.write(f"{na // 3:12d}{na:12d}{nb:12d}\n")
        f.write(" ".join(_fmt(x) for x in np.asarray(box, np.float64)) + "\n")
        for xyz in r.reshape(-1, 3):
            f.write(" ".join(_fmt(x) for x in xyz) + "\n")


# ------------------------------------------------------------------------------------------------
# VMD frames
# ------------------------------------------------------------------------------------------------
def _read_frame(lines, i: int, path: str):
    n = int(lines[i].split()[0])
    head = lines[i + 1].split()
    if len(head) < 4:
        raise ValueError(f"{path}: line {i + 2}: BOX line expected")
    box = np.array([_num(t) for t in head[1:4]])
    if i + 2 + n > len(lines):
        raise ValueError(f"{path}: frame at line {i + 1} holds fewer than {n} atom lines")
    sym, xyz = [], np.empty((n, 3))
    for j in range(n):
        t = lines[i + 2 + j].split()
        sym.append(t[0])
        xyz[j] = [_num(t[1]), _num(t[2]), _num(t[3])]
    return n, box, sym, xyz


def read_xyz_frames(path: str) -> list:
    """Every frame of a VMD file: a list of dict(box (3,), symbols, xyz (na, 3)), in the file's
    units (Angstrom for positions and boxes, a.u. for forces and velocities)."""
    lines = [ln for ln in open(path).read().split("\n") if ln.strip()]
    out, i = [], 0
    while i < len(lines):
        n, box, sym, xyz = _read_frame(lines, i, path)
        out.append(dict(box=box, symbols=sym, xyz=xyz))
        i += 2 + n
    return out


def read_reftraj(path: str, nb: int, nframes: int, na: int = None):
    """A reference trajectory (src/main.f90:360-400): for every bead k, nframes VMD position frames
    (Angstrom).  na: the run's atom count; a frame with more atoms is an error, one with fewer sets
    na (as the reference does).  Returns (r (nframes, nb, na, 3) bohr, boxes (nframes, 3) bohr); a
    frame's box is its BOX line (the last bead's, which the reference keeps)."""
    lines = [ln for ln in open(path).read().split("\n") if ln.strip()]
    frames, boxes, i = {}, np.zeros((nframes, 3)), 0
    for k in range(nb):
        for f in range(nframes):
            if i >= len(lines):
                raise ValueError(f"{path}: {nb} beads x {nframes} frames expected, the file ends at bead {k + 1} "
                                 f"frame {f + 1}")
            n, box, _, xyz = _read_frame(lines, i, path)
            if na is not None and n > na:
                raise ValueError(f"{path}: frame with {n} atoms, more than the run's {na}")
            na = n if na is None or n < na else na
            frames[(f, k)] = xyz[:na] * (1.0 / TOA)
            boxes[f] = box * (1.0 / TOA)
            i += 2 + n
    r = np.empty((nframes, nb, na, 3))
    for (f, k), xyz in frames.items():
        r[f, k] = xyz[:na]
    return r, boxes


def _write_frame(f, rows, box_bohr, isotope: str):
    lab = _LABELS[isotope]
    na = rows.shape[0]
    f.write(f"{na:12d}\n")
    f.write(" BOX " + " ".join(_fmt(TOA * x) for x in box_bohr) + "\n")
    for a in range(na):
        f.write(f" {lab[a % 3]} " + " ".join(_fmt(x) for x in rows[a]) + "\n")


def write_vmd_bead(f, r, ib: int, box, isotope: str = "H2O", wrap: bool = True):
    """One bead's positions (print_vmd_bead, src/io.f90:549-634), Angstrom.  wrap (the reference
    does it outside reftraj mode): each oxygen is shifted into the box as
    x - L (nint(x / L) - 1/2), its hydrogens with it."""
    rb = np.array(r[ib], np.float64)
    L = np.asarray(box, np.float64)
    if wrap:
        o = rb[0::3].copy()
        d1, d2 = rb[1::3] - o, rb[2::3] - o
        o = o - L * (_nint(o * (1.0 / L)) - 0.5)
        rb[0::3], rb[1::3], rb[2::3] = o, d1 + o, d2 + o
    _write_frame(f, TOA * rb, L, isotope)


def write_vmd_bead_forces(f, dvdr, dvdr2, ib: int, box, isotope: str = "H2O"):
    """One bead's forces -(dvdr + dvdr2) in Hartree/bohr (print_vmd_bead_forces, src/io.f90:636-700)."""
    _write_frame(f, -(np.asarray(dvdr)[ib] + np.asarray(dvdr2)[ib]), np.asarray(box, np.float64), isotope)


def write_vmd_bead_vels(f, p, mass, ib: int, box, isotope: str = "H2O"):
    """One bead's velocities p/m in a.u. (print_vmd_bead_vels, src/io.f90:702-770)."""
    _write_frame(f, np.asarray(p)[ib] / np.asarray(mass)[:, None], np.asarray(box, np.float64), isotope)


def write_vmd_full_forces(f, dvdr, dvdr2, box, isotope: str = "H2O"):
    """Centroid forces -sum_k (dvdr + dvdr2) / nb in Hartree/bohr (print_vmd_full_forces,
    force_matching/src/io.f90:409-477)."""
    d = np.asarray(dvdr) + np.asarray(dvdr2)
    _write_frame(f, -d.sum(axis=0) / d.shape[0], np.asarray(box, np.float64), isotope)


# ------------------------------------------------------------------------------------------------
# force-only drivers
# ------------------------------------------------------------------------------------------------
def unwrap_beads(r, box):
    """Every bead of every atom moved by whole box vectors to the minimum image of its bead 1, so
    that the ring polymer is whole: position dumps written outside reftraj mode wrap each bead's
    oxygens separately (src/io.f90:590-611), a reference trajectory is read as written."""
    r = np.array(r, np.float64)
    L = np.asarray(box, np.float64)
    d = r[1:] - r[:1]
    r[1:] = r[:1] + (d - L * np.rint(d / L))
    return r


def replay_snapshots(ctx, r_traj, boxes, z, out_dir: str = None, isotope: str = "H2O",
                     unwrap: bool = False) -> dict:
    """Every frame of a reference trajectory through the forces of one RPMD step (the reftraj > 0
    branch, src/main.f90:891-898): full_forces on the GPU context ctx (its parameters: contraction,
    multiple time steps, TEM).  unwrap: make each frame's ring polymers whole first (unwrap_beads),
    for frames taken from position dumps written outside reftraj mode.  Returns dict(v (nframes,)
    Hartree, v_kjmol_per_molecule, dvdr, dvdr2 (nframes, nb, na, 3)).  out_dir: also write each
    bead's force dump vmd_bead-<k>.frc (print_vmd_bead_forces) with one frame per snapshot."""
    r_traj = np.asarray(r_traj, np.float64)
    if unwrap:
        r_traj = np.stack([unwrap_beads(r_traj[i], boxes[i]) for i in range(r_traj.shape[0])])
    nf, nb, na = r_traj.shape[:3]
    v = np.empty(nf)
    d1, d2 = np.empty_like(r_traj), np.empty_like(r_traj)
    files = None
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        files = [open(os.path.join(out_dir, f"vmd_bead-{k + 1}.frc"), "w") for k in range(nb)]
    try:
        for i in range(nf):
            o = ctx.full_forces(np.ascontiguousarray(r_traj[i]), na, nb, z, boxes[i])
            v[i], d1[i], d2[i] = o["v"], o["dvdr"], o["dvdr2"]
            if files:
                for k in range(nb):
                    write_vmd_bead_forces(files[k], d1[i], d2[i], k, boxes[i], isotope)
    finally:
        for fh in files or []:
            fh.close()
    return dict(v=v, v_kjmol_per_molecule=v * TOKJMOL / (na // 3), dvdr=d1, dvdr2=d2)


FORCE_PARTS = ("coulomb", "oo", "stretch", "bend")


def force_match(ctx, prm: dict, r_traj, boxes, z, out_dir: str = None, isotope: str = "H2O") -> dict:
    """The force-matching driver (force_matching/src/main.f90:336-366) on the GPU context ctx: per
    frame the gradient of the whole potential (forces, iopt 0: Ewald, O-O and intramolecular terms)
    and of its parts: Coulomb (iopt 2), O-O (iopt 3), stretch and bend (iopt 4 with the bend
    constant bpot, then the stretch constant apot, set to zero: the two parts of
    force_matching/src/intra_morse.f90:126-146 are the apot and the bpot terms of the gradient).
    Every bead is evaluated (the reference's force-matching forces evaluates bead 1, which is the
    whole ring at nb = 1, its use).  The context's parameters are restored afterwards.  Returns
    dict(total, parts (4 x nframes x nb x na x 3), v) of gradients dvdr; out_dir: also write the
    centroid forces to vmd.frc, vmd.frc1 .. vmd.frc4 and vmd.frcT (print_vmd_full_forces)."""
    if prm.get("alp", 0.0) == 0.0:
        raise ValueError("force matching needs the anharmonic intramolecular potential (alp != 0), as the "
                         "reference's force-matching program does")
    r_traj = np.asarray(r_traj, np.float64)
    nf, nb, na = r_traj.shape[:3]
    total = np.empty_like(r_traj)
    parts = np.empty((4,) + r_traj.shape)
    v = np.empty(nf)
    variants = [(2, prm), (3, prm), (4, dict(prm, bpot=0.0)), (4, dict(prm, apot=0.0))]
    try:
        for i in range(nf):
            r = np.ascontiguousarray(r_traj[i])
            ctx.set_params(prm)
            v[i], total[i], _ = ctx.forces(r, nb, na, boxes[i], z, 0)
            for j, (iopt, pj) in enumerate(variants):
                ctx.set_params(pj)
                _, parts[j, i], _ = ctx.forces(r, nb, na, boxes[i], z, iopt)
    finally:
        ctx.set_params(prm)
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        zero = np.zeros((nb, na, 3))
        names = ["vmd.frc"] + [f"vmd.frc{j + 1}" for j in range(4)] + ["vmd.frcT"]
        fh = [open(os.path.join(out_dir, n), "w") for n in names]
        try:
            for i in range(nf):
                write_vmd_full_forces(fh[0], total[i], zero, boxes[i], isotope)
                for j in range(4):
                    write_vmd_full_forces(fh[1 + j], parts[j, i], zero, boxes[i], isotope)
                write_vmd_full_forces(fh[5], parts[:, i].sum(axis=0), zero, boxes[i], isotope)
        finally:
            for f in fh:
                f.close()
    return dict(total=total, parts=parts, v=v)
