"""Write tests/golden/gentest_frames.npz from the reference's own regression fixture.

Run in the container that has the reference mounted (``python tests/golden/make_gentest_fixture.py``).
The output is data only: positions (Angstrom, as printed), forces (Hartree/bohr, = -(dvdr+dvdr2))
and velocities (a.u.) of selected bead frames of test/cmp/gentest.cmp/vmd_bead-{1,2}.{xyz,frc,vel},
the BOX lines, and the starting lattice vmd_start.xyz.  The namelist values of that run
(test/cmp/gentest.cmp/input and param_scott) are written to gentest_params.json.

Frames: the file holds 2 trajectories x 12 frames, one frame every pb = 20 NVE steps
(src/md_dynamics.f90:499-548).  Frames 0,1 and 12,13 are consecutive pairs inside one
trajectory (20 steps apart); 11 is the last frame of trajectory 1.

gentest_traj.npz: positions (Angstrom) and forces of all 24 frames of both beads, with the BOX
lines.  gentest_text/: the reference's own text of the restart file eq.xyz (print_structure,
src/io.f90:1-17) and of the first frame of each bead's position and force dump (print_vmd_bead,
print_vmd_bead_forces, src/io.f90:549-700), for the readers' format tests.
"""
import json
import os
import re

import numpy as np

REF = "/root/reference/test/cmp/gentest.cmp"
OUT = os.path.dirname(os.path.abspath(__file__))
FRAMES = [0, 1, 11, 12, 13]


def read_frames(path):
    frames, boxes = [], []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        n = int(lines[i].split()[0])
        box = [float(x) for x in lines[i + 1].split()[1:4]]
        xyz = np.array([[float(x) for x in ln.split()[1:4]] for ln in lines[i + 2:i + 2 + n]])
        frames.append(xyz)
        boxes.append(box)
        i += 2 + n
    return np.array(frames), np.array(boxes)


def parse_namelist(path):
    vals = {}
    txt = open(path).read()
    for m in re.finditer(r"(\w+)\s*=\s*([^\n]+)", txt):
        k, v = m.group(1), m.group(2).strip().rstrip(",")
        v = v.replace("d0", "").replace("D0", "").replace("d-", "e-").replace("d+", "e+")
        vals[k] = v.strip("'\"")
    return vals


def main():
    out = {}
    for kind in ("xyz", "frc", "vel"):
        per_bead = []
        for b in (1, 2):
            fr, box = read_frames(os.path.join(REF, f"vmd_bead-{b}.{kind}"))
            per_bead.append(fr[FRAMES])
            if kind == "xyz" and b == 1:
                out["box_A"] = box[FRAMES]
        out[kind] = np.stack(per_bead, axis=1)  # (frame, bead, atom, 3)
    start, sbox = read_frames(os.path.join(REF, "vmd_start.xyz"))
    out["start_xyz"] = start[0]
    out["start_box_A"] = sbox[0]
    out["frames"] = np.array(FRAMES)
    np.savez_compressed(os.path.join(OUT, "gentest_frames.npz"), **out)
    params = {"input": parse_namelist(os.path.join(REF, "input")),
              "param": parse_namelist(os.path.join(REF, "param_scott")),
              "initial_energy_per_atom": -0.00420,
              "ewald_printed": {"wrcut_bohr": 13.1191, "walpha": 0.2395, "rkmax": 1.5046, "kmax": 7,
                                "rcut_bohr": 14.679, "box_bohr": 29.358}}
    with open(os.path.join(OUT, "gentest_params.json"), "w") as f:
        json.dump(params, f, indent=1)
    # the GLE drift matrix shipped with the reference (params/GLE-A: ns, then (ns+1) rows)
    rows = open("/root/reference/params/GLE-A").read().split("\n")
    ns = int(rows[0].split()[0])
    A = np.array([[float(x) for x in rows[1 + i].split()] for i in range(ns + 1)])
    np.savetxt(os.path.join(OUT, "gle_A_ns4.txt"), A, fmt="%.9e")
    allf = {}
    for kind in ("xyz", "frc"):
        per_bead = []
        for b in (1, 2):
            fr, box = read_frames(os.path.join(REF, f"vmd_bead-{b}.{kind}"))
            per_bead.append(fr)
            if kind == "xyz" and b == 1:
                allf["box_A"] = box
        allf[kind] = np.stack(per_bead, axis=1)  # (24, bead, atom, 3)
    np.savez_compressed(os.path.join(OUT, "gentest_traj.npz"), **allf)
    txt = os.path.join(OUT, "gentest_text")
    os.makedirs(txt, exist_ok=True)
    with open(os.path.join(REF, "eq.xyz")) as f, open(os.path.join(txt, "eq.xyz"), "w") as g:
        g.write(f.read())
    for b in (1, 2):
        for kind in ("xyz", "frc"):
            lines = open(os.path.join(REF, f"vmd_bead-{b}.{kind}")).read().split("\n")
            n = int(lines[0].split()[0])
            with open(os.path.join(txt, f"vmd_bead-{b}.{kind}.frame0"), "w") as g:
                g.write("\n".join(lines[:n + 2]) + "\n")
    print("wrote", os.path.join(OUT, "gentest_frames.npz"), os.path.join(OUT, "gentest_traj.npz"), txt)


if __name__ == "__main__":
    main()
