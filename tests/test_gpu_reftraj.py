"""The force-only drivers of rpmd_mainz_amd/reftraj.py on the GPU.

* All 24 frames of the reference's gentest run (test/cmp/gentest.cmp: q-TIP4P/F-type model,
  125 molecules, 2 beads, RP contraction of Ewald and O-O to the centroid, TEM correction, mts 5)
  as a two-bead reference trajectory in the reference's file layout, replayed through the RPMD
  step's forces: every bead's forces -(dvdr + dvdr2) against the forces the reference printed for
  that frame (vmd_bead-{1,2}.frc), relative RMS < 1e-11 per frame; the written force dumps read
  back to the same numbers.
* The force-matching driver on single-bead frames: the whole gradient and its Coulomb, O-O,
  stretch and bend parts against the oracle's forces (iopt 0, 2, 3 and 4 with bpot or apot set to
  zero), and the written vmd.frc* files.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import RpmdParams
from rpmd_mainz_amd import configs as cf
from rpmd_mainz_amd import reftraj as rt

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gentest_params():
    par = json.load(open(os.path.join(GOLD, "gentest_params.json")))
    model = {k: float(v) for k, v in par["param"].items()}
    box0 = cf.box_size(125, float(par["input"]["rho"]), model["wmass"], (5, 5, 5))
    prm = cf.make_params(model, box0, rcut_A=9.0, ecut=0.0, sig_A=4.5, nbdf1=1, nbdf2=1, mts=5)
    prm = cf.setup_ewald(375, box0, prm)
    mass, z = cf.masses_charges(model, 375)
    return prm, mass, z


def _write_traj(path, fx, frames):
    """the frames as a reference trajectory: all frames of bead 1, then of bead 2, positions and box
    in Angstrom as the dumps print them"""
    with open(path, "w") as f:
        for b in range(fx["xyz"].shape[1]):
            for i in frames:
                x = fx["xyz"][i, b]
                f.write(f"{x.shape[0]:12d}\n BOX " + " ".join(repr(float(v)) for v in fx["box_A"][i]) + "\n")
                for a in range(x.shape[0]):
                    f.write(" " + "OHH"[a % 3] + " " + " ".join(repr(float(v)) for v in x[a]) + "\n")


def test_replay_all_gentest_frames(capi, tmp_path):
    fx = np.load(os.path.join(GOLD, "gentest_traj.npz"))
    nf = fx["xyz"].shape[0]
    assert nf == 24
    prm, mass, z = _gentest_params()
    path = tmp_path / "traj.xyz"
    _write_traj(path, fx, range(nf))
    r_traj, boxes = rt.read_reftraj(str(path), nb=2, nframes=nf, na=375)
    ctx = capi.RPMDContext(prm, 375, 2, "fp64")
    print("case replay 24 gentest frames", flush=True)
    out = rt.replay_snapshots(ctx, r_traj, boxes, z, out_dir=str(tmp_path / "out"), unwrap=True)
    ctx.close()
    worst = 0.0
    for i in range(nf):
        F = -(out["dvdr"][i] + out["dvdr2"][i])
        ref = fx["frc"][i]
        rms = math.sqrt(((F - ref) ** 2).sum() / (ref ** 2).sum())
        worst = max(worst, rms)
        assert rms < 1e-11, (i, rms)
    print(f"  worst frame force RMS_rel {worst:.2e}", flush=True)
    assert np.all(np.isfinite(out["v"])) and np.all(out["v_kjmol_per_molecule"] < 0.0)
    for b in (1, 2):
        fr = rt.read_xyz_frames(str(tmp_path / "out" / f"vmd_bead-{b}.frc"))
        assert len(fr) == nf
        for i in range(nf):
            assert np.array_equal(fr[i]["xyz"], -(out["dvdr"][i, b - 1] + out["dvdr2"][i, b - 1]))


def test_force_match_parts_against_oracle(capi, oracle, tmp_path):
    fx = np.load(os.path.join(GOLD, "gentest_traj.npz"))
    prm, mass, z = _gentest_params()
    frames = [0, 7, 23]
    r_traj = np.stack([fx["xyz"][i, :1] * (1.0 / cf.TOA) for i in frames])
    boxes = np.stack([fx["box_A"][i] * (1.0 / cf.TOA) for i in frames])
    ctx = capi.RPMDContext(prm, 375, 1, "fp64")
    print("case force matching 3 frames", flush=True)
    out = rt.force_match(ctx, prm, r_traj, boxes, z, out_dir=str(tmp_path))
    ctx.close()
    variants = [(0, prm), (2, prm), (3, prm), (4, dict(prm, bpot=0.0)), (4, dict(prm, apot=0.0))]

    def rel(a, b):
        return math.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300))

    files = {n: rt.read_xyz_frames(str(tmp_path / n)) for n in ["vmd.frc", "vmd.frc1", "vmd.frc2", "vmd.frc3",
                                                                 "vmd.frc4", "vmd.frcT"]}
    for i in range(len(frames)):
        got = [out["total"][i]] + [out["parts"][j, i] for j in range(4)]
        for (iopt, pj), g in zip(variants, got):
            v, d, _ = oracle.forces(RpmdParams.from_dict(pj), r_traj[i], boxes[i], z, iopt)
            assert rel(g, d) < 1e-12, (frames[i], iopt, rel(g, d))
        # the four parts add up to the whole (to rounding), and the files hold the centroid forces
        assert rel(out["parts"][:, i].sum(axis=0), out["total"][i]) < 1e-13
        assert np.array_equal(files["vmd.frc"][i]["xyz"], -out["total"][i][0])
        for j in range(4):
            assert np.array_equal(files[f"vmd.frc{j + 1}"][i]["xyz"], -out["parts"][j, i][0])
        assert np.allclose(files["vmd.frcT"][i]["xyz"], -out["total"][i][0], rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        rt.force_match(None, dict(prm, alp=0.0), r_traj, boxes, z)
