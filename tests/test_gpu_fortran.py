"""The Fortran host layer (rpmd_mainz_amd/fortran/rpmd_hip.f90) driving whole steps: the
reference-named evolve_pi drop-in with the thermostat handed over by rpmd_set_thermostat (PILE-L
and GLE, the reference's ran2 deviates replayed, GLE auxiliary momenta from the oracle's
therm_gle_init), and the device-resident API through its bind(C) interfaces.  Each is compared
bit for bit with the same calls from Python through the C-ABI, and with the oracle's evolve_pi."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle.oracle import RpmdParams
from rpmd_mainz_amd import configs as cf

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def harness():
    from rpmd_mainz_amd.build import HARNESS
    if not os.path.exists(HARNESS):
        pytest.skip("Fortran harness not built (amdflang absent at build time)")
    return HARNESS


def _header(f, mode, na, nb, iopt, prm, box, z, r):
    f.write(struct.pack("4i", mode, na, nb, iopt))
    f.write(struct.pack("11d", prm["oo_eps"], prm["oo_sig"], prm["oo_gam"], prm["rcut"], prm["alpha"], prm["ecut"],
                        prm["wm"], prm["wh"], prm["wrcut"], prm["walpha"], prm["rkmax"]))
    f.write(struct.pack("2i", prm["kmax"], prm["iamcub"]))
    f.write(struct.pack("7d", prm["apot"], prm["bpot"], prm["alp"], prm["alpb"], prm["theta"], prm["reoh"],
                        prm["sig"]))
    f.write(struct.pack("3i", prm["nbdf1"], prm["nbdf2"], prm["mts"]))
    for a in (box, z, r):
        f.write(np.ascontiguousarray(a, np.float64).tobytes())


def _step_input(path, mode, iopt, prm, box, z, r, therm, nsteps, dt, beta, mass, p, d1, d2, noise=None, A=None,
                gp=None):
    nb, na = r.shape[:2]
    ns = -1 if A is None else A.shape[0] - 1
    noise = np.zeros(0) if noise is None else np.asarray(noise, np.float64)
    with open(path, "wb") as f:
        _header(f, mode, na, nb, iopt, prm, box, z, r)
        f.write(therm.ljust(3)[:3].encode())
        f.write(struct.pack("3i", nsteps, ns, 0 if gp is None else 1))
        f.write(struct.pack("3d", prm["ttau"], dt, beta))
        f.write(struct.pack("q", noise.size))
        for a in (mass, p, d1, d2, noise):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        if A is not None:
            f.write(np.ascontiguousarray(A, np.float64).tobytes())   # rows
            if gp is not None:
                f.write(np.ascontiguousarray(gp, np.float64).tobytes())


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("therm", ["PRO", "GLE"])
def test_fortran_evolve_pi_drop_in(capi, oracle, harness, tmp_path, therm):
    """One evolve_pi step through the Fortran drop-in (thermostat from rpmd_set_thermostat, the
    reference's deviates replayed) equals the C-ABI step from Python bit for bit and the oracle's
    step to 1e-12 (positions) / 1e-10 (momenta)."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8, ecut=1e-6, mts=2, therm=therm, ttaufs=100.0)
    nb, na = r.shape[:2]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(4711))
    ff = oracle.full_forces(P, r, box, z)
    s_or, s_gp = oracle.rng(99), oracle.rng(99)
    g, A, gp, n1 = None, None, None, 1
    if therm == "GLE":
        A = np.loadtxt(os.path.join(GOLD, "gle_A_ns4.txt"))
        n1 = A.shape[0]
        g = oracle.gle_init(A, prm["ttau"], na, nb, 0.5 * dt, s_or, beta)
        oracle.gaussians(s_gp, n1 * 3 * na * nb)
        gp = oracle.gle_channels(g)
    per_step = (2 * prm["mts"] if therm == "PRO" else 2 * n1) * 3 * na * nb
    noise = oracle.gaussians(s_gp, per_step)
    o = oracle.evolve_pi(P, p, r, ff["dvdr"], ff["dvdr2"], dt, mass, box, z, beta, s=s_or, g=g)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _step_input(fin, 3, 7, prm, box, z, r, therm, 1, dt, beta, mass, p, ff["dvdr"], ff["dvdr2"], noise, A, gp)
    print(f"case fortran evolve_pi {therm}", flush=True)
    subprocess.run([harness, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, dtype=np.float64)
    n = 3 * na * nb
    v_f = out[0]
    pf, rf = out[22:22 + n].reshape(nb, na, 3), out[22 + n:22 + 2 * n].reshape(nb, na, 3)
    d1f = out[22 + 2 * n:22 + 3 * n].reshape(nb, na, 3)
    # the same step from Python through the C-ABI
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    if therm == "GLE":
        ctx.gle_init(A, na, nb, 0.5 * dt, beta, gp=gp)
    pc, rc, dc1, dc2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    c = ctx.evolve_pi(pc, rc, dc1, dc2, dt, mass, na, nb, box, z, beta, noise=noise)
    ctx.close()
    assert np.array_equal(pf, pc) and np.array_equal(rf, rc) and np.array_equal(d1f, dc1) and v_f == c["v"]
    assert _rel(rf, o["r"]) < 1e-12 and _rel(pf, o["p"]) < 1e-10
    assert abs(v_f - o["v"]) < 1e-10 * abs(o["v"])
    if g is not None:
        oracle.gle_free(g)


def test_fortran_device_resident_api(capi, harness, tmp_path):
    """state_upload / use_graph / run_steps / state_download through the bind(C) interfaces from a
    Fortran program: the same bits as the same calls from Python (PILE-L, device noise stream)."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8, ecut=1e-6, therm="PRO", ttaufs=100.0)
    nb, na = r.shape[:2]
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = np.random.default_rng(5).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    zero = np.zeros_like(r)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _step_input(fin, 4, 31, prm, box, z, r, "PRO", 5, dt, beta, mass, p, zero, zero)
    print("case fortran device-resident api", flush=True)
    subprocess.run([harness, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, dtype=np.float64)
    n = 3 * na * nb
    en = out[:4]
    pf, rf = out[13:13 + n].reshape(nb, na, 3), out[13 + n:13 + 2 * n].reshape(nb, na, 3)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=31)
    ctx.use_graph(True)
    ctx.run_steps(5)
    st = ctx.state_download()
    ctx.close()
    assert np.array_equal(pf, st["p"]) and np.array_equal(rf, st["r"]) and en[0] == st["v"]


def test_fortran_evolve_pi_npt_drop_in(capi, oracle, harness, tmp_path):
    """evolve_pi with nbaro = 1 through the Fortran drop-in (rpmd_set_barostat 'MCI', the reference's
    Gaussian deviates and ran2 uniforms replayed): bit for bit the C-ABI step from Python, the
    oracle's step to 1e-12; the box and the move counters as the oracle's."""
    from oracle.oracle import OrcBaro
    r, box, z, mass, prm = cf.water_system("qspcfw", (3, 3, 3), 8, ecut=1e-6, mts=2, therm="PRO", ttaufs=100.0)
    nb, na = r.shape[:2]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.5 * cf.TOFS
    taub, patm, dv = 1.0, 1e-8, 40.0
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(4711))
    ff = oracle.full_forces(P, r, box, z)
    s_or, s_gp = oracle.rng(98), oracle.rng(98)
    noise = oracle.gaussians(s_gp, 2 * prm["mts"] * 3 * na * nb)
    us = np.array([oracle.ran2(s_gp, 0.0, 1.0), oracle.ran2(s_gp, -1.0, 1.0), oracle.ran2(s_gp, 0.0, 1.0)])
    vl0 = np.diag([1e-3, 2e-3, -1e-3])
    B = OrcBaro(2, taub, patm, dv, 0, 0)
    o = oracle.evolve_pi_npt(P, p, r, ff["dvdr"], ff["dvdr2"], dt, mass, box, z, beta, vl0, B, s=s_or)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _step_input(fin, 5, 7, prm, box, z, r, "PRO", 1, dt, beta, mass, p, ff["dvdr"], ff["dvdr2"], noise)
    with open(fin, "ab") as f:
        f.write(b"MCI")
        f.write(struct.pack("3d", taub, patm, dv))
        f.write(struct.pack("i", us.size))
        f.write(us.tobytes())
        f.write(np.ascontiguousarray(vl0.T).tobytes())
    print("case fortran evolve_pi nbaro=1 MCI", flush=True)
    subprocess.run([harness, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, dtype=np.float64)
    n = 3 * na * nb
    pf, rf = out[22:22 + n].reshape(nb, na, 3), out[22 + n:22 + 2 * n].reshape(nb, na, 3)
    boxf, cnt = out[22 + 4 * n:22 + 4 * n + 3], out[22 + 4 * n + 3:]
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    ctx.set_barostat("MCI", taub, patm, dv)
    pc, rc, dc1, dc2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    c = ctx.evolve_pi_npt(pc, rc, dc1, dc2, dt, mass, na, nb, box, z, beta, vl0, noise=noise, uniforms=us)
    counts = ctx.barostat_counts()
    ctx.close()
    assert np.array_equal(pf, pc) and np.array_equal(rf, rc) and out[0] == c["v"] and np.array_equal(boxf, c["box"])
    assert tuple(int(x) for x in cnt) == counts == (B.nacc, B.nmove) == (1, 1)
    assert _rel(rf, o["r"]) < 1e-12 and _rel(pf, o["p"]) < 1e-10 and _rel(boxf, o["box"]) < 1e-14
