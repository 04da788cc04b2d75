"""GPU parity: the HIP path through the C-ABI against the oracle and the reference's golden frames.

FP64 mode: energies, gradients and virials within 1e-11 relative of the oracle (summation order
differs; the oracle itself matches the reference Fortran to 1e-15).  Mixed mode: force RMS error
relative to the oracle's total force <= 1e-5 (the bar of the FP64 vs mixed comparison).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import RpmdParams
from rpmd_mainz_amd import configs as cf

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def capi():
    from rpmd_mainz_amd import capi
    return capi


SYSTEMS = {
    "spcfw_list": ("qspcfw", (3, 3, 3), 4, (1, 1, 1), {}),
    "tip4pf_tem": ("qtip4pf", (4, 4, 4), 5, (1, 1, 1), dict(sig_A=4.5)),
    "tip4pf_nc": ("qtip4pf", (4, 4, 4), 2, (1, 1, 1.3), dict(sig_A=4.5)),
    "spcfw_ecut": ("qspcfw", (6, 6, 6), 4, (1, 1, 1), dict(ecut=1e-6)),
    "spcfw_nc1": ("qspcfw", (6, 6, 6), 1, (1, 1, 1.2), {}),
    "spcfw_model": ("qspcfw", (5, 5, 5), 2, (1, 1, 1), dict(intra_model=cf.INTRA_SPCFW)),
    "buck_list": ("qspcfw", (4, 4, 4), 4, (1, 1, 1), dict(oo_gam=13.5)),
    "buck_nc": ("qspcfw", (5, 5, 5), 1, (1, 1, 1.2), dict(oo_gam=13.5)),
    "smear_cubic": ("smear", (4, 4, 4), 2, (1, 1, 1), {}),
    "smear_nc": ("smear", (6, 6, 6), 3, (1, 1, 1.2), {}),
    "smear_nc1": ("smear", (3, 3, 3), 2, (1, 1, 1.3), {}),
}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _escale(prm, box, z, v):
    """Magnitude of the summands of an Ewald energy: the self term alpha/sqrt(pi) sum q^2 (per
    bead); the total is a small difference of such terms, so round-off scales with this."""
    sc = 1.0 / box[0] if prm["iamcub"] else 1.0
    return max(abs(v), prm["walpha"] / math.sqrt(math.pi) * float(np.sum(np.asarray(z) ** 2)) * sc)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_forces_fp64_match_oracle(capi, oracle, name):
    model, ncell, nb, aspect, extra = SYSTEMS[name]
    r, box, z, mass, prm = cf.water_system(model, ncell, nb, aspect=aspect, **extra)
    na = r.shape[1]
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    P = RpmdParams.from_dict(prm)
    for iopt in (0, 1, 2, 3, 4, 5):
        if iopt == 5 and prm["sig"] == 0.0:
            continue
        print(f"case {name} iopt {iopt} na {na} nb {nb}", flush=True)
        v1, d1, w1 = ctx.forces(r, nb, na, box, z, iopt)
        v2, d2, w2 = oracle.forces(P, r, box, z, iopt)
        assert abs(v1 - v2) <= 1e-12 * _escale(prm, box, z, v2), (iopt, v1, v2)
        assert _rel(d1, d2) <= 1e-11, (iopt, _rel(d1, d2))
        assert _rel(w1, w2) <= 1e-10, (iopt, _rel(w1, w2))
    ctx.close()


def test_forces_config2_shape(capi, oracle):
    """1728 q-SPC/Fw, cubic ecut=1e-6 (kmax 15, 6^3 reference Ewald cells) at 2 beads."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (12, 12, 12), 2, ecut=1e-6, intra_model=cf.INTRA_SPCFW)
    na = r.shape[1]
    assert prm["kmax"] == 15
    ctx = capi.RPMDContext(prm, na, 2, "fp64")
    P = RpmdParams.from_dict(prm)
    for iopt in (0, 2):
        print(f"case config2 iopt {iopt}", flush=True)
        v1, d1, w1 = ctx.forces(r, 2, na, box, z, iopt)
        v2, d2, w2 = oracle.forces(P, r, box, z, iopt)
        assert abs(v1 - v2) <= 1e-12 * _escale(prm, box, z, v2)
        assert _rel(d1, d2) <= 1e-11
    # mixed precision against the FP64 oracle: relative force RMS
    ctx.set_precision("mixed")
    v3, d3, w3 = ctx.forces(r, 2, na, box, z, 0)
    v2, d2, w2 = oracle.forces(P, r, box, z, 0)
    rms = math.sqrt(((d3 - d2) ** 2).sum() / (d2 ** 2).sum())
    print("mixed force RMS_rel", rms, "energy rel", abs(v3 - v2) / abs(v2), flush=True)
    assert rms <= 1e-5
    ctx.close()


def _gentest():
    fx = np.load(os.path.join(GOLD, "gentest_frames.npz"))
    par = json.load(open(os.path.join(GOLD, "gentest_params.json")))
    model = {k: float(v) for k, v in par["param"].items()}
    box0 = cf.box_size(125, float(par["input"]["rho"]), model["wmass"], (5, 5, 5))
    prm = cf.make_params(model, box0, rcut_A=9.0, ecut=0.0, sig_A=4.5, nbdf1=1, nbdf2=1, mts=5)
    prm = cf.setup_ewald(375, box0, prm)
    mass, z = cf.masses_charges(model, 375)
    return fx, prm, mass, z


def _frame(fx, i):
    box = fx["box_A"][i] / cf.TOA
    r = fx["xyz"][i] / cf.TOA
    d = r[1] - r[0]
    d -= box * np.round(d / box)
    r[1] = r[0] + d
    return np.ascontiguousarray(r), box


def test_full_forces_gentest_frames(capi):
    fx, prm, mass, z = _gentest()
    ctx = capi.RPMDContext(prm, 375, 2, "fp64")
    for i in range(len(fx["frames"])):
        print(f"case gentest frame {i}", flush=True)
        r, box = _frame(fx, i)
        out = ctx.full_forces(r, 375, 2, z, box)
        F = -(out["dvdr"] + out["dvdr2"])
        ref = fx["frc"][i]
        rms = math.sqrt(((F - ref) ** 2).sum() / (ref ** 2).sum())
        assert rms < 1e-11, (i, rms)
    ctx.close()


def test_evolve_pi_gentest_nve_window(capi):
    """20 NVE evolve_pi steps on the GPU (mts 5, RP contraction to 1 mode, TEM) reproduce the
    reference's next printed bead frame."""
    fx, prm, mass, z = _gentest()
    ctx = capi.RPMDContext(prm, 375, 2, "fp64")
    r, box = _frame(fx, 0)
    p = np.ascontiguousarray(fx["vel"][0] * mass[None, :, None])
    out = ctx.full_forces(r, 375, 2, z, box)
    dvdr, dvdr2 = out["dvdr"], out["dvdr2"]
    dt, beta = 0.5 * cf.TOFS, 1.0 / (cf.KBOLTZ * 298.0)
    for k in range(20):
        print(f"case gentest nve step {k}", flush=True)
        ctx.evolve_pi(p, r, dvdr, dvdr2, dt, mass, 375, 2, box, z, beta)
    dd = r - fx["xyz"][1] / cf.TOA
    dd -= box * np.round(dd / box)
    assert np.abs(dd).max() < 1e-10
    v, vref = p / mass[None, :, None], fx["vel"][1]
    assert _rel(v, vref) < 1e-10
    ctx.close()


def test_realft_matches_oracle(capi, oracle):
    ctx = capi.RPMDContext(cf.water_system("qspcfw", (2, 2, 2), 1)[4], 24, 64, "fp64")
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 7, 8, 32, 64):
        x = rng.normal(size=(n, 24 * 3))
        for mode in (1, -1):
            print(f"case realft n {n} mode {mode}", flush=True)
            y = ctx.realft(x.copy(), 72, n, mode)
            yo = oracle.realft(x, mode)
            assert np.abs(y - yo).max() < 1e-14 * max(1.0, np.abs(yo).max())
    with pytest.raises(capi.RPMDError):
        ctx.realft(np.zeros((2, 3)), 3, 2, 0)
    ctx.close()


def test_evolve_pi_pile_replay(capi, oracle):
    """PILE-L (PRO) step with the reference's ran2/gaussian stream replayed on the device."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8, ecut=1e-6, mts=2, therm="PRO", ttaufs=100.0)
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    s0 = oracle.rng(4711)
    p = oracle.sample_rpmd(nb, na, mass, beta, s0)
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    po, ro, do1, do2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    pg, rg, dg1, dg2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    s_or = oracle.rng(99)
    s_gp = oracle.rng(99)
    for k in range(3):
        print(f"case pile replay step {k}", flush=True)
        noise = oracle.gaussians(s_gp, 2 * prm["mts"] * 3 * na * nb)
        o = oracle.evolve_pi(P, po, ro, do1, do2, dt, mass, box, z, beta, s=s_or)
        po, ro, do1, do2 = o["p"], o["r"], o["dvdr"], o["dvdr2"]
        g = ctx.evolve_pi(pg, rg, dg1, dg2, dt, mass, na, nb, box, z, beta, noise=noise)
        assert abs(g["v"] - o["v"]) < 1e-10 * abs(o["v"])
    assert _rel(rg, ro) < 1e-12
    assert _rel(pg, po) < 1e-10
    ctx.close()


def _noise_per_step(therm, na, nb, mts, n1=1):
    """Gaussian deviates one evolve_pi step draws (src/evolve_pi.f90:75-77, 100-108, 304-310)."""
    d = 3 * na * nb
    return {"PRO": 2 * mts * d, "PRL": 2 * d, "PRG": 2 * (d - 2), "PRM": 2 * mts * (d - 2),
            "GLE": 2 * n1 * d}[therm]


@pytest.mark.parametrize("therm", ["PRG", "PRM", "PRL", "GLE"])
def test_evolve_pi_thermostat_replay(capi, oracle, therm):
    """Bussi global (PRG), PILE-G (PRM), local Langevin (PRL) and GLE steps on the device with the
    reference's ran2/gaussian stream replayed, against the oracle's evolve_pi."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8, ecut=1e-6, mts=2, therm=therm, ttaufs=100.0)
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    s0 = oracle.rng(4711)
    p = oracle.sample_rpmd(nb, na, mass, beta, s0)
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    s_or, s_gp = oracle.rng(99), oracle.rng(99)
    g, n1 = None, 1
    if therm == "GLE":
        A = np.loadtxt(os.path.join(GOLD, "gle_A_ns4.txt"))
        n1 = A.shape[0]
        g = oracle.gle_init(A, prm["ttau"], na, nb, 0.5 * dt, s_or, beta)
        oracle.gaussians(s_gp, n1 * 3 * na * nb)          # the draws of therm_gle_init
        ctx.gle_init(A, na, nb, 0.5 * dt, beta, gp=oracle.gle_channels(g))
    po, ro, do1, do2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    pg, rg, dg1, dg2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    for k in range(3):
        print(f"case {therm} replay step {k}", flush=True)
        noise = oracle.gaussians(s_gp, _noise_per_step(therm, na, nb, prm["mts"], n1))
        o = oracle.evolve_pi(P, po, ro, do1, do2, dt, mass, box, z, beta, s=s_or, g=g)
        po, ro, do1, do2 = o["p"], o["r"], o["dvdr"], o["dvdr2"]
        out = ctx.evolve_pi(pg, rg, dg1, dg2, dt, mass, na, nb, box, z, beta, noise=noise)
        assert abs(out["v"] - o["v"]) < 1e-10 * abs(o["v"])
    assert _rel(rg, ro) < 1e-12
    assert _rel(pg, po) < 1e-10
    if g is not None:
        oracle.gle_free(g)
    ctx.close()


@pytest.mark.parametrize("therm", ["PRO", "PRM", "PRG", "GLE"])
def test_device_noise_thermostats_sample_temperature(capi, therm):
    """Device noise stream: from a cool start (0.3 of the thermal momenta) the bead kinetic energy of a ring polymer of
    molecules relaxes towards the canonical value 3 na nb / (2 beta_n) within a few tau; a check
    that every thermostat couples with the right sign and variance, run through run_steps
    (hipGraph)."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (3, 3, 3), 8, ecut=1e-6, therm=therm, ttaufs=10.0,
                                           intra_model=cf.INTRA_SPCFW)
    na, nb = r.shape[1], r.shape[0]
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    p0 = 0.3 * np.random.default_rng(1).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    ctx.state_upload(p0, r, mass, z, na, nb, box, beta, dt, seed=2024)
    if therm == "GLE":
        A = np.loadtxt(os.path.join(GOLD, "gle_A_ns4.txt"))
        ctx.gle_init(A, na, nb, 0.5 * dt, beta, seed=5)
    ctx.use_graph(True)
    kin = []
    ctx.run_steps(400)
    for _ in range(20):
        ctx.run_steps(20)
        st = ctx.state_download()
        kin.append(float(np.sum(st["p"] ** 2 / (2.0 * mass[None, :, None]))))
    target = 3 * na * nb / (2.0 * beta / nb)
    ratio = np.mean(kin) / target
    print(f"{therm}: <K>/K_target = {ratio:.3f}")
    assert 0.8 < ratio < 1.2
    ctx.close()


def test_config1_pile_step_by_step(capi, oracle):
    """Config 1: 216 q-SPC/Fw (6^3), 32 beads, PILE-L 298 K (tau 100 fs), dt 0.25 fs, ecut 1e-6,
    FP64; 20 steps with the reference's ran2 stream replayed, energies and forces compared with
    the oracle after every step."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (6, 6, 6), 32, ecut=1e-6, therm="PRO", ttaufs=100.0,
                                           intra_model=cf.INTRA_SPCFW, amp=0.0)
    assert prm["kmax"] == 10 and abs(box[0] * cf.TOA - 18.643) < 0.01
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(1))
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    s_or, s_gp = oracle.rng(2), oracle.rng(2)
    po, ro, d1, d2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    pg, rg, g1, g2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    for k in range(20):
        noise = oracle.gaussians(s_gp, 2 * 3 * na * nb)
        o = oracle.evolve_pi(P, po, ro, d1, d2, dt, mass, box, z, beta, s=s_or)
        po, ro, d1, d2 = o["p"], o["r"], o["dvdr"], o["dvdr2"]
        g = ctx.evolve_pi(pg, rg, g1, g2, dt, mass, na, nb, box, z, beta, noise=noise)
        assert abs(g["v"] - o["v"]) < 1e-10 * _escale(prm, box, z, o["v"]), k
        assert _rel(g1, d1) < 1e-9 and _rel(g2, d2) < 1e-9, k
    assert _rel(rg, ro) < 1e-11
    ctx.close()


def test_sample_rpmd_matches_oracle(capi, oracle):
    """sample_rpmd + pshift on the device: the reference's stream replayed gives the oracle's
    momenta; the device stream gives zero total momentum and the Maxwell-Boltzmann variance."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8)
    na, nb = r.shape[1], r.shape[0]
    beta = 1.0 / (cf.KBOLTZ * 298.0)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    noise = oracle.gaussians(oracle.rng(21), 3 * na * nb)
    pg = ctx.sample_rpmd(nb, na, mass, beta, noise=noise)
    po = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(21))
    assert _rel(pg, po) < 1e-14
    pd = ctx.sample_rpmd(nb, na, mass, beta, seed=3)
    assert np.abs(pd.sum(axis=(0, 1))).max() < 1e-9 * np.abs(pd).sum()
    var = float(np.mean(pd ** 2 / (mass[None, :, None] / (beta / nb))))
    assert abs(var - 1.0) < 0.05
    ctx.close()


def test_device_resident_steps_match_oracle(capi, oracle):
    """NVE steps run with all state in HBM agree with the oracle's evolve_pi."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8, ecut=1e-6, intra_model=cf.INTRA_SPCFW)
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = np.random.default_rng(7).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt)
    ctx.run_steps(4)
    g = ctx.state_download()
    po, ro, d1, d2 = p, r, ff["dvdr"], ff["dvdr2"]
    for _ in range(4):
        o = oracle.evolve_pi(P, po, ro, d1, d2, dt, mass, box, z, beta)
        po, ro, d1, d2 = o["p"], o["r"], o["dvdr"], o["dvdr2"]
    assert _rel(g["r"], ro) < 1e-12
    assert _rel(g["p"], po) < 1e-10
    assert abs(g["v"] - o["v"]) < 1e-10 * abs(o["v"])
    ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_stream_layouts_give_identical_steps(capi, prec):
    """Device-resident PILE-L steps (the fused step: O-O Verlet list, both Ewald branches,
    intramolecular forces) give the same bits on one, two or three streams, launched eagerly or
    replayed from a graph: the streams change only where kernels run, not what they add."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 8, ecut=1e-6, intra_model=cf.INTRA_SPCFW,
                                           therm="PRO", ttaufs=100.0)
    na, nb = r.shape[1], r.shape[0]
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = np.random.default_rng(11).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    out = []
    for streams, graph in [(1, False), (2, True), (3, False), (3, True)]:
        print(f"case streams {streams} graph {graph} {prec}", flush=True)
        ctx = capi.RPMDContext(prm, na, nb, prec)
        ctx.set_streams(streams)
        ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=11)
        ctx.use_graph(graph)
        ctx.run_steps(6)
        out.append(ctx.state_download())
        ctx.close()
    for g in out[1:]:
        assert np.array_equal(g["r"], out[0]["r"])
        assert np.array_equal(g["p"], out[0]["p"])
        assert g["v"] == out[0]["v"]


def test_contracted_mts_noncubic_matches_oracle(capi, oracle):
    """Config-4 form at small size: q-SPC/Fw in a 1:1:1.2 box (non-cubic Ewald), 64 beads, mts 4,
    Ewald on 7 contracted modes and LJ on 15 (odd, as src/evolve_pi.f90:344-347 requires): two
    evolve_pi steps on the device against the oracle."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), 64, aspect=(1, 1, 1.2), nbdf1=7, nbdf2=15, mts=4,
                                           intra_model=cf.INTRA_SPCFW)
    assert prm["iamcub"] == 0
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 1.0 * cf.TOFS
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(3))
    ff = oracle.full_forces(P, r, box, z)
    g = capi.RPMDContext(prm, na, nb, "fp64").full_forces(r, na, nb, z, box)
    assert _rel(g["dvdr"], ff["dvdr"]) < 1e-11 and _rel(g["dvdr2"], ff["dvdr2"]) < 1e-11
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    po, ro, d1, d2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    pg, rg, g1, g2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    for k in range(2):
        print(f"case contracted mts step {k}", flush=True)
        o = oracle.evolve_pi(P, po, ro, d1, d2, dt, mass, box, z, beta)
        po, ro, d1, d2 = o["p"], o["r"], o["dvdr"], o["dvdr2"]
        out = ctx.evolve_pi(pg, rg, g1, g2, dt, mass, na, nb, box, z, beta)
        assert abs(out["v"] - o["v"]) < 1e-10 * _escale(prm, box, z, o["v"])
    assert _rel(rg, ro) < 1e-12
    assert _rel(pg, po) < 1e-10
    ctx.close()


def test_config4_shape_device_resident(capi):
    """Config-4 shape: 4096 q-SPC/Fw (16^3) in a 1:1:1.2 box, 64 beads, mts 4, Ewald on 7 modes
    and LJ on 15, mixed precision, stepped in HBM; the contracted forces of the mixed path agree
    with the FP64 path to the 1e-5 RMS bar."""
    r, box, z, mass, prm = cf.water_system("qspcfw", (16, 16, 16), 64, aspect=(1, 1, 1.2), nbdf1=7, nbdf2=15,
                                           mts=4, intra_model=cf.INTRA_SPCFW, amp=0.02)
    na, nb = r.shape[1], r.shape[0]
    assert prm["kmax"] <= 15 and prm["iamcub"] == 0
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 1.0 * cf.TOFS
    ctx = capi.RPMDContext(prm, na, nb, "mixed")
    f_mx = ctx.full_forces(r, na, nb, z, box)
    ctx.set_precision("fp64")
    f_64 = ctx.full_forces(r, na, nb, z, box)
    err = math.sqrt(((f_mx["dvdr"] - f_64["dvdr"]) ** 2).sum() / (f_64["dvdr"] ** 2).sum())
    print(f"config4 contracted LF force rms rel error mixed vs fp64: {err:.3e}")
    assert err < 1e-5
    ctx.set_precision("mixed")
    p = np.random.default_rng(4).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=3)
    ctx.run_steps(2)
    st = ctx.state_download()
    assert np.isfinite(st["v"]) and np.isfinite(st["p"]).all()
    ctx.close()


def _replica_specs(n, seed, sizes=(128, 256, 512, 1024), beads=(8, 16, 32, 64)):
    """Config-3 style draws: molecule counts, bead counts, densities 0.95-1.05 g/cm^3 and
    temperatures 250-350 K from a seed; every replica in its own cubic box, filled by a simple,
    body- or face-centred cubic lattice."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nm = int(rng.choice(sizes))
        ncell, basis = cf.cubic_cells(nm)
        out.append(dict(nm=nm, ncell=ncell, basis=basis, nb=int(rng.choice(beads)),
                        rho=float(rng.uniform(0.95, 1.05)), temp=float(rng.uniform(250.0, 350.0)), seed=100 + i))
    return out


def _replica_ctx(capi, spec, prec, therm="PRO"):
    r, box, z, mass, prm = cf.water_system("qspcfw", spec["ncell"], spec["nb"], rho=spec["rho"], basis=spec["basis"],
                                           seed=spec["seed"], ecut=1e-6, intra_model=cf.INTRA_SPCFW, therm=therm,
                                           ttaufs=100.0)
    assert prm["iamcub"] == 1
    na, nb = r.shape[1], r.shape[0]
    beta, dt = 1.0 / (cf.KBOLTZ * spec["temp"]), 0.25 * cf.TOFS
    p = np.random.default_rng(spec["seed"]).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    ctx = capi.RPMDContext(prm, na, nb, prec)
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=spec["seed"])
    return ctx


def test_replica_batch_equals_independent_runs(capi):
    """Uneven replicas (different na, nb, density, temperature) stepped by one batch graph give
    bit-identical states to the same replicas stepped one by one."""
    specs = _replica_specs(5, 7, sizes=(8, 27, 64), beads=(4, 8, 16))
    solo = []
    for sp in specs:
        c = _replica_ctx(capi, sp, "fp64")
        c.use_graph(True)
        c.run_steps(6)
        solo.append(c.state_download())
        c.close()
    ctxs = [_replica_ctx(capi, sp, "fp64") for sp in specs]
    batch = capi.ReplicaBatch(ctxs)
    batch.run_steps(3)
    batch.run_steps(3)
    for c, ref in zip(ctxs, solo):
        st = c.state_download()
        assert np.array_equal(st["r"], ref["r"]) and np.array_equal(st["p"], ref["p"])
        assert st["v"] == ref["v"]
    batch.close()
    for c in ctxs:
        c.close()


def test_replica_batch_config3_shape(capi):
    """Config 3: 48 replicas drawn from a seed (128-1024 molecules, 8-64 beads, own cubic boxes,
    kmax 9-13, 250-350 K), mixed precision, stepped together for 48 steps: no list overflows, every
    state stays finite, near its target temperature (0.8-1.25 of it) and with its ring polymers
    whole (no bead more than 2 bohr from its centroid; thermal spreads reach about 1.2 bohr)."""
    specs = _replica_specs(48, 2024)
    ctxs = [_replica_ctx(capi, sp, "mixed") for sp in specs]
    batch = capi.ReplicaBatch(ctxs)
    for k in range(4):
        batch.run_steps(12)
        print(f"config 3: {12 * (k + 1)} batch steps", flush=True)
    for c, sp in zip(ctxs, specs):
        st = c.state_download()
        assert np.isfinite(st["v"]) and np.isfinite(st["p"]).all()
        assert 9 <= c.prm["kmax"] <= 13
        t = c.estimators()["temperature_K"]
        spread = np.sqrt(((st["r"] - st["r"].mean(0)[None]) ** 2).sum(-1)).max()
        assert 0.8 * sp["temp"] < t < 1.25 * sp["temp"], (sp, t)
        assert spread < 2.0, (sp, spread)
    batch.close()
    for c in ctxs:
        c.close()


def _rigid_case(therm="NONE", ncell=(3, 3, 3), nb=4, seed=1, aspect=(1, 1, 1)):
    r, box, z, mass, prm = cf.water_system("tip4p2005", ncell, nb, ecut=1e-6, rigid=True, therm=therm,
                                           ttaufs=50.0, seed=seed, amp=0.1, aspect=aspect)
    return r, box, z, mass, prm


def _rigid_start(oracle, P, r, mass, beta, dt):
    nb, na = r.shape[:2]
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(11))
    for k in range(nb):
        p[k] = oracle.rattle_s2(r[k], p[k], mass, dt, np.zeros((na, 3)), P)[0]
    return p


def _bond_residual(r, prm):
    rehh = 2.0 * math.sin(0.5 * prm["theta"]) * prm["reoh"]
    m = r.reshape(r.shape[0], -1, 3, 3)
    out = 0.0
    for (i, j, d0) in ((0, 1, prm["reoh"]), (0, 2, prm["reoh"]), (1, 2, rehh)):
        v = m[:, :, j] - m[:, :, i]
        out = max(out, float(np.abs((v * v).sum(-1) - d0 * d0).max()) / (d0 * d0))
    return out


@pytest.mark.parametrize("therm", ["NONE", "PRG", "PRL"])
def test_evolve_rig_pi_matches_oracle(capi, oracle, therm):
    """Rigid TIP4P/2005 RPMD steps (RATTLE on the device, FP64) with the reference's ran2 stream
    replayed, against the oracle's evolve_rig_pi."""
    r, box, z, mass, prm = _rigid_case(therm)
    nb, na = r.shape[:2]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.5 * cf.TOFS
    p = _rigid_start(oracle, P, r, mass, beta, dt)
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    s_or, s_gp = oracle.rng(57), oracle.rng(57)
    nm = na // 3
    per_step = {"NONE": 0, "PRG": 2 * (3 * na * nb - 3 * nm - 2), "PRL": 2 * 3 * na * nb}[therm]
    o = dict(p=p, r=r, dvdr=ff["dvdr"], dvdr2=ff["dvdr2"], vir=np.zeros((3, 3)))
    pg, rg, dg1, dg2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    vg = np.zeros((3, 3))
    for k in range(3):
        print(f"case rigid {therm} step {k}", flush=True)
        noise = oracle.gaussians(s_gp, per_step) if per_step else None
        o = oracle.evolve_rig_pi(P, o["p"], o["r"], o["dvdr"], o["dvdr2"], dt, mass, box, z, beta, o["vir"], s_or)
        g = ctx.evolve_rig_pi(pg, rg, dg1, dg2, dt, mass, na, nb, box, z, beta, vir=vg, noise=noise)
        vg = g["vir"]
        assert abs(g["v"] - o["v"]) < 1e-10 * _escale(prm, box, z, o["v"])
        assert np.abs(vg - o["vir"]).max() < 1e-9 * np.abs(o["vir"]).max()
    assert _rel(rg, o["r"]) < 1e-12
    assert _rel(pg, o["p"]) < 1e-9
    assert _rel(dg1, o["dvdr"]) < 1e-9
    assert _bond_residual(rg, prm) <= 1.0001e-6
    ctx.close()


def test_rigid_device_resident_config5_shape(capi, oracle):
    """The rigid TIP4P/2005 workload shape (1024 molecules, 32 beads, M-site Ewald, mixed-precision
    forces, RATTLE in FP64) stepped in HBM with a hipGraph: every constraint holds to the RATTLE
    tolerance and the bead-averaged energy stays finite; a 125-molecule FP64 run agrees with the
    oracle after 4 steps."""
    r, box, z, mass, prm = _rigid_case("PRG", ncell=(8, 8, 16), nb=32, aspect=(2, 2, 1))
    assert prm["iamcub"] == 1 and abs(box[0] * cf.TOA - 31.318) < 0.01
    nb, na = r.shape[:2]
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.5 * cf.TOFS
    P = RpmdParams.from_dict(prm)
    p = _rigid_start(oracle, P, r, mass, beta, dt)
    ctx = capi.RPMDContext(prm, na, nb, "mixed")
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=9)
    ctx.use_graph(True)
    ctx.run_steps(10)
    st = ctx.state_download()
    assert np.isfinite(st["v"]) and np.isfinite(st["p"]).all()
    assert _bond_residual(st["r"], prm) <= 1.0001e-6
    ctx.close()
    # FP64, small: device-resident steps against the oracle
    r, box, z, mass, prm = _rigid_case("NONE", ncell=(5, 5, 5), nb=4)
    nb, na = r.shape[:2]
    P = RpmdParams.from_dict(prm)
    p = _rigid_start(oracle, P, r, mass, beta, dt)
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt)
    ctx.use_graph(True)
    ctx.run_steps(4)
    g = ctx.state_download()
    o = dict(p=p, r=r, dvdr=ff["dvdr"], dvdr2=ff["dvdr2"], vir=np.zeros((3, 3)))
    for _ in range(4):
        o = oracle.evolve_rig_pi(P, o["p"], o["r"], o["dvdr"], o["dvdr2"], dt, mass, box, z, beta, o["vir"])
    assert _rel(g["r"], o["r"]) < 1e-12
    assert _rel(g["p"], o["p"]) < 1e-9
    assert abs(g["v"] - o["v"]) < 1e-10 * _escale(prm, box, z, o["v"])
    ctx.close()


def test_error_paths(capi):
    r, box, z, mass, prm = cf.water_system("qspcfw", (2, 2, 2), 2)
    ctx = capi.RPMDContext(prm, 24, 2, "fp64")
    with pytest.raises(capi.RPMDError) as e:
        ctx.forces(r, 2, 24, box, z, 7)
    assert e.value.code == 13
    bad = dict(prm, nbdf1=2, nbdf2=0)
    ctx.set_params(bad)
    with pytest.raises(capi.RPMDError) as e:
        ctx.full_forces(r, 24, 2, z, box)
    assert "ODD" in str(e.value)
    ctx.close()


def _write_harness_input(path, mode, na, nb, iopt, prm, box, z, r):
    import struct
    with open(path, "wb") as f:
        f.write(struct.pack("4i", mode, na, nb, iopt))
        f.write(struct.pack("11d", prm["oo_eps"], prm["oo_sig"], prm["oo_gam"], prm["rcut"], prm["alpha"],
                            prm["ecut"], prm["wm"], prm["wh"], prm["wrcut"], prm["walpha"], prm["rkmax"]))
        f.write(struct.pack("2i", prm["kmax"], prm["iamcub"]))
        f.write(struct.pack("7d", prm["apot"], prm["bpot"], prm["alp"], prm["alpb"], prm["theta"], prm["reoh"],
                            prm["sig"]))
        f.write(struct.pack("3i", prm["nbdf1"], prm["nbdf2"], prm["mts"]))
        f.write(np.asarray(box, np.float64).tobytes())
        f.write(np.asarray(z, np.float64).tobytes())
        f.write(np.ascontiguousarray(r, np.float64).tobytes())


def test_fortran_host_through_iso_c_binding(capi, oracle, tmp_path):
    """The reference-named Fortran entry points (rpmd_hip.f90) called from a Fortran program give
    the same numbers as the C-ABI called from Python and as the oracle."""
    import subprocess
    from rpmd_mainz_amd.build import HARNESS
    if not os.path.exists(HARNESS):
        pytest.skip("Fortran harness not built (amdflang absent at build time)")
    r, box, z, mass, prm = cf.water_system("qtip4pf", (4, 4, 4), 5, sig_A=4.5)
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    for iopt in (0, 5):
        print(f"case fortran harness iopt {iopt}", flush=True)
        fin, fout = tmp_path / f"in{iopt}.bin", tmp_path / f"out{iopt}.bin"
        _write_harness_input(fin, 0, na, nb, iopt, prm, box, z, r)
        subprocess.run([HARNESS, str(fin), str(fout)], check=True, timeout=120)
        out = np.fromfile(fout, dtype=np.float64)
        v_f, vir_f, d_f = out[0], out[1:10], out[10:].reshape(nb, na, 3)
        v_c, d_c, w_c = ctx.forces(r, nb, na, box, z, iopt)
        assert v_f == v_c and np.array_equal(d_f, d_c)
        v_o, d_o, w_o = oracle.forces(P, r, box, z, iopt)
        assert abs(v_f - v_o) < 1e-11 * abs(v_o) and _rel(d_f, d_o) < 1e-11
    ctx.close()


def test_smeared_charges_step_and_tem_error(capi, oracle):
    """Gaussian (smeared) charges through the RPMD step: three NVE evolve_pi steps of the param_smear
    model (M-site, Morse stretch and bend) against the oracle's; the TEM correction with smeared
    charges stops in the reference (src/correct_tem.f90:14-18) and returns RPMD_ERR_SMEARED."""
    r, box, z, mass, prm = cf.water_system("smear", (4, 4, 4), 4)
    na, nb = r.shape[1], r.shape[0]
    P = RpmdParams.from_dict(prm)
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(4711))
    ff = oracle.full_forces(P, r, box, z)
    ctx = capi.RPMDContext(prm, na, nb, "fp64")
    po, ro, do1, do2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    pg, rg, dg1, dg2 = p.copy(), r.copy(), ff["dvdr"].copy(), ff["dvdr2"].copy()
    for k in range(3):
        print(f"case smeared NVE step {k}", flush=True)
        o = oracle.evolve_pi(P, po, ro, do1, do2, dt, mass, box, z, beta)
        po, ro, do1, do2 = o["p"], o["r"], o["dvdr"], o["dvdr2"]
        g = ctx.evolve_pi(pg, rg, dg1, dg2, dt, mass, na, nb, box, z, beta)
        assert abs(g["v"] - o["v"]) < 1e-10 * abs(o["v"])
    assert _rel(rg, ro) < 1e-12
    assert _rel(pg, po) < 1e-10
    ctx.close()
    r, box, z, mass, prm = cf.water_system("smear", (4, 4, 4), 2, sig_A=4.5)
    ctx = capi.RPMDContext(prm, r.shape[1], 2, "fp64")
    with pytest.raises(capi.RPMDError) as e:
        ctx.forces(r, 2, r.shape[1], box, z, 5)
    assert e.value.code == 9
    ctx.close()
