"""The rigid workload as stated: 1024 TIP4P/2005 molecules (M-site Ewald), 32 beads, NVT at 298 K
with PILE-L or GLE, dt 0.5 fs, RATTLE in FP64 at tolerance 1e-10, mixed-precision forces, stepped
in HBM with graphs.  The thermostats on the rigid step are this library's addition (the reference's
rigid step knows only the Bussi and local Langevin thermostats), so they are pinned by invariants:

* every constraint holds to the tolerance: |d^2 - d0^2| / d0^2 <= 1e-10 for O-H1, O-H2, H1-H2, and
  the bond velocities u . (v_b - v_a) vanish to the same order;
* the kinetic energy samples the canonical value of the constrained ring polymer:
  <K> = (3 na nb - 3 nm nb - 3) / (2 beta_n), beta_n = beta / nb (3 constraints per molecule and
  bead; the centroid's total momentum is removed), within the statistics of the run.
"""
import math
import os

import numpy as np
import pytest

from oracle.oracle import RpmdParams
from rpmd_mainz_amd import configs as cf

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _residuals(r, v, prm):
    """largest relative squared-length error and largest |u . dv| / (|u| |v|_max) of the constraints"""
    rehh = 2.0 * math.sin(0.5 * prm["theta"]) * prm["reoh"]
    m = r.reshape(r.shape[0], -1, 3, 3)
    w = v.reshape(v.shape[0], -1, 3, 3)
    pos, vel = 0.0, 0.0
    vmax = float(np.abs(w).max())
    for (i, j, d0) in ((0, 1, prm["reoh"]), (0, 2, prm["reoh"]), (1, 2, rehh)):
        u = m[:, :, j] - m[:, :, i]
        pos = max(pos, float(np.abs((u * u).sum(-1) - d0 * d0).max()) / (d0 * d0))
        vel = max(vel, float(np.abs((u * (w[:, :, j] - w[:, :, i])).sum(-1)).max()) / (d0 * vmax))
    return pos, vel


@pytest.mark.parametrize("therm", ["PRO", "GLE"])
def test_config5_rigid_tip4p2005_thermostatted(capi, oracle, therm):
    tol = 1e-10
    r, box, z, mass, prm = cf.water_system("tip4p2005", (8, 8, 8), 32, basis="bcc", ecut=1e-6, rigid=True,
                                           therm=therm, ttaufs=100.0, amp=0.1, rat_tol=tol)
    assert prm["iamcub"] == 1 and abs(box[0] * cf.TOA - 31.318) < 0.01 and prm["alpha"] < 1.0
    nb, na = r.shape[:2]
    nm = na // 3
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.5 * cf.TOFS
    P = RpmdParams.from_dict(prm)
    p = oracle.sample_rpmd(nb, na, mass, beta, oracle.rng(21))
    for k in range(nb):
        p[k] = oracle.rattle_s2(r[k], p[k], mass, dt, np.zeros((na, 3)), P)[0]
    # the start (beads displaced as rigid bodies) holds far more spring energy than the canonical
    # state; PILE-L brings the kinetic energy down to it in about 200 steps.  Every run starts from
    # 600 such steps, then runs 200 more steps of the thermostat under test before sampling.
    eq = capi.RPMDContext(dict(prm, therm=cf.THERM["PRO"]), na, nb, "mixed")
    eq.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=8)
    eq.use_graph(True)
    eq.run_steps(600)
    st0 = eq.state_download()
    eq.close()
    ctx = capi.RPMDContext(prm, na, nb, "mixed")
    if therm == "GLE":
        A = np.loadtxt(os.path.join(GOLD, "gle_A_ns4.txt"))
        ctx.gle_init(A, na, nb, 0.5 * dt, beta, seed=4)
    ctx.state_upload(st0["p"], st0["r"], mass, z, na, nb, box, beta, dt, seed=9)
    ctx.use_graph(True)
    ctx.run_steps(200)
    kin, worst = [], (0.0, 0.0)
    for i in range(20):
        ctx.run_steps(25)
        st = ctx.state_download()
        v = st["p"] / mass[None, :, None]
        res = _residuals(st["r"], v, prm)
        worst = (max(worst[0], res[0]), max(worst[1], res[1]))
        kin.append(float(np.sum(st["p"] ** 2 / (2.0 * mass[None, :, None]))))
        assert np.isfinite(st["v"])
    ctx.close()
    betan = beta / nb
    k_exp = (3 * na * nb - 3 * nm * nb - 3) / (2.0 * betan)
    k_mean = float(np.mean(kin))
    print(f"{therm}: constraint residual pos {worst[0]:.2e} vel {worst[1]:.2e}; "
          f"<K>/K_canonical {k_mean / k_exp:.4f} (+- {np.std(kin) / k_exp:.4f})", flush=True)
    assert worst[0] <= 1.0001 * tol
    assert worst[1] <= 1e-9
    assert abs(k_mean / k_exp - 1.0) < 0.02
