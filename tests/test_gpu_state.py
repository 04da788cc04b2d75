"""Device-resident state and captured graphs: replays stay bit-identical to eager steps whatever the
order of solo runs, replica-batch runs and host-synchronous force calls on the same context.

Each step exchanges the roles of some buffers (r and rst; in contracted steps dvdr or dvdrl and
tmp), and host-synchronous calls (forces, full_forces, sample_rpmd) share tables and lists with
the state.  These tests interleave them and compare the final state with plain eager stepping.
"""
import numpy as np
import pytest

from rpmd_mainz_amd import configs as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from rpmd_mainz_amd import capi
    return capi


def _system(nb=8, therm="PRO", **kw):
    r, box, z, mass, prm = cf.water_system("qspcfw", (4, 4, 4), nb, ecut=1e-6, intra_model=cf.INTRA_SPCFW,
                                           therm=therm, ttaufs=100.0, **kw)
    na = r.shape[1]
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = np.random.default_rng(21).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    return r, p, box, z, mass, prm, beta, dt, na, nb


def _same(a, b):
    return (np.array_equal(a["r"], b["r"]) and np.array_equal(a["p"], b["p"]) and np.array_equal(a["dvdr"], b["dvdr"])
            and a["v"] == b["v"])


def _eager(capi, n, sysargs, prec="fp64"):
    r, p, box, z, mass, prm, beta, dt, na, nb = sysargs
    c = capi.RPMDContext(prm, na, nb, prec)
    c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=5)
    c.use_graph(False)
    c.run_steps(n)
    st = c.state_download()
    c.close()
    return st


@pytest.mark.parametrize("order", [(2, 1, 2), (3, 4, 1), (1, 2, 3)])
def test_solo_and_batch_graphs_interleaved(capi, order):
    """Solo graph steps, then a replica batch holding the same context, then solo graph steps
    again: the same bits as the same number of eager steps (a graph captured with the buffer roles
    of an even step count is never replayed after an odd one)."""
    sa = _system()
    r, p, box, z, mass, prm, beta, dt, na, nb = sa
    ref = _eager(capi, sum(order), sa)
    c = capi.RPMDContext(prm, na, nb, "fp64")
    c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=5)
    c.use_graph(True)
    batch = capi.ReplicaBatch([c])
    print(f"case order {order}", flush=True)
    c.run_steps(order[0])
    batch.run_steps(order[1])
    c.run_steps(order[2])
    assert _same(c.state_download(), ref)
    batch.close()
    c.close()


def test_host_calls_between_steps_leave_the_state(capi):
    """forces / full_forces / sample_rpmd on other positions (another box as well) between graph
    steps of a device-resident state: the state continues as if they had not been called."""
    sa = _system()
    r, p, box, z, mass, prm, beta, dt, na, nb = sa
    ref = _eager(capi, 6, sa)
    c = capi.RPMDContext(prm, na, nb, "fp64")
    c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=5)
    c.use_graph(True)
    c.run_steps(3)
    r2 = r + np.random.default_rng(3).normal(size=r.shape) * 0.05
    box2 = box * 1.03
    c.full_forces(r2, na, nb, z, box2)
    c.forces(r2[:2], 2, na, box2, z, 0)
    c.sample_rpmd(nb, na, mass * 1.5, beta)
    c.run_steps(3)
    assert _same(c.state_download(), ref)
    c.close()


def test_contracted_steps_in_a_graph(capi):
    """RP-contracted MTS steps (Ewald on 5 modes, LJ on 7 of 16 beads, mts 2) replayed from a graph
    give the same bits as eager steps; the mode maps live on the device, so nothing is uploaded
    per step."""
    sa = _system(nb=16, therm="PRO", nbdf1=5, nbdf2=7, mts=2)
    r, p, box, z, mass, prm, beta, dt, na, nb = sa
    ref = _eager(capi, 5, sa)
    c = capi.RPMDContext(prm, na, nb, "fp64")
    c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=5)
    c.use_graph(True)
    c.run_steps(5)
    assert _same(c.state_download(), ref)
    c.close()


def test_batch_with_a_contracted_replica(capi):
    """A replica batch may hold contracted replicas beside plain ones; both match solo eager runs."""
    s1 = _system(nb=16, nbdf1=5, nbdf2=7, mts=2)
    s2 = _system(nb=8)
    refs = [_eager(capi, 4, s) for s in (s1, s2)]
    ctxs = []
    for s in (s1, s2):
        r, p, box, z, mass, prm, beta, dt, na, nb = s
        c = capi.RPMDContext(prm, na, nb, "fp64")
        c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=5)
        ctxs.append(c)
    batch = capi.ReplicaBatch(ctxs)
    batch.run_steps(4)
    for c, ref in zip(ctxs, refs):
        assert _same(c.state_download(), ref)
    batch.close()
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("case", ["flexible", "contracted", "rigid"])
def test_device_estimators_match_state(capi, case):
    """rpmd_estimators (kinetic and spring energy, H_RP per atom, temperature, virial_ke) computed on
    the device after steps, against the same estimators of the downloaded state (pinned against the
    reference's kinetic_rpmd and virial_ke on the CPU), without touching the state or its graph."""
    from test_oracle import np_estimators
    if case == "rigid":
        r, box, z, mass, prm = cf.water_system("tip4p2005", (4, 4, 4), 8, ecut=1e-6, rigid=True, therm="PRO",
                                               ttaufs=100.0, amp=0.05)
        nb, na = r.shape[:2]
        beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.5 * cf.TOFS
        p = np.zeros_like(r)
    else:
        extra = dict(nbdf1=5, nbdf2=7, mts=2) if case == "contracted" else {}
        r, p, box, z, mass, prm, beta, dt, na, nb = _system(nb=16 if extra else 8, **extra)
    c = capi.RPMDContext(prm, na, nb, "fp64")
    c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=5)
    c.use_graph(True)
    c.run_steps(4)
    e = c.estimators()
    st = c.state_download()
    c.run_steps(2)           # the graph is still valid after the estimators
    c.close()
    ref = np_estimators(st["p"], st["r"], st["dvdr"], st["dvdr2"], mass, beta, st["v"],
                        3 * (na // 3) if case == "rigid" else 0)
    for k, v in ref.items():
        assert abs(e[k] - v) <= 1e-12 * max(abs(v), 1e-300), (k, e[k], v)


def _mixed_replicas(capi, n, seed, beads=(2, 8, 16, 32), fp64_at=None):
    from test_gpu_parity import _replica_ctx, _replica_specs
    specs = _replica_specs(n, seed, sizes=(27, 64, 128, 256), beads=beads)
    return specs, [_replica_ctx(capi, sp, "fp64" if i == fp64_at else "mixed") for i, sp in enumerate(specs)]


@pytest.mark.parametrize("case", ["batched", "with_fp64_replica"])
def test_batched_force_launches_equal_independent_runs(capi, case):
    """Mixed-precision replicas of uneven size (27-256 molecules, 2-32 beads, own cubic boxes and
    temperatures) stepped by the batch, whose force kernels then run as one launch per kernel for
    every replica (and whose stream layout and pair-walk cap the autotuner picks): the same bits as
    each replica stepped alone.  With one FP64 replica the batch steps each replica on its own
    stream instead, with the same result.  Includes an odd number of solo steps of one replica
    between batch calls."""
    fp64_at = 2 if case == "with_fp64_replica" else None
    specs, ctxs = _mixed_replicas(capi, 6, 11, fp64_at=fp64_at)
    solo = []
    for i, sp in enumerate(specs):
        from test_gpu_parity import _replica_ctx
        c = _replica_ctx(capi, sp, "fp64" if i == fp64_at else "mixed")
        c.use_graph(True)
        c.run_steps(9)
        solo.append(c.state_download())
        c.close()
    batch = capi.ReplicaBatch(ctxs)
    print(f"case batched force launches {case}", flush=True)
    batch.run_steps(4)
    ctxs[1].run_steps(1)      # one replica one step ahead: the batch brings it back to even roles
    batch.run_steps(4)
    # replica 1 has had 9 steps; the others 8, then one more each
    for i, c in enumerate(ctxs):
        if i != 1:
            c.run_steps(1)
    for c, ref in zip(ctxs, solo):
        st = c.state_download()
        assert np.array_equal(st["r"], ref["r"]) and np.array_equal(st["p"], ref["p"])
        assert st["v"] == ref["v"]
    batch.close()
    for c in ctxs:
        c.close()


def test_aged_verlet_lists_stay_exact(capi, oracle):
    """Device-resident FP64 PILE-L steps over several lifetimes of the real-space list (rebuilt only
    when a bead site leaves the sphere of its build, bead centroid and spread, by half the skin), from
    collapsed ring polymers that spread out as the run goes: after each stretch the state's gradient
    equals the oracle's full_forces at the state's positions, so no pair within the cut-off was
    missing from an aged list."""
    from oracle.oracle import RpmdParams
    r, box, z, mass, prm = cf.water_system("qspcfw", (6, 6, 6), 8, ecut=1e-6, intra_model=cf.INTRA_SPCFW,
                                           therm="PRO", ttaufs=100.0)
    na, nb = r.shape[1], r.shape[0]
    beta, dt = 1.0 / (cf.KBOLTZ * 298.0), 0.25 * cf.TOFS
    p = np.random.default_rng(31).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    P = RpmdParams.from_dict(prm)
    c = capi.RPMDContext(prm, na, nb, "fp64")
    c.state_upload(p, r, mass, z, na, nb, box, beta, dt, seed=31)
    c.use_graph(True)
    moved = 0.0
    for k in range(5):
        c.run_steps(24)
        st = c.state_download()
        ff = oracle.full_forces(P, st["r"], box, z)
        g, g0 = st["dvdr"] + st["dvdr2"], ff["dvdr"] + ff["dvdr2"]
        err = np.abs(g - g0).max() / np.abs(g0).max()
        moved = max(moved, float(np.abs(st["r"] - r).max()))
        print(f"stretch {k}: gradient max rel {err:.2e}, largest move {moved:.3f} bohr", flush=True)
        assert err < 1e-10
    # some bead moved by more than half the skin (0.7 A = 1.3 bohr): the lists aged and were rebuilt
    assert moved > 1.0
    c.close()


def test_replica_batch_lattices_stay_physical(capi):
    """The replica batch's system sizes (128 and 1024 molecules on body-centred, 256 on face-centred,
    512 on simple cubic lattices in cubic boxes) stepped 160 times at 250-350 K from Maxwell-Boltzmann
    bead momenta: no list overflows (every pair within reach fits each site's list), the kinetic
    temperatures stay near their targets and no ring polymer tears apart (largest bead distance from
    its centroid under 2 bohr; a lattice that starts oxygens 2 angstrom apart heats such runs to
    1.5-2 times the target and stretches hydrogens to 5 bohr within 40 steps)."""
    import bench
    specs = [(128, 64, 1.02, 252.0), (256, 8, 0.97, 347.0), (512, 32, 1.0, 298.0), (1024, 16, 1.04, 274.0)]
    ctxs = []
    for i, (nm, nb, rho, temp) in enumerate(specs):
        ncell, basis = cf.cubic_cells(nm)
        r, box, z, mass, prm = cf.water_system("qspcfw", ncell, nb, rho=rho, basis=basis, seed=100 + i, ecut=1e-6,
                                               intra_model=cf.INTRA_SPCFW, therm="PRO", ttaufs=100.0)
        na = r.shape[1]
        beta = 1.0 / (cf.KBOLTZ * temp)
        p = np.random.default_rng(100 + i).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
        c = capi.RPMDContext(prm, na, nb, "mixed")
        c.state_upload(p, r, mass, z, na, nb, box, beta, bench.DT_FS * cf.TOFS, seed=100 + i)
        ctxs.append(c)
    batch = capi.ReplicaBatch(ctxs)
    for k in range(4):
        batch.run_steps(40)
        print(f"replica lattices: {40 * (k + 1)} steps", flush=True)
    for c, (nm, nb, rho, temp) in zip(ctxs, specs):
        t = c.estimators()["temperature_K"]
        r = c.state_download()["r"]
        spread = np.sqrt(((r - r.mean(0)[None]) ** 2).sum(-1)).max()
        print(f"{nm} x {nb} at {temp} K: T {t:.1f} K, largest spread {spread:.3f} bohr", flush=True)
        assert 0.8 * temp < t < 1.25 * temp
        assert spread < 2.0
    batch.close()
    for c in ctxs:
        c.close()
