"""Benchmark: RPMD force-and-propagate steps of 1728 q-SPC/Fw H2O x 32 beads on one MI355X.

Workload (the configuration the headline metric is quoted on): 1728 q-SPC/Fw molecules on a 12^3
cubic lattice with seeded random orientations, rho = 0.997 g/cm^3, nb = 32 beads, 298 K, PILE-L
(tau = 100 fs), dt = 0.25 fs, mts = 1, cubic Ewald ecut = 1e-6 (kmax 15), O-O LJ rcut = 9 A on the
bead-centroid list, harmonic-bend SPC/Fw intramolecular term.  A "step" is one evolve_pi step
(all 32 bead force evaluations plus the normal-mode propagation), state resident in HBM.

Prints ONE JSON line: value = bead-force-evaluations/s of the mixed-precision path (or FP64 when
--prec fp64), with the FP64 GPU figure, ns/day, force error of the mixed path against FP64,
the roofline of the dominant kernel (HIP events on the library's stream) and the CPU baseline
(the reference's own Fortran force path, compiled from source, timed on this host's cores; its
first evaluation also serves as the reference forces the GPU paths' errors are measured against).
Extra keys: the NVE H_RP drift of both precisions over 10 ps from a PILE-equilibrated state
(device estimators), and throughput lines for the other workloads: 216 x 32 FP64, 48 uneven
replicas, 4096 x 64 contracted MTS in a 1:1:1.2 box, 1024 x 32 rigid TIP4P/2005.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from rpmd_mainz_amd import capi  # noqa: E402
from rpmd_mainz_amd import configs as cf  # noqa: E402

NB = 32
NCELL = (12, 12, 12)
DT_FS = 0.25
TEMP = 298.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def workload(seed: int = 2024):
    r, box, z, mass, prm = cf.water_system("qspcfw", NCELL, NB, rho=0.997, seed=seed, amp=0.0, ecut=1e-6,
                                           intra_model=cf.INTRA_SPCFW, therm="PRO", ttaufs=100.0, mts=1)
    beta = 1.0 / (cf.KBOLTZ * TEMP)
    rng = np.random.default_rng(seed + 7)
    na = r.shape[1]
    p = rng.normal(size=r.shape) * np.sqrt(mass / (beta / NB))[None, :, None]
    p -= mass[None, :, None] * p.sum(axis=(0, 1)) / (NB * mass.sum())
    return r, p, box, z, mass, prm, beta, na


def count_pairs(r0: np.ndarray, box, rc: float) -> int:
    """in-cutoff site pairs of one bead (periodic KD tree)"""
    from scipy.spatial import cKDTree
    pts = np.mod(r0, box)
    t = cKDTree(pts, boxsize=box)
    return int(t.count_neighbors(t, rc) - len(pts)) // 2


def kset_size(prm) -> int:
    """k vectors of the reference's cubic loop (src/ewald_cubic.f90:292-520), kx >= 0"""
    K, rk2max = prm["kmax"], prm["rkmax"] ** 2
    rl = 2.0 * math.pi
    n = 0
    for kx in range(K + 1):
        for ky in range(-K, K + 1):
            for kz in range(-K, K + 1):
                if kx == 0 and ky == 0 and kz == 0:
                    continue
                rk2 = (rl * kx) ** 2 + (rl * ky) ** 2 + (rl * kz) ** 2
                if (ky == 0 and kz == 0) or rk2 <= rk2max:
                    n += 1
    return n


def run_gpu(prec: str, steps: int, warmup: int, r, p, box, z, mass, prm, beta, na, use_graph: bool, streams: int):
    ctx = capi.RPMDContext(prm, na, NB, prec)
    ctx.set_streams(streams)
    dt = DT_FS * cf.TOFS
    ctx.state_upload(p, r, mass, z, na, NB, box, beta, dt, seed=77)
    ctx.use_graph(use_graph)
    ctx.run_steps(warmup)
    t0 = time.perf_counter()
    ctx.run_steps(steps)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    # kernel timing over a separate eager window (HIP events on the library stream; the library runs
    # every kernel on one stream while it times them)
    ctx.prof_begin()
    nprof = max(2, min(steps, 50))
    ctx.run_steps(nprof)
    prof = ctx.prof_end()
    st = ctx.state_download()
    ctx.close()
    return ms, prof, st


def equilibrate(nsteps: int, r, p, box, z, mass, prm, beta, na):
    """PILE-L steps (mixed precision) from the lattice start, so that an NVE window does not begin
    with the relaxation of the randomly oriented lattice."""
    ctx = capi.RPMDContext(prm, na, NB, "mixed")
    ctx.state_upload(p, r, mass, z, na, NB, box, beta, DT_FS * cf.TOFS, seed=5)
    ctx.use_graph(True)
    ctx.run_steps(nsteps)
    st = ctx.state_download()
    ctx.close()
    return st["r"], st["p"]


def nve_drift(prec: str, nsteps: int, every: int, r, p, box, z, mass, prm, beta, na):
    """Slope of H_RP per atom (hartree / ps) over an NVE window run in HBM with hipGraphs; H_RP
    (src/md_eq.f90:255-278) from the device estimators, no state download."""
    q = dict(prm)
    q["therm"] = 0
    ctx = capi.RPMDContext(q, na, NB, prec)
    dt = DT_FS * cf.TOFS
    ctx.state_upload(p, r, mass, z, na, NB, box, beta, dt, seed=1)
    ctx.use_graph(True)
    ts, hs, temps = [], [], []
    for i in range(nsteps // every):
        ctx.run_steps(every)
        e = ctx.estimators()
        ts.append((i + 1) * every * DT_FS * 1e-3)
        hs.append(e["h_rp_per_atom"])
        temps.append(e["temperature_K"])
    ctx.close()
    ts, hs = np.array(ts), np.array(hs)
    slope, se, tau = drift_fit(ts, hs)
    resid = hs - np.polyval(np.polyfit(ts, hs, 1), ts)
    return dict(slope_hartree_per_atom_ps=slope, slope_stderr=se, resid_corr_samples=tau, h0=float(hs[0]),
                spread=float(hs.max() - hs.min()), fluct_rms=float(np.sqrt(np.mean(resid ** 2))),
                mean_temperature_K=float(np.mean(temps)))


def drift_fit(ts, hs):
    """Least-squares slope of H_RP(t) and its standard error, the residuals' autocorrelation taken
    into account: var(slope) = s^2 (2 tau + 1) / sum (t - tbar)^2, tau = the residual
    autocorrelation summed up to its first non-positive value (in samples)."""
    if len(ts) < 3:
        return float("nan"), float("nan"), float("nan")
    c = np.polyfit(ts, hs, 1)
    res = hs - np.polyval(c, ts)
    n = len(res)
    var = float(np.mean(res ** 2))
    tau = 0.0
    for lag in range(1, n // 2):
        rho = float(np.mean(res[:-lag] * res[lag:])) / var if var > 0 else 0.0
        if rho <= 0.0:
            break
        tau += rho
    s2 = float(np.sum(res ** 2)) / (n - 2)
    se = math.sqrt(s2 * (2.0 * tau + 1.0) / float(np.sum((ts - ts.mean()) ** 2)))
    return float(c[0]), se, tau


def step_flops(r0: np.ndarray, box, prm, na: int, nb: int) -> float:
    """algorithmic flops of one force evaluation of nb beads (cubic box): 67 per in-cut-off real-space
    site pair, 8 per (k vector, site) for each of the two k-space contractions, 45 per O-O pair within
    the LJ cut-off, 150 per molecule for the intramolecular terms; pairs counted on bead r0"""
    p_ew = count_pairs(r0, box, prm["wrcut"] * box[0])
    p_lj = count_pairs(r0[0::3], box, prm["rcut"])
    return nb * (67.0 * p_ew + 16.0 * kset_size(prm) * na + 45.0 * p_lj + 150.0 * (na // 3))


def replicas(nrep: int, steps: int, seed: int = 2024):
    """Config 3: nrep q-SPC/Fw replicas, sizes and conditions drawn from seed, stepped together by
    one replica-batch graph in mixed precision.  Returns bead-force-evals/s over all replicas."""
    rng = np.random.default_rng(seed)
    ctxs, sum_nb, flops = [], 0, 0.0
    for i in range(nrep):
        nm = int(rng.choice([128, 256, 512, 1024]))
        nb = int(rng.choice([8, 16, 32, 64]))
        rho, temp = float(rng.uniform(0.95, 1.05)), float(rng.uniform(250.0, 350.0))
        ncell, basis = cf.cubic_cells(nm)
        r, box, z, mass, prm = cf.water_system("qspcfw", ncell, nb, rho=rho, basis=basis, seed=100 + i, ecut=1e-6,
                                               intra_model=cf.INTRA_SPCFW, therm="PRO", ttaufs=100.0)
        na = r.shape[1]
        beta = 1.0 / (cf.KBOLTZ * temp)
        p = np.random.default_rng(100 + i).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
        c = capi.RPMDContext(prm, na, nb, "mixed")
        c.state_upload(p, r, mass, z, na, nb, box, beta, DT_FS * cf.TOFS, seed=100 + i)
        ctxs.append(c)
        sum_nb += nb
        flops += step_flops(r[0], box, prm, na, nb)
    batch = capi.ReplicaBatch(ctxs)
    # untimed: one eager step, the graphs captured, the autotuner's four steps of each form, one
    # replay of the kept graph
    batch.run_steps(11)
    t0 = time.perf_counter()
    batch.run_steps(steps)
    dt = time.perf_counter() - t0
    tuning = batch.tuning()
    batch.close()
    for c in ctxs:
        c.close()
    ms = dt * 1e3 / steps
    achieved = flops / (ms * 1e-3) / 1e12
    roof = dict(bound="valu", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s", frac=achieved / PEAK_FP32_TFLOPS,
                flops_per_step=flops,
                note="whole batch step: the algorithmic flops of every replica's force evaluation (real-space "
                     "pairs, both k-space contractions, O-O pairs, intramolecular terms; the headline roofline's "
                     "counts) over the batch's mean step time, against the FP32 vector peak")
    return dict(replicas=nrep, beads_total=sum_nb, steps=steps, ms_per_step=ms, roofline=roof,
                bead_force_evals_s=steps * sum_nb / dt, dtype="f32-pair/f64-acc", tuning=tuning,
                workload="config 3: q-SPC/Fw replicas of 128-1024 molecules, 8-64 beads, rho 0.95-1.05, "
                         "250-350 K, own cubic boxes (kmax 9-13) filled by bcc/fcc/sc lattices, PILE-L, one batch hipGraph per 2 steps (captured in the warm-up); "
                         "batched force launches or the replicas' own steps, whichever the autotuner timed faster")


def _ref_forces_worker(args):
    """one process of the all-core baseline: the reference's forces(iopt=0) on a slice of beads"""
    prm, rs, box, z = args
    from oracle.oracle import RefPath, RpmdParams
    ref = RefPath()
    t0 = time.perf_counter()
    ref.forces(RpmdParams.from_dict(prm), rs, box, z, 0)
    return time.perf_counter() - t0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(r, box, z, prm, nbs: int, all_cores: bool = True):
    """The reference's own forces(iopt=0) (oracle/_ref, Fortran from source): one thread, and the
    beads spread over the host's cores (the reference's force path is serial; one process per core,
    each with its own copy of the reference's COMMON state).  Runs before the GPU is touched.
    Returns (baseline dict, the first evaluation's (v, dvdr) as the reference forces)."""
    try:
        from oracle.oracle import LIB_REF, RefPath, RpmdParams
    except Exception:
        return None, None
    if not os.path.exists(LIB_REF):
        return None, None
    ref = RefPath()
    P = RpmdParams.from_dict(prm)
    rs = np.ascontiguousarray(r[:nbs])
    calls, t0, first = 0, time.perf_counter(), None
    while True:
        v, d, _ = ref.forces(P, rs, box, z, 0)
        if first is None:
            first = (v, d)
        calls += 1
        dt = time.perf_counter() - t0
        if dt > 10.0 or calls >= 3:
            break
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    out = dict(value=calls * nbs / dt, unit="bead-force-evals/s", cores=1, kind="reference",
               sample=f"reference Fortran forces(iopt=0), {calls} calls on {nbs} beads of the 1728-molecule "
                      f"system ({dt:.1f} s, single thread, amdflang -O2 build of the reference sources)",
               cpu_model=cpu_model(), nproc=ncpu)
    if all_cores and ncpu and ncpu > 1:
        nw = min(ncpu, 16, nbs)
        chunks = [list(range(i, nbs, nw)) for i in range(nw)]
        jobs = [(dict(prm), np.ascontiguousarray(rs[c]), box, z) for c in chunks]
        ctx = mp.get_context("fork")
        t0 = time.perf_counter()
        with ctx.Pool(nw) as pool:
            pool.map(_ref_forces_worker, jobs)
        wall = time.perf_counter() - t0
        out["all_cores"] = dict(value=nbs / wall, cores=nw,
                                sample=f"the same {nbs} beads split over {nw} processes (one per core), "
                                       f"{wall:.1f} s wall including process start")
    return out, first


def force_errors(d, d_ref):
    return dict(rms_rel=math.sqrt(((d - d_ref) ** 2).sum() / (d_ref ** 2).sum()),
                max_abs=float(np.abs(d - d_ref).max()),
                max_rel=float(np.abs(d - d_ref).max() / np.abs(d_ref).max()))


def config_line(name, model, ncell, nb, dt_fs, prec, steps, warmup, seed=11, aspect=(1, 1, 1), gle=False, **kw):
    """ms/step, bead-force-evals/s and ns/day of one workload stepped in HBM with graphs"""
    r, box, z, mass, prm = cf.water_system(model, ncell, nb, aspect=aspect, seed=seed, **kw)
    na = r.shape[1]
    beta = 1.0 / (cf.KBOLTZ * TEMP)
    p = np.random.default_rng(seed).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
    ctx = capi.RPMDContext(prm, na, nb, prec)
    if gle:
        ctx.gle_init(np.loadtxt(os.path.join(ROOT, "tests", "golden", "gle_A_ns4.txt")), na, nb,
                     0.5 * dt_fs * cf.TOFS, beta, seed=seed)
    ctx.state_upload(p, r, mass, z, na, nb, box, beta, dt_fs * cf.TOFS, seed=seed)
    ctx.use_graph(True)
    ctx.run_steps(warmup)
    t0 = time.perf_counter()
    ctx.run_steps(steps)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.close()
    return dict(workload=name, precision=prec, ms_per_step=ms, bead_force_evals_s=nb * 1e3 / ms,
                ns_per_day=1e3 / ms * dt_fs * 0.0864, steps=steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--streams", type=int, default=2, choices=[1, 2, 3],
                    help="2: real-space Ewald on a second stream beside the k-space kernels; 3: also the O-O and "
                         "intramolecular terms on a third; 1: one stream")
    ap.add_argument("--cpu-beads", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nve-steps", type=int, default=40000,
                    help="NVE window per precision for the H_RP drift (40000 steps = 10 ps; 0: skip)")
    ap.add_argument("--equil-steps", type=int, default=8000, help="PILE-L steps before the NVE window")
    ap.add_argument("--replicas", type=int, default=48, help="config-3 replica count (0: skip)")
    ap.add_argument("--no-configs", action="store_true", help="skip the other workloads' lines")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    r, p, box, z, mass, prm, beta, na = workload()
    # the CPU baseline first, before the GPU is touched (its worker processes are plain forks)
    cpu, ref_forces = None, None
    if not a.no_cpu:
        print(f"[bench] CPU baseline: reference forces on {a.cpu_beads} beads", file=sys.stderr, flush=True)
        cpu, ref_forces = cpu_baseline(r, box, z, prm, a.cpu_beads)
    res = {}
    precs = ["fp64", "mixed"] if a.prec == "both" else [a.prec]
    for prec in precs:
        print(f"[bench] {prec}: warmup {a.warmup}, timing {a.steps} steps", file=sys.stderr, flush=True)
        res[prec] = run_gpu(prec, a.steps, a.warmup, r, p, box, z, mass, prm, beta, na, not a.no_graph, a.streams)
    # force errors on the configuration the CPU baseline evaluated: against the reference Fortran's own
    # forces (the intramolecular model selected by the reference's rule there), and mixed against FP64
    err = {}
    ra = np.ascontiguousarray(r[:a.cpu_beads])
    nbe = ra.shape[0]
    if ref_forces is not None:
        for prec in precs:
            c = capi.RPMDContext(dict(prm, intra_model=cf.INTRA_AUTO), na, nbe, prec)
            v_g, d_g, _ = c.forces(ra, nbe, na, box, z, 0)
            c.close()
            err[f"{prec}_vs_reference_fortran"] = dict(force_errors(d_g, ref_forces[1]),
                                                       energy_rel=abs(v_g - ref_forces[0]) / abs(ref_forces[0]))
    if "mixed" in res and "fp64" in res:
        c64 = capi.RPMDContext(prm, na, NB, "fp64")
        rr = res["fp64"][2]["r"]
        _, d64, _ = c64.forces(rr, NB, na, box, z, 0)
        c64.set_precision("mixed")
        _, dmx, _ = c64.forces(rr, NB, na, box, z, 0)
        c64.close()
        err["mixed_vs_fp64_gpu"] = force_errors(dmx, d64)
    head = "mixed" if "mixed" in res else "fp64"
    ms, prof, st = res[head]
    steps_s = 1e3 / ms
    bfe = steps_s * NB
    # roofline of the dominant kernel: algorithmic flops per launch / mean launch time
    nk = kset_size(prm)
    sites = st["r"][0].copy()
    rc_bohr = prm["wrcut"] * box[0]
    p_ew = count_pairs(sites, box, rc_bohr)
    # per launch (all NB beads): 67 flop per in-cut-off real-space site pair; 8 flop per (k, site) for
    # the structure factor and for the force projection (kx >= 0 set of the reference)
    flops = {"ewald_real": 67.0 * p_ew * NB, "k_sk": 8.0 * nk * na * NB, "k_kforce": 8.0 * nk * na * NB}
    dom = max(("ewald_real", "k_sk", "k_kforce"), key=lambda k: prof[k][0])
    t_ms = prof[dom][0] / max(prof[dom][1], 1)
    achieved = flops[dom] / (t_ms * 1e-3) / 1e12
    unit_note = {
        ("fp64", "ewald_real"): ("valu", PEAK_FP64_TFLOPS, "FP64 vector ALU"),
        ("fp64", "k_sk"): ("mfma", PEAK_FP64_TFLOPS, "FP64 MFMA (v_mfma_f64_16x16x4f64)"),
        ("fp64", "k_kforce"): ("valu", PEAK_FP64_TFLOPS, "FP64 vector FMA (Horner)"),
        ("mixed", "ewald_real"): ("valu", PEAK_FP32_TFLOPS, "FP32 vector ALU (157.3 is the packed-FP32 vector peak; "
                                                            "packed FP32 is not used, so 78.6 is the issue limit)"),
        ("mixed", "k_sk"): ("mfma", PEAK_FP32_TFLOPS, "FP16 MFMA with split FP32 operands, FP32-equivalent flops"),
        ("mixed", "k_kforce"): ("mfma", PEAK_FP32_TFLOPS, "FP16 MFMA with split FP32 operands, FP32-equivalent flops"),
    }[(head, dom)]
    bound, peak, note = unit_note
    kernel_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items()}
    traffic, kernel_gbs = None, {}
    tpath = os.path.join(ROOT, "profiles", f"traffic_1728x32_{head}.json")
    if os.path.exists(tpath):
        tk = json.load(open(tpath)).get("kernels", {})
        if dom in tk:
            traffic = tk[dom]["bytes_per_launch"]
        for k, v in tk.items():
            if k in kernel_ms and kernel_ms[k] > 0:
                kernel_gbs[k] = dict(bytes_per_launch=v["bytes_per_launch"], ms_per_launch=kernel_ms[k],
                                     achieved_GBs=v["bytes_per_launch"] / (kernel_ms[k] * 1e-3) / 1e9,
                                     frac_of_hbm_peak=v["bytes_per_launch"] / (kernel_ms[k] * 1e-3) / 1e9 / PEAK_HBM_GBS)
    sflops = step_flops(sites, box, prm, na, NB)
    whole = dict(flops_per_step=sflops, achieved=sflops / (ms * 1e-3) / 1e12, unit="TFLOP/s",
                 frac_of_fp32_peak=sflops / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                 note="the step's algorithmic flops (all force terms, as the replica batch's figure) over the "
                      "measured step time; the k-space contractions run on the FP16 MFMA")
    roof = dict(bound=bound, kernel=dom, achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak,
                whole_step=whole,
                traffic=traffic, flops_per_launch=flops[dom], ms_per_launch=t_ms,
                note=f"{note}; algorithmic flops (not the tile-padded count the kernel issues) over the mean launch "
                     f"time from HIP events on the library stream",
                kernel_ms=kernel_ms, kernel_hbm=kernel_gbs)
    drift = None
    if a.nve_steps > 0:
        drift = {}
        print(f"[bench] PILE-L equilibration: {a.equil_steps} steps", file=sys.stderr, flush=True)
        re_, pe = equilibrate(a.equil_steps, r, p, box, z, mass, prm, beta, na)
        for prec in precs:
            print(f"[bench] NVE drift window {prec}: {a.nve_steps} steps", file=sys.stderr, flush=True)
            drift[prec] = nve_drift(prec, a.nve_steps, 200, re_, pe, box, z, mass, prm, beta, na)
        if "fp64" in drift and "mixed" in drift:
            s64, smx = drift["fp64"]["slope_hartree_per_atom_ps"], drift["mixed"]["slope_hartree_per_atom_ps"]
            e64, emx = drift["fp64"]["slope_stderr"], drift["mixed"]["slope_stderr"]
            drift["ratio_mixed_fp64"] = abs(smx) / max(abs(s64), 1e-300)
            # the slopes are compared in units of their combined standard error: a ratio of two slopes
            # that are both within their errors of zero says nothing
            drift["difference_in_stderr"] = abs(smx - s64) / max(math.hypot(e64, emx), 1e-300)
            drift["mixed_within_2x_fp64_bound"] = abs(smx) <= 2.0 * max(abs(s64) + 2.0 * e64, 2.0 * emx)
        drift["window_ps"] = a.nve_steps * DT_FS * 1e-3
        drift["start"] = f"after {a.equil_steps} PILE-L steps ({a.equil_steps * DT_FS * 1e-3:.1f} ps) from the lattice"
    rep = None
    if a.replicas > 0:
        print(f"[bench] config-3 replicas: {a.replicas}", file=sys.stderr, flush=True)
        rep = replicas(a.replicas, 10)
    others = None
    if not a.no_configs:
        others = {}
        print("[bench] other workloads", file=sys.stderr, flush=True)
        others["config1"] = config_line("216 q-SPC/Fw x 32, PILE-L, dt 0.25 fs, FP64", "qspcfw", (6, 6, 6), 32, 0.25,
                                        "fp64", 200, 3, ecut=1e-6, intra_model=cf.INTRA_SPCFW, therm="PRO",
                                        ttaufs=100.0)
        others["config4"] = config_line("4096 q-SPC/Fw x 64, 1:1:1.2 box, mts 4, Ewald on 7 and LJ on 15 contracted "
                                        "modes, PILE-L, outer dt 1 fs, mixed", "qspcfw", (16, 16, 16), 64, 1.0, "mixed",
                                        30, 3, aspect=(1, 1, 1.2), nbdf1=7, nbdf2=15, mts=4, ecut=1e-6,
                                        intra_model=cf.INTRA_SPCFW, therm="PRO", ttaufs=100.0)
        for th in ("PRO", "GLE"):
            others[f"config5_{th}"] = config_line(
                f"1024 rigid TIP4P/2005 x 32, {th}, RATTLE 1e-10, dt 0.5 fs, mixed", "tip4p2005", (8, 8, 8), 32, 0.5,
                "mixed", 50, 3, basis="bcc", ecut=1e-6, rigid=True, therm=th, ttaufs=100.0, rat_tol=1e-10,
                amp=0.02, gle=(th == "GLE"))
    out = {
        "metric": "bead-force-evals/s (1728 q-SPC/Fw H2O x 32 beads, RPMD step)",
        "value": bfe,
        "unit": "bead-force-evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32-pair/f64-acc",
        "data": "synthetic (seeded 12^3 lattice, random orientations, Maxwell-Boltzmann momenta)",
        "config": {"workload": "1728 q-SPC/Fw x 32 beads, cubic Ewald ecut 1e-6, LJ 9 A, PILE-L 298 K, "
                               "dt 0.25 fs, mts 1", "molecules": 1728, "beads": NB, "precision": head,
                   "graph": not a.no_graph, "streams": a.streams},
        "ns_per_day": steps_s * DT_FS * 0.0864,
        "fp64_gpu": ({"bead_force_evals_s": 1e3 / res["fp64"][0] * NB, "ms_per_step": res["fp64"][0],
                      "ns_per_day": 1e3 / res["fp64"][0] * DT_FS * 0.0864,
                      "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in res["fp64"][1].items()}}
                     if "fp64" in res else None),
        "mixed_vs_fp64": (res["fp64"][0] / res["mixed"][0]) if ("fp64" in res and "mixed" in res) else None,
        "force_error": err,
        "nve_drift": drift,
        "replicas": rep,
        "other_workloads": others,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    if a.copy_probe_mb > 0:
        copy_probe(a.copy_probe_mb)


def copy_probe(mib: int):
    """A plain device-to-device copy of known size (mib MiB read and written, 4 times) for scaling
    the HBM byte counters of a profiler run: hipMemcpy between two device buffers, which the HIP
    runtime performs with its copy kernel."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(4):
        if hip.hipMemcpy(b, a, C.c_size_t(n), C.c_int(3)):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy probe: 4 copies of {mib} MiB", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
