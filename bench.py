"""Benchmark: RPMD force-and-propagate steps of 1728 q-SPC/Fw H2O x 32 beads on one MI355X.

Workload (the configuration the headline metric is quoted on): 1728 q-SPC/Fw molecules on a 12^3
cubic lattice with seeded random orientations, rho = 0.997 g/cm^3, nb = 32 beads, 298 K, PILE-L
(tau = 100 fs), dt = 0.25 fs, mts = 1, cubic Ewald ecut = 1e-6 (kmax 15), O-O LJ rcut = 9 A on the
bead-centroid list, harmonic-bend SPC/Fw intramolecular term.  A "step" is one evolve_pi step
(all 32 bead force evaluations plus the normal-mode propagation), state resident in HBM.

Prints ONE JSON line: value = bead-force-evaluations/s of the mixed-precision path (or FP64 when
--prec fp64), with the FP64 GPU figure, ns/day, force error of the mixed path against FP64,
the roofline of the dominant kernel (HIP events on the library's stream) and the CPU baseline
(the reference's own Fortran force path, compiled from source, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import math
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


def h_rp(st, mass, beta, nb, na) -> float:
    """H_RP per atom as src/md_eq.f90:247-278 prints it: (nb v + K + V_springs) / na."""
    p, r = st["p"], st["r"]
    tk = float(np.sum(p ** 2 / (2.0 * mass[None, :, None])))
    wn = nb / beta
    d = r - np.roll(r, 1, axis=0)
    vrp = float(np.sum(0.5 * mass[None, :, None] * wn ** 2 * d ** 2))
    return (nb * st["v"] + tk + vrp) / na


def equilibrate(nsteps: int, r, p, box, z, mass, prm, beta, na):
    """PILE-L steps (mixed precision) from the lattice start, so that an NVE window does not begin
    with the violent relaxation of the randomly oriented lattice."""
    ctx = capi.RPMDContext(prm, na, NB, "mixed")
    ctx.state_upload(p, r, mass, z, na, NB, box, beta, DT_FS * cf.TOFS, seed=5)
    ctx.use_graph(True)
    ctx.run_steps(nsteps)
    st = ctx.state_download()
    ctx.close()
    return st["r"], st["p"]


def nve_drift(prec: str, nsteps: int, every: int, r, p, box, z, mass, prm, beta, na):
    """Slope of H_RP per atom (hartree / ps) over an NVE window run in HBM with hipGraphs."""
    q = dict(prm)
    q["therm"] = 0
    ctx = capi.RPMDContext(q, na, NB, prec)
    dt = DT_FS * cf.TOFS
    ctx.state_upload(p, r, mass, z, na, NB, box, beta, dt, seed=1)
    ctx.use_graph(True)
    ts, hs = [], []
    for i in range(nsteps // every):
        ctx.run_steps(every)
        st = ctx.state_download()
        ts.append((i + 1) * every * DT_FS * 1e-3)
        hs.append(h_rp(st, mass, beta, NB, na))
    ctx.close()
    ts, hs = np.array(ts), np.array(hs)
    slope = float(np.polyfit(ts, hs, 1)[0]) if len(ts) > 1 else float("nan")
    return dict(slope_hartree_per_atom_ps=slope, h0=float(hs[0]), spread=float(hs.max() - hs.min()))


def replicas(nrep: int, steps: int, seed: int = 2024):
    """Config 3: nrep q-SPC/Fw replicas, sizes and conditions drawn from seed, stepped together by
    one replica-batch graph in mixed precision.  Returns bead-force-evals/s over all replicas."""
    rng = np.random.default_rng(seed)
    lattice = {128: ((4, 4, 8), (2, 2, 1)), 256: ((4, 8, 8), (2, 1, 1)), 512: ((8, 8, 8), (1, 1, 1)),
               1024: ((8, 8, 16), (2, 2, 1))}
    ctxs, sum_nb = [], 0
    for i in range(nrep):
        nm = int(rng.choice([128, 256, 512, 1024]))
        nb = int(rng.choice([8, 16, 32, 64]))
        rho, temp = float(rng.uniform(0.95, 1.05)), float(rng.uniform(250.0, 350.0))
        ncell, aspect = lattice[nm]
        r, box, z, mass, prm = cf.water_system("qspcfw", ncell, nb, rho=rho, aspect=aspect, seed=100 + i, ecut=1e-6,
                                               intra_model=cf.INTRA_SPCFW, therm="PRO", ttaufs=100.0)
        na = r.shape[1]
        beta = 1.0 / (cf.KBOLTZ * temp)
        p = np.random.default_rng(100 + i).normal(size=r.shape) * np.sqrt(mass / (beta / nb))[None, :, None]
        c = capi.RPMDContext(prm, na, nb, "mixed")
        c.state_upload(p, r, mass, z, na, nb, box, beta, DT_FS * cf.TOFS, seed=100 + i)
        ctxs.append(c)
        sum_nb += nb
    batch = capi.ReplicaBatch(ctxs)
    batch.run_steps(2)
    t0 = time.perf_counter()
    batch.run_steps(steps)
    dt = time.perf_counter() - t0
    batch.close()
    for c in ctxs:
        c.close()
    return dict(replicas=nrep, beads_total=sum_nb, steps=steps, ms_per_step=dt * 1e3 / steps,
                bead_force_evals_s=steps * sum_nb / dt, dtype="f32-pair/f64-acc",
                workload="config 3: q-SPC/Fw replicas of 128-1024 molecules, 8-64 beads, rho 0.95-1.05, "
                         "250-350 K, own cubic boxes (kmax 9-13), PILE-L, one batch hipGraph per 2 steps")


def cpu_baseline(r, box, z, prm, nbs: int):
    """The reference's own forces(iopt=0) (oracle/_ref, Fortran from source), 1 thread."""
    try:
        from oracle.oracle import LIB_REF, RefPath, RpmdParams
    except Exception:
        return None
    if not os.path.exists(LIB_REF):
        return None
    ref = RefPath()
    P = RpmdParams.from_dict(prm)
    rs = np.ascontiguousarray(r[:nbs])
    calls, t0 = 0, time.perf_counter()
    while True:
        ref.forces(P, rs, box, z, 0)
        calls += 1
        dt = time.perf_counter() - t0
        if dt > 10.0 or calls >= 3:
            break
    return dict(value=calls * nbs / dt, unit="bead-force-evals/s", cores=1, kind="reference",
                sample=f"reference Fortran forces(iopt=0), {calls} calls on {nbs} beads of the 1728-molecule "
                       f"system ({dt:.1f} s, single thread, amdflang -O2 build of the reference sources)")


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
    ap.add_argument("--nve-steps", type=int, default=2000, help="NVE window per precision for H_RP drift (0: skip)")
    ap.add_argument("--replicas", type=int, default=48, help="config-3 replica count (0: skip)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    r, p, box, z, mass, prm, beta, na = workload()
    res = {}
    precs = ["fp64", "mixed"] if a.prec == "both" else [a.prec]
    for prec in precs:
        print(f"[bench] {prec}: warmup {a.warmup}, timing {a.steps} steps", file=sys.stderr, flush=True)
        res[prec] = run_gpu(prec, a.steps, a.warmup, r, p, box, z, mass, prm, beta, na, not a.no_graph, a.streams)
    # force error of the mixed path against FP64 on the same configuration
    err = None
    if "mixed" in res and "fp64" in res:
        rr = res["fp64"][2]["r"]
        c64 = capi.RPMDContext(prm, na, NB, "fp64")
        _, d64, _ = c64.forces(rr, NB, na, box, z, 0)
        c64.set_precision("mixed")
        _, dmx, _ = c64.forces(rr, NB, na, box, z, 0)
        c64.close()
        err = dict(rms_rel=math.sqrt(((dmx - d64) ** 2).sum() / (d64 ** 2).sum()),
                   max_abs=float(np.abs(dmx - d64).max()))
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
    peak = PEAK_FP64_TFLOPS if head == "fp64" else PEAK_FP32_TFLOPS
    unit_note = {
        ("fp64", "ewald_real"): "FP64 vector ALU",
        ("fp64", "k_sk"): "FP64 MFMA (v_mfma_f64_16x16x4f64)",
        ("fp64", "k_kforce"): "FP64 vector FMA (Horner)",
        ("mixed", "ewald_real"): "FP32 vector ALU (packed FP32 not used)",
        ("mixed", "k_sk"): "FP16 MFMA with split FP32 operands (v_mfma_f32_16x16x32_f16), FP32-equivalent flops",
        ("mixed", "k_kforce"): "FP16 MFMA with split FP32 operands (v_mfma_f32_16x16x32_f16), FP32-equivalent flops",
    }[(head, dom)]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_1728x32_{head}.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath)).get("kernels", {}).get(dom)
        if t:
            traffic = t["bytes_per_launch"]
    roof = dict(bound="mfma", kernel=dom, achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak,
                traffic=traffic, flops_per_launch=flops[dom], ms_per_launch=t_ms,
                note=f"{unit_note}; algorithmic flops (not the tile-padded count the kernel issues) over the "
                     f"mean launch time from HIP events on the library stream; peak = MI355X "
                     f"{'FP64 78.6' if head == 'fp64' else 'FP32 157.3'} TFLOP/s",
                kernel_ms={k: v[0] / max(v[1], 1) for k, v in prof.items()})
    drift = None
    if a.nve_steps > 0:
        drift = {}
        print(f"[bench] PILE-L equilibration: {a.nve_steps} steps", file=sys.stderr, flush=True)
        re, pe = equilibrate(a.nve_steps, r, p, box, z, mass, prm, beta, na)
        for prec in precs:
            print(f"[bench] NVE drift window {prec}: {a.nve_steps} steps", file=sys.stderr, flush=True)
            drift[prec] = nve_drift(prec, a.nve_steps, 100, re, pe, box, z, mass, prm, beta, na)
        if "fp64" in drift and "mixed" in drift:
            drift["ratio_mixed_fp64"] = (abs(drift["mixed"]["slope_hartree_per_atom_ps"]) /
                                         max(abs(drift["fp64"]["slope_hartree_per_atom_ps"]), 1e-300))
        drift["window_ps"] = a.nve_steps * DT_FS * 1e-3
        drift["start"] = f"after {a.nve_steps} PILE-L steps from the lattice"
    rep = None
    if a.replicas > 0:
        print(f"[bench] config-3 replicas: {a.replicas}", file=sys.stderr, flush=True)
        rep = replicas(a.replicas, 10)
    cpu = None if a.no_cpu else cpu_baseline(r, box, z, prm, a.cpu_beads)
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
        "mixed_force_error": err,
        "nve_drift": drift,
        "replicas": rep,
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
