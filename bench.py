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


def run_gpu(prec: str, steps: int, warmup: int, r, p, box, z, mass, prm, beta, na, use_graph: bool):
    ctx = capi.RPMDContext(prm, na, NB, prec)
    dt = DT_FS * cf.TOFS
    ctx.state_upload(p, r, mass, z, na, NB, box, beta, dt, seed=77)
    ctx.use_graph(use_graph)
    ctx.run_steps(warmup)
    t0 = time.perf_counter()
    ctx.run_steps(steps)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    # kernel timing over a separate eager window (HIP events on the library stream)
    ctx.prof_begin()
    nprof = max(2, min(steps, 5))
    ctx.run_steps(nprof)
    prof = ctx.prof_end()
    st = ctx.state_download()
    ctx.close()
    return ms, prof, st


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-beads", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
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
        res[prec] = run_gpu(prec, a.steps, a.warmup, r, p, box, z, mass, prm, beta, na, not a.no_graph)
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
    flops = {"ewald_real": 67.0 * p_ew * NB, "k_sk": 8.0 * nk * na * NB, "k_kforce": 8.0 * nk * na * NB}
    dom = max(("ewald_real", "k_sk", "k_kforce"), key=lambda k: prof[k][0])
    t_ms = prof[dom][0] / max(prof[dom][1], 1)
    achieved = flops[dom] / (t_ms * 1e-3) / 1e12
    roof = dict(bound="mfma", kernel=dom, achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                frac=achieved / PEAK_FP64_TFLOPS, traffic=None,
                note="FP64 vector-FMA bound; MI355X FP64 peak (vector = matrix = 78.6 TFLOP/s)",
                kernel_ms={k: v[0] / max(v[1], 1) for k, v in prof.items()})
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
                   "graph": not a.no_graph},
        "ns_per_day": steps_s * DT_FS * 0.0864,
        "fp64_gpu": ({"bead_force_evals_s": 1e3 / res["fp64"][0] * NB, "ms_per_step": res["fp64"][0],
                      "ns_per_day": 1e3 / res["fp64"][0] * DT_FS * 0.0864} if "fp64" in res else None),
        "mixed_vs_fp64": (res["fp64"][0] / res["mixed"][0]) if ("fp64" in res and "mixed" in res) else None,
        "mixed_force_error": err,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
