"""Summaries of rocprofv3 output for profiles/ (text only).

  python -m rpmd_mainz_amd.profsummary stats <kernel_stats.csv> "<command>"          > kernels.md
  python -m rpmd_mainz_amd.profsummary pmc <fetch.csv> <write.csv> <copy_mib> "<commands>" > traffic.json

stats: the per-kernel table of a --kernel-trace --stats run.
pmc:   FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of two --pmc runs, averaged per kernel and scaled
       by what the same counters read for the plain device copy of copy_mib MiB that the run also
       made (bench.py --copy-probe-mb), so that the copy reads as exactly its own bytes.
"""
import csv
import json
import sys

# bench.py's kernel names -> the kernels that implement them (mixed, then FP64)
KERNEL_KEYS = {"ewald_real": ["k_ewald_vl"], "k_sk": ["k_sk_h16", "k_sk_mfma"],
               "k_kforce": ["k_kforce_h16", "k_kforce"], "k_sk_final": ["k_sk_final"]}


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "")


def stats(path: str, command: str, top: int = 30):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("# rocprofv3 kernel statistics\n")
    print(f"Command: `{command}`\n")
    print("| kernel | calls | total (us) | average (us) | % |")
    print("|---|---|---|---|---|")
    for r in rows[:top]:
        print(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e3:.1f} | "
              f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")


def _per_kernel(path: str, counter: str):
    acc = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = _short(r["Kernel_Name"])
        acc.setdefault(k, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def _copy_dispatch(path: str, counter: str, copy_kib: float) -> float:
    """the counter of the runtime's copy kernel dispatch that moved the calibration buffer: the
    copyBuffer dispatch with the largest grid (the probe's copies are the largest in the run)"""
    best, grid = None, -1
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter or "copyBuffer" not in r["Kernel_Name"]:
            continue
        g = int(r["Grid_Size"])
        if g > grid:
            best, grid = float(r["Counter_Value"]), g
    if best is None:
        raise SystemExit("no copy kernel in " + path)
    return best


def pmc(fetch_csv: str, write_csv: str, copy_mib: float, command: str):
    f = _per_kernel(fetch_csv, "FETCH_SIZE")
    w = _per_kernel(write_csv, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0
    cf, cw = _copy_dispatch(fetch_csv, "FETCH_SIZE", copy_kib), _copy_dispatch(write_csv, "WRITE_SIZE", copy_kib)
    sf, sw = copy_kib / cf, copy_kib / cw
    out = dict(command=command, copy_kernel="__amd_rocclr_copyBuffer", copy_mib=copy_mib, copy_fetch_kib=cf,
               copy_write_kib=cw, fetch_scale=sf, write_scale=sw, kernels={})
    for key, subs in KERNEL_KEYS.items():
        ks = [k for sub in subs for k in f if sub in k and k in w]
        if not ks:
            continue
        k = ks[0]
        fb, wb = f[k] * sf * 1024.0, w[k] * sw * 1024.0
        out["kernels"][key] = dict(kernel=k, fetch_kib_raw=f[k], write_kib_raw=w[k], fetch_bytes=fb,
                                   write_bytes=wb, bytes_per_launch=fb + wb)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2], sys.argv[3])
    elif sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3], float(sys.argv[4]), sys.argv[5])
