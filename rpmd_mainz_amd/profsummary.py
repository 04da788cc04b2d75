"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) as a markdown table:
python -m rpmd_mainz_amd.profsummary <results.db> "<command that produced it>" > summary.md"""
import sqlite3
import sys


def main(db_path: str, command: str, top: int = 30):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    print("# rocprofv3 kernel statistics\n")
    print(f"Command: `{command}`\n")
    print("| kernel | calls | total (us) | average (us) | % |")
    print("|---|---|---|---|---|")
    for name, calls, tot, avg, pct in rows[:top]:
        short = name.split("(")[0]
        print(f"| `{short}` | {calls} | {tot:.1f} | {avg:.2f} | {pct:.2f} |")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
