"""Summarise rocprofv3 result databases (rocpd SQLite) into short text tables.

  python -m diablo3_amd.rocpd_summary trace <run_results.db>
      per-kernel dispatch count, mean/min/max duration, registers, LDS
  python -m diablo3_amd.rocpd_summary counters <run_results.db> [<copy_db> <copy_bytes>]
      per-kernel mean counter value per dispatch; with a copy-kernel database of
      known size, the calibration factor (true bytes / counter bytes) and the
      calibrated bytes per dispatch

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch.
"""
from __future__ import annotations

import sqlite3
import sys


def trace(db: str) -> str:
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), avg(duration), min(duration), max(duration), max(vgpr_count), "
        "max(accum_vgpr_count), max(sgpr_count), max(lds_size), max(scratch_size) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    out = ["| kernel | dispatches | mean us | min us | max us | VGPR | AGPR | SGPR | LDS B | scratch B |",
           "|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{r[0][:90]}` | {r[1]} | {r[2] / 1e3:.2f} | {r[3] / 1e3:.2f} | {r[4] / 1e3:.2f} "
                   f"| {r[5]} | {r[6]} | {r[7]} | {r[8]} | {r[9]} |")
    return "\n".join(out)


def counter_means(db: str):
    con = sqlite3.connect(db)
    return con.execute(
        "select kernel_name, counter_name, count(*), avg(value) from counters_collection "
        "group by kernel_name, counter_name order by kernel_name").fetchall()


def counters(db: str, copy_db: str = None, copy_bytes: float = None) -> str:
    factor = None
    lines = []
    if copy_db:
        for name, cname, n, v in counter_means(copy_db):
            if name.startswith("copy_kernel"):
                # a copy reads and writes copy_bytes each; counters are KiB
                factor = copy_bytes / (v * 1024.0)
                lines.append(f"calibration {cname}: copy of {copy_bytes:.0f} B reads {v:.1f} KiB "
                             f"-> factor {factor:.4f} true bytes per counted byte")
    lines.append("| kernel | counter | dispatches | mean KiB/dispatch | calibrated MB/dispatch |")
    lines.append("|---|---|---|---|---|")
    for name, cname, n, v in counter_means(db):
        if name.startswith("__amd"):
            continue
        cal = f"{v * 1024.0 * factor / 1e6:.1f}" if factor else "-"
        lines.append(f"| `{name[:90]}` | {cname} | {n} | {v:.1f} | {cal} |")
    return "\n".join(lines)


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace(argv[1]))
    elif len(argv) >= 2 and argv[0] == "counters":
        print(counters(argv[1], argv[2] if len(argv) > 2 else None,
                       float(argv[3]) if len(argv) > 3 else None))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
