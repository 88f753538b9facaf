"""Summaries of rocprofv3 output (the rocpd SQLite results database).

  python -m mpk_amd.rocprof_summary stats <dir>
      per-kernel calls, average duration, VGPRs, LDS and scratch bytes
  python -m mpk_amd.rocprof_summary counters <dir>
      average of every collected counter per kernel
  python -m mpk_amd.rocprof_summary pmc <fetch_dir> <write_dir>
      HBM bytes per dispatch of every kernel from a FETCH_SIZE run and a
      WRITE_SIZE run (both counters are KiB per dispatch), each scaled by
      what the counter reads for the calibration copy kernel k_dcopy
      (mpkx_dcopy), which reads and writes 8*n bytes (its grid size is n,
      rounded up to the 256-thread workgroup).
"""
import glob
import json
import os
import sqlite3
import sys
from collections import defaultdict


def _db(d):
    hits = sorted(glob.glob(os.path.join(d, "**", "*.db"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no rocprofv3 .db under {d}")
    return sqlite3.connect(hits[0])


def short(name):
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0][:100]


def stats(d):
    c = _db(d)
    rows = c.execute("select name, count(*), avg(duration), max(vgpr_count), max(accum_vgpr_count), "
                     "max(lds_size), max(scratch_size), sum(duration) from kernels group by name "
                     "order by sum(duration) desc").fetchall()
    total = sum(r[7] for r in rows)
    return [{"kernel": short(r[0]), "calls": r[1], "avg_us": r[2] / 1e3, "vgpr": r[3], "agpr": r[4],
             "lds": r[5], "scratch": r[6], "pct": 100.0 * r[7] / total} for r in rows]


def _counter(d, name):
    c = _db(d)
    acc = defaultdict(list)
    for kname, grid, val in c.execute("select kernel_name, grid_size, value from counters_collection "
                                      "where counter_name = ?", (name,)):
        acc[kname].append((grid, val))
    return acc


def pmc(fetch_dir, write_dir):
    f = _counter(fetch_dir, "FETCH_SIZE")
    w = _counter(write_dir, "WRITE_SIZE")
    cal = [k for k in f if "k_dcopy" in k]
    if not cal:
        raise RuntimeError("calibration kernel k_dcopy not found")
    cal = cal[0]
    grid = max(g for g, _ in f[cal])
    copy_bytes = 8.0 * grid
    fscale = copy_bytes / (sum(v for _, v in f[cal]) / len(f[cal]) * 1024.0)
    wscale = copy_bytes / (sum(v for _, v in w[cal]) / len(w[cal]) * 1024.0)
    out = {"calibration": {"kernel": short(cal), "bytes_each_way": copy_bytes,
                           "fetch_scale": fscale, "write_scale": wscale}, "kernels": {}}
    for k in f:
        fb = sum(v for _, v in f[k]) / len(f[k]) * 1024.0 * fscale
        wb = sum(v for _, v in w[k]) / len(w[k]) * 1024.0 * wscale if k in w else 0.0
        out["kernels"][short(k)] = {"dispatches": len(f[k]), "fetch_bytes": round(fb),
                                    "write_bytes": round(wb), "hbm_bytes": round(fb + wb)}
    return out


def counters(d):
    """Average value per dispatch of every collected counter, per kernel."""
    c = _db(d)
    out = defaultdict(dict)
    for kname, cname, val in c.execute("select kernel_name, counter_name, avg(value) from counters_collection "
                                       "group by kernel_name, counter_name"):
        out[short(kname)][cname] = val
    return out


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        print(f'{"kernel":70s} {"calls":>6s} {"avg_us":>10s} {"vgpr":>5s} {"lds":>7s} {"scratch":>7s} {"pct":>6s}')
        for r in stats(sys.argv[2]):
            print(f'{r["kernel"][:70]:70s} {r["calls"]:6d} {r["avg_us"]:10.3f} {r["vgpr"]:5d} {r["lds"]:7d} '
                  f'{r["scratch"]:7d} {r["pct"]:6.2f}')
    elif sys.argv[1] == "counters":
        for k, v in counters(sys.argv[2]).items():
            print(k[:70], " ".join(f"{c}={x:.4g}" for c, x in sorted(v.items())))
    elif sys.argv[1] == "pmc":
        print(json.dumps(pmc(sys.argv[2], sys.argv[3]), indent=1))
