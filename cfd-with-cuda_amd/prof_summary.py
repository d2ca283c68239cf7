"""Summarise a rocprofv3 --kernel-trace run (SQLite output) as a markdown
table of per-kernel device time, or rocprofv3 --pmc runs as per-kernel
counter averages.

    python cfd-with-cuda_amd/prof_summary.py <results.db> "<title>" "<command>"
    python cfd-with-cuda_amd/prof_summary.py --pmc <fetch.db> <write.db> <copy_bytes_each_way> <out.json> "<command>"

For --pmc the two databases hold FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) of `bench.py --spmv-probe`; the largest dispatch of the library's
plain copy kernel (fsx_bench_copy, 256 MiB each way) gives the factor that
turns each counter into bytes.

Durations in the rocpd kernels view are nanoseconds. "Working calls" are the
dispatches longer than 20 us: the solvers enqueue iterations in chunks and the
iterations after the stop flag is set return at once (a few microseconds),
so the working average is the figure to compare with a kernel's timing.
"""
import collections
import re
import sqlite3
import sys


def main():
    db, title, cmd = sys.argv[1], sys.argv[2], sys.argv[3]
    rows = sqlite3.connect(db).execute("select name, duration from kernels").fetchall()
    agg = collections.defaultdict(list)
    for name, dur in rows:
        name = re.sub(r"\(.*", "", name).replace("void ", "")
        agg[name].append(dur / 1000.0)
    tot = sum(sum(v) for v in agg.values())
    print(f"# {title}\n\nCommand: `{cmd}`\n\nSource: rocprofv3 --kernel-trace --stats (rocpd SQLite kernels view), "
          f"durations in microseconds.\n")
    print("| kernel | calls | total ms | avg us | % of GPU time | working calls (>20 us) | avg us (working) |")
    print("|---|---|---|---|---|---|---|")
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        w = [x for x in v if x > 20.0]
        print(f"| `{name}` | {len(v)} | {sum(v) / 1e3:.2f} | {sum(v) / len(v):.2f} | {100 * sum(v) / tot:.1f} | "
              f"{len(w)} | {(sum(w) / len(w) if w else 0):.2f} |")
    print(f"\nTotal kernel time {tot / 1e3:.1f} ms over {len(rows)} dispatches.")


def counter_avgs(db, counter):
    rows = sqlite3.connect(db).execute(
        "select name, avg(counter_value), count(*) from pmc_events where counter_name = ? group by name",
        (counter,)).fetchall()
    return {re.sub(r"\(.*", "", n).replace("void ", ""): (v, c) for n, v, c in rows}


def pmc_main(argv):
    import json
    fdb, wdb, copy_bytes, out_json, cmd = argv[0], argv[1], float(argv[2]), argv[3], argv[4]
    fetch, write = counter_avgs(fdb, "FETCH_SIZE"), counter_avgs(wdb, "WRITE_SIZE")
    copies = [k for k in fetch if k.endswith("k_copy")]
    cal = max(copies, key=lambda k: fetch[k][0])
    f_fac = copy_bytes / (fetch[cal][0] * 1024.0)
    w_fac = copy_bytes / (write[cal][0] * 1024.0)
    out = {"command": cmd, "calibration_kernel": cal, "copy_bytes_each_way": copy_bytes,
           "fetch_factor": f_fac, "write_factor": w_fac, "kernels": {}}
    print(f"# HBM traffic per launch (rocprofv3 --pmc), 64^3 cavity Z SpMV\n\nCommand: `{cmd}`\n")
    print(f"Calibration: `{cal}` copies {copy_bytes / 2**20:.0f} MiB each way; FETCH_SIZE reads "
          f"{fetch[cal][0]:.0f} KiB, WRITE_SIZE {write[cal][0]:.0f} KiB; factors {f_fac:.4f} (fetch), "
          f"{w_fac:.4f} (write).\n")
    print("| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per launch (scaled) |")
    print("|---|---|---|---|---|")
    for k in sorted(fetch, key=lambda k: -fetch[k][0]):
        if k not in write:
            continue
        b = fetch[k][0] * 1024.0 * f_fac + write[k][0] * 1024.0 * w_fac
        out["kernels"][k] = {"fetch_kib": fetch[k][0], "write_kib": write[k][0], "hbm_bytes_per_launch": b}
        if fetch[k][0] > 1024:
            print(f"| `{k}` | {fetch[k][1]} | {fetch[k][0]:.0f} | {write[k][0]:.0f} | {b:.4g} |")
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--pmc":
        pmc_main(sys.argv[2:])
    else:
        main()
