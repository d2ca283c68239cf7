"""Summarise a rocprofv3 --kernel-trace run (SQLite output) as a markdown
table of per-kernel device time, or rocprofv3 --pmc runs as per-kernel
counter averages.

    python cfd-with-cuda_amd/prof_summary.py <results.db> "<title>" "<command>"
    python cfd-with-cuda_amd/prof_summary.py --steps <results.db> <after> <nsteps> "<title>" "<command>"
    python cfd-with-cuda_amd/prof_summary.py --pmc <fetch.db> <write.db> <copy_bytes_each_way> <out.json> "<command>"

--steps reports the last <nsteps> time steps after the last dispatch of the
kernel named <after> (a time step ends with fsx::k_maxacc_finish), per step.

For --pmc the two databases hold FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) of `bench.py --probe`; the largest dispatch of the library's
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


def clean(name):
    """Kernel name without its argument list."""
    m = re.match(r"_ZN3fsx(\d+)", name)
    if m:   # fsx::<kernel><int template arguments>, which binutils c++filt may not take (bf16 parameters)
        n = int(m.group(1))
        kern = name[m.end():m.end() + n]
        args = re.match(r"I((?:Li-?\d+E)+)E", name[m.end() + n:])
        targs = re.findall(r"Li(-?\d+)E", args.group(1)) if args else []
        return f"fsx::{kern}" + (f"<{', '.join(targs)}>" if targs else "")
    name = name.replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", name).replace("void ", "")


def main():
    db, title, cmd = sys.argv[1], sys.argv[2], sys.argv[3]
    rows = sqlite3.connect(db).execute("select name, duration from kernels").fetchall()
    agg = collections.defaultdict(list)
    for name, dur in rows:
        agg[clean(name)].append(dur / 1000.0)
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


def steps_main(argv):
    db, after, nsteps, title, cmd = argv[0], argv[1], int(argv[2]), argv[3], argv[4]
    rows = sqlite3.connect(db).execute("select name, start, end from kernels order by start").fetchall()
    rows = [(clean(n), s, e) for n, s, e in rows]
    last = max(i for i, (n, _, _) in enumerate(rows) if after in n)
    rest = rows[last + 1:]
    ends = [i for i, (n, _, _) in enumerate(rest) if n.endswith("k_maxacc_finish")]
    first = ends[-nsteps - 1] + 1 if len(ends) > nsteps else 0
    win = rest[first:ends[-1] + 1]
    agg = collections.defaultdict(list)
    for n, s0, e0 in win:
        agg[n].append((e0 - s0) / 1000.0)
    tot = sum(sum(v) for v in agg.values())
    wall = (win[-1][2] - win[0][1]) / 1e6
    print(f"# {title}\n\nCommand: `{cmd}`\n\nSource: rocprofv3 --kernel-trace --stats (rocpd SQLite kernels view): the "
          f"last {nsteps} time steps after the last `{after}` dispatch, per step. Durations in microseconds.\n")
    print(f"Per step: {tot / 1e3 / nsteps:.3f} ms of kernel time, {wall / nsteps:.3f} ms from the first dispatch to the "
          f"last end (the profiler's own per-dispatch cost included), {len(win) / nsteps:.0f} dispatches.\n")
    print("| kernel | calls per step | ms per step | avg us | % of kernel time |")
    print("|---|---|---|---|---|")
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{name}` | {len(v) / nsteps:.1f} | {sum(v) / 1e3 / nsteps:.3f} | {sum(v) / len(v):.2f} | "
              f"{100 * sum(v) / tot:.1f} |")


def counter_avgs(db, counter):
    """kernel -> (mean counter value per dispatch, dispatches, largest value)"""
    rows = sqlite3.connect(db).execute(
        "select name, avg(counter_value), count(*), max(counter_value) from pmc_events where counter_name = ? "
        "group by name", (counter,)).fetchall()
    return {clean(n): (v, c, m) for n, v, c, m in rows}


def pmc_main(argv):
    import json
    fdb, wdb, copy_bytes, out_json, cmd = argv[0], argv[1], float(argv[2]), argv[3], argv[4]
    fetch, write = counter_avgs(fdb, "FETCH_SIZE"), counter_avgs(wdb, "WRITE_SIZE")
    # the calibration copy is the largest k_copy dispatch (the time step's own
    # vector copies are smaller), so its factor comes from the largest value
    copies = [k for k in fetch if k.endswith("k_copy")]
    cal = max(copies, key=lambda k: fetch[k][2])
    f_fac = copy_bytes / (fetch[cal][2] * 1024.0)
    w_fac = copy_bytes / (write[cal][2] * 1024.0)
    out = {"command": cmd, "calibration_kernel": cal, "copy_bytes_each_way": copy_bytes,
           "fetch_factor": f_fac, "write_factor": w_fac, "kernels": {}}
    print(f"# HBM traffic per launch (rocprofv3 --pmc), 64^3 cavity: one time step and the Z SpMVs\n\nCommand: `{cmd}`\n")
    print(f"Calibration: `{cal}` copies {copy_bytes / 2**20:.0f} MiB each way; FETCH_SIZE reads "
          f"{fetch[cal][2]:.0f} KiB, WRITE_SIZE {write[cal][2]:.0f} KiB; factors {f_fac:.4f} (fetch), "
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
    elif len(sys.argv) > 1 and sys.argv[1] == "--steps":
        steps_main(sys.argv[2:])
    else:
        main()
