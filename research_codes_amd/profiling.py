"""Summaries of rocprofv3 output for the interaction kernels.

Reads the CSV files that `rocprofv3 --kernel-trace --stats --output-format csv`
and `rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE --output-format csv` write, and
reduces them to the figures the benchmark reports:

* per-kernel launch counts and durations, with the mean over the last K
  launches of each kernel (the timed steps of a bench.py run);
* HBM bytes per launch of one kernel from FETCH_SIZE / WRITE_SIZE (KiB per
  dispatch), each scaled by what the same counter reads for a plain
  device-to-device copy of known size (bench.py --copy-calibration runs three
  256 MiB copies first);
* any other counters of one `--pmc` pass per launch of one kernel (e.g. the
  hardware-counted VALU flops beside the algorithmic ones);
* the kernel time of the last K full KDK steps of a run (each step starts
  at a predict launch), by stage: list builds, tree rebuild, functors, glue.

Usage:
  python -m research_codes_amd.profiling trace <kernel_stats.csv> <kernel_trace.csv> [last]
  python -m research_codes_amd.profiling traffic <fetch counter_collection.csv> \
      <write counter_collection.csv> <kernel substring> [last]
  python -m research_codes_amd.profiling counters <counter_collection.csv> <kernel substring> [last]
  python -m research_codes_amd.profiling kdk <kernel_trace.csv> [last]
"""
from __future__ import annotations

import csv
import statistics
import sys
from collections import OrderedDict, defaultdict

COPY_BYTES = 256 * 1024 * 1024     # one calibration copy reads and writes this much
COPY_KERNEL = "copyBuffer"


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def kernel_trace(stats_csv: str, trace_csv: str, last: int = 5) -> list[dict]:
    """Per kernel: calls, mean over all launches and over the last `last` (ms)."""
    durs: dict[str, list[float]] = defaultdict(list)
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            durs[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    rows = []
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            d = durs.get(r["Name"], [])
            rows.append({"kernel": _short(r["Name"]), "calls": int(r["Calls"]),
                         "avg_ms": float(r["AverageNs"]) * 1e-6,
                         "last_avg_ms": statistics.fmean(d[-last:]) if d else float("nan"),
                         "percent": float(r["Percentage"])})
    return rows


def _per_dispatch(path: str) -> "OrderedDict[int, tuple[str, float]]":
    out: "OrderedDict[int, tuple[str, float]]" = OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            k = int(r["Dispatch_Id"])
            name, v = out.get(k, (r["Kernel_Name"], 0.0))
            out[k] = (name, v + float(r["Counter_Value"]))
    return out


def traffic(fetch_csv: str, write_csv: str, kernel: str, last: int = 5) -> dict:
    """HBM bytes per launch of `kernel` (substring of its name), calibrated.

    The calibration copies are the first run of consecutive copy dispatches
    of the program (bench.py issues them before anything else); the scale is
    COPY_BYTES over their mean reading."""
    res = {}
    for label, path in (("fetch", fetch_csv), ("write", write_csv)):
        d = _per_dispatch(path)
        copies = []
        for name, v in d.values():
            if COPY_KERNEL in name and v * 1024 > COPY_BYTES / 4:
                copies.append(v)
            elif copies:
                break
        if not copies:
            raise ValueError(f"{path}: no calibration copy dispatch found")
        scale = COPY_BYTES / (statistics.fmean(copies) * 1024)
        vals = [v for name, v in d.values() if kernel in name][-last:]
        if not vals:
            raise ValueError(f"{path}: no dispatch of {kernel}")
        res[label] = {"kib_per_launch": statistics.fmean(vals), "scale": scale,
                      "bytes_per_launch": statistics.fmean(vals) * 1024 * scale}
    res["bytes_per_launch"] = res["fetch"]["bytes_per_launch"] + res["write"]["bytes_per_launch"]
    return res


def counters(csv_path: str, kernel: str, last: int = 5) -> dict:
    """Mean per launch of every counter in one counter-collection file, over
    the last `last` dispatches of `kernel` (substring of its name)."""
    per: "OrderedDict[int, dict]" = OrderedDict()
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            if kernel not in r["Kernel_Name"]:
                continue
            d = per.setdefault(int(r["Dispatch_Id"]), {})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    if not per:
        raise ValueError(f"{csv_path}: no dispatch of {kernel}")
    rows = list(per.values())[-last:]
    names = sorted({k for r in rows for k in r})
    return {k: statistics.fmean(r.get(k, 0.0) for r in rows) for k in names}


_KDK_STAGES = (
    ("long list (EPJ + SPJ): one walk into rooms (count + fill on overflow)", ("list_kernel<2",)),
    ("symmetry list: candidates in one walk, exact members", ("list_kernel<1", "member_kernel<1")),
    ("gather list: candidates in one walk, exact members", ("list_kernel<0", "member_kernel<0")),
    ("list bookkeeping (walk boxes, offsets, packing the lists)", ("walk_box", "offsets_kernel", "member_copy",
                                                                   "room_copy")),
    ("scans and sorts (rocPRIM)", ("rocprim",)),
    ("tree rebuild (keys, levels, moments, node records, groups)",
     ("split_kernel", "moments_kernel", "obox_kernel", "pack_", "bbox_", "keys_kernel",
      "unsorted_kernel", "root_kernel", "group_", "head_flag", "fill_kernel")),
    ("Morton reorder of the state, unchanged-input tests of the rebuilds",
     ("permute_kernel", "equal_kernel", "flag_zero_kernel")),
    ("density functor", ("density_kernel",)),
    ("gravity functor", ("gravity_kernel",)),
    ("hydro functor", ("hydro_kernel",)),
    ("step glue (EOS, write-backs, predict/correct, softening, reductions, retry-round bookkeeping)",
     ("sphk::", "density_commit_kernel")),
    ("copies and fills", ("copyBuffer", "fillBuffer")),
)


def kdk_stages(trace_csv: str, last: int = 5) -> tuple[float, float, list[tuple[str, float, float]]]:
    """Wall time and kernel time per KDK step over the last `last` steps
    (from the last `last` predict launches to the end of the trace), and per
    stage (calls per step, ms per step); kernels matching no stage are torch's
    own (comparisons, list copies)."""
    with open(trace_csv, newline="") as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if "predict_kernel" in r["Kernel_Name"]]
    if len(starts) < last:
        raise ValueError(f"{len(starts)} KDK steps in the trace, {last} asked for")
    seg = rows[starts[-last]:]
    wall = (int(seg[-1]["End_Timestamp"]) - int(seg[0]["Start_Timestamp"])) / last / 1e6
    acc, cnt = defaultdict(float), defaultdict(int)
    for r in seg:
        name = r["Kernel_Name"]
        stage = next((s for s, keys in _KDK_STAGES if any(k in name for k in keys)),
                     "torch ops (comparisons, list copies)")
        acc[stage] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        cnt[stage] += 1
    out = sorted(((s, cnt[s] / last, acc[s] / last) for s in acc), key=lambda t: -t[2])
    return wall, sum(acc.values()) / last, out


def main(argv: list[str]) -> int:
    if len(argv) >= 3 and argv[0] == "trace":
        last = int(argv[3]) if len(argv) > 3 else 5
        print(f"| kernel | calls | avg all (ms) | avg last {last} (ms) | % time |")
        print("|---|---|---|---|---|")
        for r in kernel_trace(argv[1], argv[2], last):
            print(f"| `{r['kernel']}` | {r['calls']} | {r['avg_ms']:.4f} | {r['last_avg_ms']:.4f} | "
                  f"{r['percent']:.2f} |")
        return 0
    if len(argv) >= 4 and argv[0] == "traffic":
        last = int(argv[4]) if len(argv) > 4 else 5
        t = traffic(argv[1], argv[2], argv[3], last)
        print(f"kernel={argv[3]}")
        for k in ("fetch", "write"):
            print(f"{k}_kib_per_launch={t[k]['kib_per_launch']:.1f}")
            print(f"{k}_scale={t[k]['scale']:.5f}")
            print(f"{k}_bytes_per_launch={t[k]['bytes_per_launch']:.0f}")
        print(f"traffic_bytes_per_launch={t['bytes_per_launch']:.0f}")
        return 0
    if len(argv) >= 3 and argv[0] == "counters":
        last = int(argv[3]) if len(argv) > 3 else 5
        print(f"kernel={argv[2]}")
        for k, v in counters(argv[1], argv[2], last).items():
            print(f"{k}={v:.6g}")
        return 0
    if len(argv) >= 2 and argv[0] == "kdk":
        last = int(argv[2]) if len(argv) > 2 else 5
        wall, kern, stages = kdk_stages(argv[1], last)
        print(f"per KDK step (last {last}): wall {wall:.2f} ms, kernels {kern:.2f} ms")
        print("| stage | launches per step | ms per step |")
        print("|---|---|---|")
        for s_, c, ms in stages:
            print(f"| {s_} | {c:.0f} | {ms:.2f} |")
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
