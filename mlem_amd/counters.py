"""HBM traffic per kernel launch from rocprofv3 counter passes.

The L2 (TCC) of gfx950 counts the requests it sends to memory by size:
TCC_EA0_RDREQ_32B / _64B / _128B for reads (their sum is TCC_EA0_RDREQ) and
TCC_EA0_WRREQ / TCC_EA0_WRREQ_64B for writes (32- or 64-byte transactions).
The bytes a launch moves between L2 and memory are therefore

    read  = 32 * RDREQ_32B + 64 * RDREQ_64B + 128 * RDREQ_128B
    write = 64 * WRREQ_64B + 32 * (WRREQ - WRREQ_64B)

Collect them in two runs of the same command (each pass holds at most four
TCC counters):

    rocprofv3 --pmc TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum \\
              TCC_EA0_RDREQ_128B_sum -d <dirR> -o run --output-format csv -- <cmd>
    rocprofv3 --pmc TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum -d <dirW> -o run \\
              --output-format csv -- <cmd>

then `python3 -m mlem_amd.counters <dirR> <dirW> [--command TEXT]` prints, per
kernel name, the dispatch count and the mean read / write bytes per dispatch
as JSON (with TEXT, the command that produced the passes, under "command").
"""
from __future__ import annotations

import csv
import json
import sys
from collections import defaultdict
from pathlib import Path

READ = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64, "TCC_EA0_RDREQ_128B_sum": 128}


def counter_files(d: Path) -> list[Path]:
    return sorted(Path(d).rglob("*counter_collection.csv"))


def per_dispatch(paths) -> dict:
    """{(kernel name, dispatch id): {counter: value}} over the CSV files."""
    out: dict = defaultdict(lambda: defaultdict(float))
    for p in paths:
        with open(p, newline="") as fh:
            for row in csv.DictReader(fh):
                key = (row["Kernel_Name"], int(row["Dispatch_Id"]))
                out[key][row["Counter_Name"]] += float(row["Counter_Value"])
    return out


def summarize(read_dir: Path, write_dir: Path) -> dict:
    """Mean counters and bytes per dispatch, per kernel name."""
    acc: dict = defaultdict(lambda: defaultdict(list))
    for d in (read_dir, write_dir):
        for (name, _), cnt in per_dispatch(counter_files(d)).items():
            for c, v in cnt.items():
                acc[name][c].append(v)
    result = {}
    for name, cnt in acc.items():
        mean = {c: sum(v) / len(v) for c, v in cnt.items()}
        k = {"dispatches": max(len(v) for v in cnt.values()), "counters": mean}
        if all(c in mean for c in READ):
            k["read_bytes"] = sum(mean[c] * b for c, b in READ.items())
            if "TCC_EA0_RDREQ_sum" in mean:
                # requests of other sizes would make the sizes not add up
                k["read_requests_unsized"] = mean["TCC_EA0_RDREQ_sum"] - sum(
                    mean[c] for c in READ)
        if "TCC_EA0_WRREQ_sum" in mean and "TCC_EA0_WRREQ_64B_sum" in mean:
            w, w64 = mean["TCC_EA0_WRREQ_sum"], mean["TCC_EA0_WRREQ_64B_sum"]
            k["write_bytes"] = 64 * w64 + 32 * (w - w64)
        if "read_bytes" in k and "write_bytes" in k:
            k["bytes"] = k["read_bytes"] + k["write_bytes"]
        result[name] = k
    return result


def kernel_traffic(summary: dict, prefix: str):
    """The entry of the first kernel whose name contains `prefix` (or None)."""
    for name in sorted(summary.get("kernels", summary)):
        if prefix in name:
            return summary.get("kernels", summary)[name]
    return None


def main(argv) -> int:
    args = list(argv[1:])
    command = None
    if "--command" in args:
        i = args.index("--command")
        command = args[i + 1]
        del args[i:i + 2]
    if len(args) != 2:
        print(__doc__, file=sys.stderr)
        return 2
    out = {"kernels": summarize(Path(args[0]), Path(args[1])),
           "formula": "read = 32*RDREQ_32B + 64*RDREQ_64B + 128*RDREQ_128B; "
                      "write = 64*WRREQ_64B + 32*(WRREQ - WRREQ_64B) (TCC_EA0_*_sum, per dispatch)"}
    if command:
        out["command"] = command
    json.dump(out, sys.stdout, indent=1, sort_keys=True)
    print()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
