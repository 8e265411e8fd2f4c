"""``python -m subsetix_amd.benchmarks [--pattern RE] [--repeat N] [--warmup N]
[--list] [--json PATH]``: run the suite and print, per case, the mean and
spread of the GPU time and the ns per input interval (__main__.py:71-86)."""

from __future__ import annotations

import argparse
import json
import re
from typing import List

import numpy as np

from . import load_cases, run_case


def main(argv: List[str] | None = None) -> List[dict]:
    p = argparse.ArgumentParser(description="subsetix interval-set benchmark suite (MI355X)",
                                formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    p.add_argument("--pattern", default="", help="regular expression on case names")
    p.add_argument("--repeat", type=int, default=None, help="timed calls per case (default: the case's own)")
    p.add_argument("--warmup", type=int, default=None, help="untimed calls per case (default: the case's own)")
    p.add_argument("--list", action="store_true", help="list the cases and exit")
    p.add_argument("--json", default=None, help="write the results to this JSON file")
    args = p.parse_args(argv)
    cases = load_cases()
    if args.pattern:
        rx = re.compile(args.pattern)
        cases = [c for c in cases if rx.search(c.name)]
    if not cases:
        raise SystemExit(f"no case matches {args.pattern!r}")
    if args.list:
        for c in cases:
            print(f"{c.name:>40s}  |  {c.description}")
        return []
    records = []
    for c in cases:
        print(f"{c.name} ...", end="", flush=True)
        res, meta, reps, warm = run_case(c, repeat=args.repeat, warmup=args.warmup)
        ms = res.gpu_times * 1e3
        rec = {"name": c.name, "description": c.description, "avg_ms": float(ms.mean()), "std_ms": float(ms.std()),
               "repeat": reps, "warmup": warm, "gpu_times": [float(t) for t in res.gpu_times],
               "cpu_times": [float(t) for t in res.cpu_times], "metadata": meta}
        line = f" {rec['avg_ms']:9.4f} ms +- {rec['std_ms']:7.4f} (repeat={reps}, warmup={warm})"
        n_in = meta.get("input_intervals")
        if n_in:
            rec["input_intervals"] = int(n_in)
            rec["ns_per_interval"] = rec["avg_ms"] * 1e6 / int(n_in)
            line += f" | {rec['ns_per_interval']:8.3f} ns/int (input={int(n_in)})"
        print(line, flush=True)
        records.append(rec)
    if args.json:
        with open(args.json, "w", encoding="utf-8") as fh:
            json.dump(records, fh, indent=1)
        print(f"results written to {args.json}")
    return records


if __name__ == "__main__":
    main()
