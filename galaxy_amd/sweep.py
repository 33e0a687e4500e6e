"""BH #2D hardware sweep: the driver at each size of the reference's sweep, every
receipt checked by the fail-closed validator, one summary table.

Follows scripts/bench-bh2d-hardware.py (DEFAULT_PARTICLES 512..65,536, :39; the
driver invoked per size with its default workload, :1392-1466; every receipt
validated, :1013-1390).  The provenance half of that script (git and cargo
hygiene, the manifest's signatures) has no counterpart here.

    python -m galaxy_amd.sweep [--sizes 512,1024,...] [--json out.json] [--markdown out.md]
"""
from __future__ import annotations

import argparse
import json
import sys

from . import driver, receipt

DEFAULT_PARTICLES = (512, 1024, 2048, 4096, 8192, 16384, 32768, 65536)


def sweep(sizes=DEFAULT_PARTICLES, log=None):
    rows = []
    for n in sizes:
        args = driver.parse_args(["--particles", str(n), "--require-hardware"])
        rec = driver.run(args)
        summary = receipt.validate_receipt(
            rec, particles=n, preset=args.preset, steps=args.steps, dt_myr=args.dt_myr, seed=args.seed,
            theta=args.theta, softening_kpc=args.softening_kpc, direct_probes=args.direct_probes,
            oracle_limit=args.oracle_limit, benchmark_warmup=args.benchmark_warmup,
            benchmark_repeats=args.benchmark_repeats)
        bench = rec["tree_build_benchmark"]
        serial = bench["bh2c_serial"]
        row = dict(summary, preset=args.preset, steps=args.steps,
                   bh2c_median_seconds=serial.get("median_seconds"),
                   device_tree_buffer_bytes=rec.get("device_tree_buffer_bytes"))
        rows.append(row)
        if log:
            sp = row["bh2c_parallel_vs_serial_speedup"]
            log(f"  n={n}: rebuild median {1e3 * row['parallel_median_seconds']:.4f} ms"
                + (f", BH #2C {1e3 * row['bh2c_median_seconds']:.4f} ms, speedup {sp:.2f}" if sp else "")
                + ", receipt valid")
    return rows


def markdown(rows) -> str:
    out = ["| particles | active cells | leaves | max depth | rebuild median (ms) | BH #2C rebuild median (ms) "
           "| BH #2D vs BH #2C | probe RMS rel. | probe max rel. |",
           "|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        sp = r["bh2c_parallel_vs_serial_speedup"]
        out.append(f"| {r['particles']:,} | {r['active_cell_count']:,} | {r['leaf_count']:,} | {r['max_depth']} | "
                   f"{1e3 * r['parallel_median_seconds']:.4f} | "
                   + (f"{1e3 * r['bh2c_median_seconds']:.4f} | {sp:.2f} | " if sp else "skipped (cap 4,096) | - | ")
                   + f"{r['direct_probe_rms_relative']:.4f} | {r['direct_probe_max_relative']:.4f} |")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description="BH #2D driver sweep with receipt validation")
    ap.add_argument("--sizes", default=",".join(str(n) for n in DEFAULT_PARTICLES))
    ap.add_argument("--json", default=None)
    ap.add_argument("--markdown", default=None)
    a = ap.parse_args(argv)
    sizes = [int(s) for s in a.sizes.split(",") if s]
    rows = sweep(sizes, log=lambda m: print(m, file=sys.stderr, flush=True))
    text = markdown(rows)
    print(text)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    if a.markdown:
        with open(a.markdown, "w") as f:
            f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
