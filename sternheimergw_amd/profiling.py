"""Text summaries of rocprofv3 results (rocpd SQLite databases).

    python -m sternheimergw_amd.profiling trace <run_results.db>
        per-kernel launches, total/average/min/max duration and time share
    python -m sternheimergw_amd.profiling pmc <run_results.db> [copy-bytes]
        per-kernel average counter value (e.g. FETCH_SIZE or WRITE_SIZE, KiB
        per dispatch); with copy-bytes, the largest copy kernel's reading is
        used to scale KiB into bytes (calibration against a copy of known size)
    python -m sternheimergw_amd.profiling happly <trace.db> <fetch.db> <write.db> <copy-bytes> [skip]
        the plane-wave H-apply as one unit (all k_pw_* kernels of one
        application): summed duration per application from the kernel trace
        and calibrated HBM bytes read and written per application, as JSON
    python -m sternheimergw_amd.profiling sq <run_results.db> [more.db ...] [--skip N]
        per-kernel average of every counter of the databases (SQ_* and the
        like, one row per kernel and counter), then, as JSON, the FP64 VALU
        flops of one H-apply: 64 lanes x (ADD + MUL + TRANS + 2 FMA) FP64
        instructions of all k_pw_* launches over the number of applications
        (full wavefronts assumed, so partial execution masks count as full)
"""
from __future__ import annotations

import collections
import sqlite3
import sys


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    name = name.replace("HIP_vector_type<double, 2u>", "double2").replace(
        "HIP_vector_type<float, 2u>", "float2")
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):  # drop the parameter list
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    name = name[:cut]
    return name if len(name) < 90 else name[:87] + "..."


def trace(db: str) -> str:
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    agg = collections.defaultdict(list)
    for name, dur in rows:
        agg[name].append(dur)
    total = sum(sum(v) for v in agg.values())
    out = ["| kernel | launches | total ms | avg us | min us | max us | share |",
           "|---|---|---|---|---|---|---|"]
    for name, durs in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        t = sum(durs)
        out.append(f"| `{_short(name)}` | {len(durs)} | {t / 1e6:.3f} | {t / len(durs) / 1e3:.2f} | "
                   f"{min(durs) / 1e3:.2f} | {max(durs) / 1e3:.2f} | {t / total:.3f} |")
    return "\n".join(out)


def pmc(db: str, copy_bytes: int = 0) -> str:
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value, grid_size from counters_collection"
                       ).fetchall()
    agg = collections.defaultdict(list)
    counter = None
    for name, cname, val, grid in rows:
        agg[name].append(val)
        counter = cname
    scale = None
    if copy_bytes:
        copies = [(max(v), n) for n, v in agg.items() if "copy" in n.lower() or "elementwise" in n.lower()]
        if copies:
            kib, _ = max(copies)
            scale = copy_bytes / (kib * 1024.0) if kib else None
    out = [f"counter: {counter}" + (f"; copy calibration factor {scale:.4f} (bytes per KiB*1024)" if scale else ""),
           "| kernel | dispatches | avg KiB | max KiB | avg bytes (calibrated) |", "|---|---|---|---|---|"]
    for name, vals in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        avg = sum(vals) / len(vals)
        cal = f"{avg * 1024 * scale:.4g}" if scale else "-"
        out.append(f"| `{_short(name)}` | {len(vals)} | {avg:.1f} | {max(vals):.1f} | {cal} |")
    return "\n".join(out)


def _pmc_per_kernel(db: str, copy_bytes: int, skip_last: int = 0):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value from counters_collection "
                       "order by dispatch_id").fetchall()
    agg = collections.defaultdict(list)
    for name, _c, val in rows:
        agg[name].append(val)
    copies = [max(v) for n, v in agg.items() if "copy" in n.lower() or "elementwise" in n.lower()]
    if skip_last:
        agg = {n: (v[:-skip_last] if "k_pw_" in n and len(v) > skip_last else v)
               for n, v in agg.items()}
    scale = copy_bytes / (max(copies) * 1024.0) if copies and max(copies) else None
    return agg, scale


def happly(trace_db: str, fetch_db: str, write_db: str, copy_bytes: int,
           skip_last: int = 0) -> dict:
    """The H-apply as one unit: the summed duration (and calibrated bytes) of
    all k_pw_* kernel launches divided by the number of applications (the
    launches of the first kernel of an application, k_pw_becp_zinv or
    k_pw_zinv).  skip_last leaves out the last applications of each database
    (the residual checks that follow the timed solve), i.e. the last
    skip_last launches of every kernel that runs once per application."""
    def per_app(rows):
        by = collections.defaultdict(list)
        for name, val in rows:
            if "k_pw_" in name:
                by[_short(name)].append(val)
        first = [k for k in by if "k_pw_becp_zinv" in k or "k_pw_zinv" in k]
        napp = max((len(by[k]) for k in first), default=0)
        if skip_last:
            by = {k: (v[:-skip_last] if len(v) >= napp and len(v) > skip_last else v)
                  for k, v in by.items()}
            napp -= skip_last
        return by, max(napp, 1)

    con = sqlite3.connect(trace_db)
    dur, napp = per_app(con.execute("select name, duration from kernels order by start").fetchall())
    out = {"kernels": sorted(dur), "applications": napp,
           "launches": {k: len(v) for k, v in dur.items()},
           "avg_us": {k: round(sum(v) / len(v) / 1e3, 3) for k, v in dur.items()},
           "happly_avg_ms": round(sum(sum(v) for v in dur.values()) / napp / 1e6, 5)}
    for tag, db in (("fetch", fetch_db), ("write", write_db)):
        con = sqlite3.connect(db)
        rows = con.execute("select kernel_name, value from counters_collection "
                           "order by dispatch_id").fetchall()
        copies = [v for n, v in rows if "copy" in n.lower() or "elementwise" in n.lower()]
        scale = copy_bytes / (max(copies) * 1024.0) if copies and max(copies) else 1.0
        by, n = per_app(rows)
        per = {k: sum(v) * 1024.0 * scale / n for k, v in by.items()}
        out[f"{tag}_factor"] = round(scale, 4)
        out[f"{tag}_bytes_per_happly"] = round(sum(per.values()))
        out[f"{tag}_bytes_per_kernel"] = {k: round(v) for k, v in per.items()}
    out["traffic_bytes_per_happly"] = out["fetch_bytes_per_happly"] + out["write_bytes_per_happly"]
    return out


def sq(dbs, skip_last: int = 0):
    """Per-kernel averages of every counter in `dbs` and the H-apply's FP64
    VALU flops (see the module text)."""
    agg = collections.defaultdict(list)
    for db in dbs:
        con = sqlite3.connect(db)
        rows = con.execute("select kernel_name, counter_name, value from counters_collection "
                           "order by dispatch_id").fetchall()
        for name, cname, val in rows:
            agg[(_short(name), cname)].append(val)
    if skip_last:
        agg = {k: (v[:-skip_last] if "k_pw_" in k[0] and len(v) > skip_last else v)
               for k, v in agg.items()}
    lines = ["| kernel | counter | dispatches | avg per dispatch |", "|---|---|---|---|"]
    for (name, cname), vals in sorted(agg.items(), key=lambda kv: (kv[0][0], kv[0][1])):
        lines.append(f"| `{name}` | {cname} | {len(vals)} | {sum(vals) / len(vals):.4g} |")
    weight = {"SQ_INSTS_VALU_ADD_F64": 1, "SQ_INSTS_VALU_MUL_F64": 1,
              "SQ_INSTS_VALU_TRANS_F64": 1, "SQ_INSTS_VALU_FMA_F64": 2}
    # applications: launches of the first kernel of an H-apply
    napp = max([len(v) for (n, c), v in agg.items()
                if ("k_pw_becp_zinv" in n or "k_pw_zinv" in n) and c in weight] or [1])
    per_kernel = collections.defaultdict(float)
    for (name, cname), vals in agg.items():
        if "k_pw_" in name and cname in weight:
            per_kernel[name] += 64.0 * weight[cname] * sum(vals) / napp
    out = {"applications": napp,
           "fp64_flops_per_kernel": {k: round(v) for k, v in per_kernel.items()},
           "fp64_flops_per_happly": round(sum(per_kernel.values()))}
    return "\n".join(lines), out


def main(argv):
    if len(argv) < 3 or argv[1] not in ("trace", "pmc", "happly", "sq"):
        print(__doc__)
        return 2
    if argv[1] == "sq":
        import json
        args = argv[2:]
        skip = 0
        if "--skip" in args:
            i = args.index("--skip")
            skip = int(args[i + 1])
            args = args[:i] + args[i + 2:]
        table, out = sq(args, skip)
        print(table)
        print(json.dumps(out, indent=1))
        return 0
    if argv[1] == "happly":
        import json
        skip = int(argv[6]) if len(argv) > 6 else 0
        print(json.dumps(happly(argv[2], argv[3], argv[4], int(argv[5]), skip), indent=1))
    elif argv[1] == "trace":
        print(trace(argv[2]))
    else:
        print(pmc(argv[2], int(argv[3]) if len(argv) > 3 else 0))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
