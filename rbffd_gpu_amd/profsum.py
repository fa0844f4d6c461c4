"""Summarise rocprofv3 result databases (one kernel-trace pass and any number of --pmc passes
over the same program) into the markdown tables kept under profiles/.

usage: python -m rbffd_gpu_amd.profsum [--per-dispatch] <prof_dir> <algorithmic_bytes_json> [kernel-substring ...]
<prof_dir> holds one subfolder per pass, each with the .db that `rocprofv3 -d <prof_dir>/<pass>
-o <pass>` wrote; the pass named kt is the kernel trace (durations, registers, LDS), all others
are counter passes. <algorithmic_bytes_json> maps kernel-name substrings to algorithmic bytes per
launch ("{}" for none). Traffic calibration: the plain copy kernel (stream_copy_kernel) in the
same program copies 1 GiB; its FETCH_SIZE / WRITE_SIZE per dispatch give the bytes per counted
KiB, which scale the other kernels' counters into HBM bytes per launch. With --per-dispatch every
dispatch of a kernel gets its own row, named `<kernel> #k` for its k-th dispatch (in launch order),
instead of one row averaged over the kernel's dispatches.
"""
import glob
import json
import os
import sqlite3
import sys
from collections import defaultdict

COPY_BYTES = 1 << 30


def db(d, name):
    g = glob.glob(os.path.join(d, name, "*.db"))
    return sqlite3.connect(g[0]) if g else None


def short(k):
    k = k.replace("void (anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    return k.split("(")[0] if "stream_copy" in k else k[: k.find(">") + 1] if ">" in k else k.split("(")[0]


def want(name, pats):
    return not pats or any(p in name for p in pats) or "stream_copy" in name


def summarize(d, alg, pats, per_dispatch=False):
    """Markdown tables for the passes under d (see the module docstring)."""
    out = []
    dur = {}
    c = db(d, "kt")
    if c:
        out.append("Kernel trace (`--kernel-trace --stats`):")
        out.append("")
        out.append("| kernel | launches | avg us | min us | max us | VGPR | AGPR | LDS B | wg | grid |")
        out.append("|---|---|---|---|---|---|---|---|---|---|")
        if per_dispatch:
            rows, seen = [], defaultdict(int)
            for name, t, v, a, lds, wg, gr in c.execute(
                    "select name, end-start, vgpr_count, accum_vgpr_count, lds_size, workgroup_x, grid_x from kernels "
                    "order by start"):
                seen[name] += 1
                rows.append((f"{short(name)} #{seen[name]}", name, 1, t, t, t, v, a, lds, wg, gr))
            rows.sort(key=lambda r: -r[3])
        else:
            rows = [(short(r[0]),) + tuple(r) for r in c.execute(
                "select name, count(*), avg(end-start), min(end-start), max(end-start), vgpr_count, accum_vgpr_count, "
                "lds_size, workgroup_x, grid_x from kernels group by name order by sum(end-start) desc")]
        for key, name, cnt, avg, mn, mx, v, a, lds, wg, gr in rows:
            if not want(name, pats):
                continue
            dur[key] = avg / 1e3
            out.append(f"| `{key}` | {cnt} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | {v} | {a} | {lds} | {wg} | {gr} |")
    cnt = defaultdict(dict)
    for p in sorted(os.path.basename(x) for x in glob.glob(os.path.join(d, "*")) if os.path.isdir(x) and not x.endswith("kt")):
        c = db(d, p)
        if not c:
            continue
        if per_dispatch:
            seen = defaultdict(int)
            for name, ctr, val in c.execute("select kernel_name, counter_name, sum(value) from counters_collection "
                                            "group by dispatch_id, counter_name order by dispatch_id, counter_name"):
                seen[name, ctr] += 1  # each dispatch reports every counter of the pass once
                cnt[f"{short(name)} #{seen[name, ctr]}"][ctr] = val
            continue
        for name, ctr, val in c.execute("select kernel_name, counter_name, avg(v) from (select kernel_name, counter_name, "
                                        "dispatch_id, sum(value) v from counters_collection group by dispatch_id, counter_name) "
                                        "group by kernel_name, counter_name"):
            cnt[short(name)][ctr] = val
    if cnt:
        ctrs = sorted({k for v in cnt.values() for k in v})
        out.append("")
        out.append("Counters (`--pmc`, one pass per group), average per dispatch, summed over XCDs:")
        out.append("")
        out.append("| kernel | " + " | ".join(ctrs) + " |")
        out.append("|---|" + "---|" * len(ctrs))
        for name, v in cnt.items():
            if not want(name, pats):
                continue
            out.append(f"| `{name}` | " + " | ".join(f"{v.get(k, float('nan')):.4g}" for k in ctrs) + " |")
        cp = cnt.get("stream_copy_kernel", {})
        if "FETCH_SIZE" in cp and "WRITE_SIZE" in cp:
            fr = COPY_BYTES / cp["FETCH_SIZE"]
            fw = COPY_BYTES / cp["WRITE_SIZE"]
            out.append("")
            out.append(f"Calibration: the plain copy kernel moves {COPY_BYTES} B each way per dispatch and reads "
                       f"FETCH_SIZE {cp['FETCH_SIZE']:.6g} KiB, WRITE_SIZE {cp['WRITE_SIZE']:.6g} KiB, so one counted "
                       f"KiB is {fr:.1f} B fetched and {fw:.1f} B written (factors {fr / 1024:.3f} and {fw / 1024:.3f}).")
            out.append("")
            out.append("| kernel | fetched B | written B | traffic B / launch | algorithmic B / launch | traffic / algorithmic | avg us | algorithmic GB/s |")
            out.append("|---|---|---|---|---|---|---|---|")
            for name, v in cnt.items():
                if "FETCH_SIZE" not in v or "stream_copy" in name or not want(name, pats):
                    continue
                rb, wb = v["FETCH_SIZE"] * fr, v["WRITE_SIZE"] * fw
                a = next((b for s, b in alg.items() if s in name), None)
                us = dur.get(name)
                out.append(f"| `{name}` | {rb:.4g} | {wb:.4g} | {rb + wb:.4g} | {a if a else '-'} | "
                           f"{(rb + wb) / a if a else float('nan'):.3f} | {us if us else float('nan'):.1f} | "
                           f"{a / (us * 1e3) if a and us else float('nan'):.0f} |")
    return "\n".join(out)


if __name__ == "__main__":
    args = sys.argv[1:]
    per = args[:1] == ["--per-dispatch"]
    if per:
        args = args[1:]
    print(summarize(args[0], json.loads(args[1]), args[2:], per_dispatch=per))
