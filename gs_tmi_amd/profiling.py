"""Counter summary of the tile kernels from rocprofv3 output.

    python -m gs_tmi_amd.profiling --stats DIR --fetch DIR --write DIR --valu DIR --lds DIR \\
        --calib-mib 512 --json OUT.json --md OUT.md --command-prefix '...'

Each DIR is the ``-d`` output directory of one rocprofv3 run of
``bench.py --copy-calibration-mib N``:

  --stats  ``--kernel-trace --stats``   average kernel durations
  --fetch  ``--pmc FETCH_SIZE``         HBM read, KiB per dispatch
  --write  ``--pmc WRITE_SIZE``         HBM write, KiB per dispatch
  --valu   ``--pmc SQ_INSTS_VALU_FMA_F32 SQ_INSTS_VALU_ADD_F32 SQ_INSTS_VALU_MUL_F32
           SQ_INSTS_VALU_TRANS_F32 SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES GRBM_GUI_ACTIVE``
  --lds    ``--pmc SQ_LDS_BANK_CONFLICT SQ_INSTS_LDS SQ_WAVE_CYCLES SQ_INSTS_SALU``

The shader clock of each kernel is its GRBM_GUI_ACTIVE count in the --valu
run over the XCD count and the dispatch's duration.  HBM bytes are the
counters scaled by what the same counter reads for the
bench's copy of a known size (the ``__amd_rocclr_copyBuffer`` dispatch of
``--calib-mib`` MiB, which reads and writes that many bytes once).  Executed
FP32 FLOPs are 64 lanes x (2 FMA + ADD + MUL + TRANS) wave instructions.

The JSON (read by bench.py) carries the SHA-256 of ``csrc/gstv.hip`` it was
measured on and, per tile kernel, ``fetch_bytes``, ``write_bytes``,
``fp32_flop`` and the average duration.
"""

from __future__ import annotations

import argparse
import collections
import csv
import glob
import hashlib
import json
import os
from pathlib import Path

SOURCE = Path(__file__).resolve().parent / "csrc" / "gstv.hip"
COPY_KERNEL = "__amd_rocclr_copyBuffer"
TILE_KERNELS = {"forward": "k_forward(", "backward": "k_backward<"}
ROWS = ["k_forward(", "k_backward<8>", "k_radix_scatter_gen<false>", "k_radix_scatter<true>",
        "k_radix_hist<true>", "k_radix_hist<false>", "k_prologue_prep", "k_finalize_epilogue", "k_prep(",
        "k_bin_count"]
VALU_PEAK_ISSUE = 1024 * 2.4e9 / 4  # wave instructions per second (1024 SIMDs, 4 cycles per wave64 VALU op)


def _csvs(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))


def _short(name):
    return name.replace("(anonymous namespace)::", "").replace("void ", "")


def kernel_stats(d):
    """{kernel name: (calls, average ns)} from a --stats run."""
    out = {}
    for f in _csvs(d, "kernel_stats.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                out[_short(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]))
    return out


def counters(d, reduce=None):
    """{kernel name: {counter: mean per dispatch}} from a --pmc run (``reduce``
    replaces the mean, e.g. max)."""
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for f in _csvs(d, "counter_collection.csv"):
        per = collections.defaultdict(float)
        names = {}
        with open(f) as fh:
            for r in csv.DictReader(fh):
                per[(r["Dispatch_Id"], r["Counter_Name"])] += float(r["Counter_Value"])
                names[r["Dispatch_Id"]] = _short(r["Kernel_Name"])
        for (disp, cname), v in per.items():
            vals[names[disp]][cname].append(v)
    red = reduce or (lambda v: sum(v) / len(v))
    return {k: {c: red(v) for c, v in cs.items()} for k, cs in vals.items()}


def num_xcc(d):
    """XCDs of the profiled GPU, from the run's agent info (8 on MI355X)."""
    for f in _csvs(d, "agent_info.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Agent_Type") == "GPU" and r.get("Num_Xcc", "") not in ("", "0"):
                    return int(r["Num_Xcc"])
    return 8


def dispatch_clocks(d, min_ns=200_000):
    """{kernel name: shader clock in GHz}: the GRBM_GUI_ACTIVE count of a
    dispatch (summed over the XCDs) over the XCD count and the dispatch's own
    duration, over the kernel's dispatches of at least ``min_ns`` (the count
    includes a start-up that dominates short dispatches)."""
    nx = num_xcc(d)
    per = collections.defaultdict(float)
    dur, names = {}, {}
    for f in _csvs(d, "counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != "GRBM_GUI_ACTIVE":
                    continue
                per[r["Dispatch_Id"]] += float(r["Counter_Value"])
                dur[r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                names[r["Dispatch_Id"]] = _short(r["Kernel_Name"])
    cnt, tim = collections.defaultdict(float), collections.defaultdict(float)
    for k, v in per.items():
        if dur[k] >= min_ns:
            cnt[names[k]] += v / nx
            tim[names[k]] += dur[k]
    return {k: cnt[k] / tim[k] for k in cnt}


def _find(table, key):
    hits = [k for k in table if k.startswith(key) or k.split("(")[0] == key.rstrip("(")]
    return hits[0] if hits else None


def summarise(stats_d, fetch_d, write_d, valu_d, lds_d, calib_mib):
    st, fe, wr, va, ld = (kernel_stats(stats_d), counters(fetch_d), counters(write_d), counters(valu_d),
                          counters(lds_d))
    # The calibration copy is the largest copy dispatch of each run (the
    # library's other copies move a few bytes).
    fe_max, wr_max = counters(fetch_d, max), counters(write_d, max)
    copy_f, copy_w = _find(fe_max, COPY_KERNEL), _find(wr_max, COPY_KERNEL)
    if copy_f is None or copy_w is None:
        raise SystemExit("no calibration copy in the FETCH_SIZE / WRITE_SIZE runs (bench.py --copy-calibration-mib)")
    clk = dispatch_clocks(valu_d)
    copy_clk = _find(clk, COPY_KERNEL)
    nbytes = calib_mib * (1 << 20)
    f_scale = nbytes / (fe_max[copy_f]["FETCH_SIZE"] * 1024.0)  # true bytes per counted byte
    w_scale = nbytes / (wr_max[copy_w]["WRITE_SIZE"] * 1024.0)
    rows = {}
    for key in ROWS:
        name = _find(st, key)
        if name is None:
            continue
        calls, avg_ns = st[name]
        row = {"name": name.split("(")[0], "calls": calls, "avg_us": avg_ns / 1e3}
        for table, cname, out, scale in ((fe, "FETCH_SIZE", "fetch_bytes", f_scale),
                                         (wr, "WRITE_SIZE", "write_bytes", w_scale)):
            k = _find(table, key)
            if k is not None and cname in table[k]:
                row[out] = table[k][cname] * 1024.0 * scale
        k = _find(va, key)
        if k is not None:
            c = va[k]
            row["valu"] = c.get("SQ_INSTS_VALU")
            fp = [c.get(n, 0.0) for n in ("SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_ADD_F32",
                                          "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_TRANS_F32")]
            row["fp32_wave_instr"] = fp
            row["fp32_flop"] = 64.0 * (2 * fp[0] + fp[1] + fp[2] + fp[3])
            row["fp32_share_of_valu"] = sum(fp) / c["SQ_INSTS_VALU"] if c.get("SQ_INSTS_VALU") else None
            row["valu_issue"] = c["SQ_INSTS_VALU"] / (VALU_PEAK_ISSUE * avg_ns * 1e-9) if c.get("SQ_INSTS_VALU") else None
        k = _find(clk, key)
        if k is not None:
            row["clock_ghz"] = clk[k]
        k = _find(ld, key)
        if k is not None:
            row["lds_bank_conflict_cycles"] = ld[k].get("SQ_LDS_BANK_CONFLICT")
            row["lds_instr"] = ld[k].get("SQ_INSTS_LDS")
        rows[key] = row
    kernels = {}
    for tag, key in TILE_KERNELS.items():
        r = next((v for k, v in rows.items() if k.startswith(key)), None)
        if r and "fetch_bytes" in r and "write_bytes" in r and "fp32_flop" in r:
            kernels[tag] = {"fetch_bytes": r["fetch_bytes"], "write_bytes": r["write_bytes"],
                            "fp32_flop": r["fp32_flop"], "avg_us": r["avg_us"],
                            "fp32_share_of_valu": r.get("fp32_share_of_valu"), "valu_issue": r.get("valu_issue"),
                            "lds_bank_conflict_cycles": r.get("lds_bank_conflict_cycles"),
                            "clock_ghz": r.get("clock_ghz")}
    return {"gstv_hip_sha256": hashlib.sha256(SOURCE.read_bytes()).hexdigest(),
            "calibration": {"copy_mib": calib_mib, "fetch_scale": f_scale, "write_scale": w_scale,
                            "copy_clock_ghz": clk[copy_clk] if copy_clk is not None else None},
            "kernels": kernels, "rows": list(rows.values())}


def markdown(s, title, commands):
    def f(x, spec):
        return "–" if x is None else format(x, spec)

    out = [f"# {title}", "", "Commands (repository root as `$R`, one MI355X):", "", "```"]
    out += commands
    out += ["```", "",
            f"Calibration: the {s['calibration']['copy_mib']} MiB copy gives true bytes = FETCH_SIZE KiB x "
            f"{s['calibration']['fetch_scale'] * 1024:.1f} and WRITE_SIZE KiB x "
            f"{s['calibration']['write_scale'] * 1024:.1f} (bytes per KiB counted).",
            f"Kernel source SHA-256: `{s['gstv_hip_sha256']}`.", "",
            "Clock: GRBM_GUI_ACTIVE of each dispatch over the XCD count and the dispatch's duration; the "
            f"calibration copy (memory-bound) reads {f(s['calibration'].get('copy_clock_ghz'), '.2f')} GHz by the "
            "same measure, against a published peak engine clock of 2.4 GHz; dispatches shorter than 200 µs are "
            "left out (–). VALU issue is wave instructions x 4 cycles over 1024 SIMDs x 2.4 GHz x the duration.", "",
            "| kernel | calls | avg µs | HBM read MB | HBM write MB | VALU wave-instr | FP32 share of VALU | "
            "executed FP32 TFLOP/s | VALU issue | clock GHz | LDS bank-conflict cycles |",
            "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in s["rows"]:
        tf = r["fp32_flop"] / (r["avg_us"] * 1e-6) / 1e12 if r.get("fp32_flop") is not None else None
        out.append("| `{}` | {} | {} | {} | {} | {} | {} | {} | {} | {} | {} |".format(
            r["name"], r["calls"], f(r["avg_us"], ".1f"),
            f(r.get("fetch_bytes", None) and r["fetch_bytes"] / 1e6, ".0f"),
            f(r.get("write_bytes", None) and r["write_bytes"] / 1e6, ".0f"),
            f(r.get("valu"), ".3e"),
            f(r.get("fp32_share_of_valu") and 100 * r["fp32_share_of_valu"], ".0f") + ("%" if r.get("fp32_share_of_valu") else ""),
            f(tf, ".1f"), f(r.get("valu_issue") and 100 * r["valu_issue"], ".0f") + ("%" if r.get("valu_issue") else ""),
            f(r.get("clock_ghz"), ".2f"), f(r.get("lds_bank_conflict_cycles"), ".2e")))
    return "\n".join(out) + "\n"


def stats_markdown(stats_d, title, commands, top=25):
    """Every kernel of the --stats run: calls, average and total time, share."""
    rows = []
    for f in _csvs(stats_d, "kernel_stats.csv"):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"# {title}", "", "Command (repository root as `$R`, one MI355X):", "", "```"] + commands + ["```", "",
           "| kernel | calls | avg µs | total ms | share |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name']).split('(')[0][:70]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.1f}% |")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    for name in ("stats", "fetch", "write", "valu", "lds"):
        ap.add_argument(f"--{name}", required=True)
    ap.add_argument("--calib-mib", type=int, required=True)
    ap.add_argument("--json", required=True)
    ap.add_argument("--md", required=True)
    ap.add_argument("--title", default="Hardware counters of the tile and binning kernels")
    ap.add_argument("--command", action="append", default=[], help="a command line to list in the markdown")
    ap.add_argument("--stats-md", default=None, help="also write the kernel-stats table of the --stats run here")
    ap.add_argument("--stats-title", default="Kernel statistics")
    ap.add_argument("--stats-command", action="append", default=[])
    args = ap.parse_args(argv)
    if args.stats_md:
        Path(args.stats_md).write_text(stats_markdown(args.stats, args.stats_title, args.stats_command))
    s = summarise(args.stats, args.fetch, args.write, args.valu, args.lds, args.calib_mib)
    Path(args.json).write_text(json.dumps(s, indent=1) + "\n")
    Path(args.md).write_text(markdown(s, args.title, args.command))
    print(json.dumps(s["kernels"], indent=1))


if __name__ == "__main__":
    main()
