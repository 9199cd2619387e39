"""Algorithmic bytes of the time step's kernels and a per-kernel roofline
table built from rocprofv3 result databases.

Algorithmic (compulsory) bytes of one launch: every per-particle array
element the kernel reads is counted once and every element it writes once;
neighbour lists, bitmasks and matrices are counted at their stored size; data
gathered from neighbours is counted once, as the per-particle array it comes
from.  The sizes are those of the workload the kernels ran on:

    n     live particles
    L     entries of the large-radius neighbour list (re_l)
    LS    entries of the small-radius list (re_s)
    nnz   off-diagonal PPE entries, nesc of them with an escaped column
    rows  stencil rows of the bucket walk (9 in 3-D), W mask words per row

Per-particle layouts (bytes): P4 packed (x, y, z, type) 32, FP64 vector 24,
mirror point 24, correction matrix 72, FP64 scalar 8, int 4, FP32 scalar 4.

Per-kernel totals of a kernel trace alone: `--trace T.db --summary`.

Usage (prints a Markdown table):

    python -m polymps_amd.roofline --trace T.db --fetch F.db --write W.db \
        --sq S.db --copy-fetch CF.db --copy-write CW.db --copy-bytes B \
        --sizes bench_line.log

T.db comes from `rocprofv3 --kernel-trace`, F.db / W.db from `--pmc
FETCH_SIZE` / `--pmc WRITE_SIZE`, S.db from `--pmc SQ_WAVES SQ_INSTS_VALU`,
CF.db / CW.db from the same counters on `bench.py --copy-calibration B`
(a device copy of B bytes); the counters of each kernel are scaled by the
copy's bytes over its counter reading.  --sizes names a file holding the
JSON line of bench.py, whose "sizes" entry gives n, L, LS, nnz and nesc.
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
from collections import defaultdict
from pathlib import Path

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, GB/s
SIMDS = 256 * 4           # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9          # peak engine clock
VALU_CYCLES = 4           # one wave64 VALU instruction per 4 cycles on a SIMD


def _f(s):
    n, L, LS = s["n"], s["L"], s["LS"]
    nnz, nesc = s["nnz"], s["nesc"]
    rows, W = s.get("rows", 9), s.get("W", 8)
    mask = rows * W * 4 * n
    return n, L, LS, nnz, nesc, mask


# kernel -> (bytes of one launch, what is counted)
def model(sizes):
    n, L, LS, nnz, nesc, mask = _f(sizes)
    nlc = sizes.get("nlc", 0)   # window index escapes
    return {
        "kMxSpmv": (5 * nnz + 2 * nlc + 16 * n,
                    "FP32 value 4 + window code 1 per entry, 2 per window escape; per row count 4, diag 4, "
                    "p 4, q 4"),
        "kCgSpmv": (9 * nnz + 2 * nlc + 28 * n,
                    "FP64 value 8 + window code 1 per entry, 2 per window escape; per row count 4, diag 8, "
                    "p 8, q 8"),
        "kTrueResidual": (9 * nnz + 2 * nlc + 36 * n,
                          "FP64 values and window codes; per row count, diag, x, b in, r out"),
        "kPpeAssemble<true, true, true>": (
            15 * nnz + 4 * nesc + 2 * nlc + 4 * L + 116 * n,
            "writes FP64 + FP32 values, codes and window codes (15 per entry), escapes; reads the large list; "
            "per particle P4, x, mirror, bc, source in (80), count, diag, b, FP32 diag / 1/diag, r out (36)"),
        "kViscosity<true>": (4 * L + 140 * n,
                             "large list; P4 32, packed velocity 32, own velocity 24, mirror 24, count 4 in; "
                             "acceleration 24 out"),
        "kViscosityWalk": (4 * n + 24 * n + 24 * n + 24 * n + 24 * n,
                           "positions, type, velocity, mirror in; acceleration out (bucket walk, no list)"),
        "kPndNeighNPCD<true>": (4 * L + 196 * n,
                                "large list; P4, mirror, NPCD, wall PND / count, PND, normal, count in (128); "
                                "PND x3, count, NPCD, NPCD^2, deviation out (68)"),
        "kCorrectionMatrix<true>": (4 * LS + 136 * n,
                                    "small list; P4, mirror, count, neighbour count in; matrix 72 out"),
        "kPressGradient<true>": (4 * LS + 164 * n,
                                 "small list; P4, mirror, pressure, matrix 72, count in; acceleration out"),
        "kCollisionPC<true>": (4 * LS + 112 * n,
                               "small list; P4, velocity, mirror, type, count in; velocity change out"),
        "kBuildList<true>": (28 * n + 32 * n + 2 * mask + 8 * n,
                             "positions, type in; P4 out, both candidate bitmasks (rows x words x 4) and counts out"),
        "kBuildList<false>": (28 * n + 32 * n + mask + 4 * n,
                              "positions, type in; P4, the large-radius bitmask and its counts out"),
        "kMaskToList<8>": (None, "bitmask + count in, list out (large or small list: per call)"),
        "kMxUpdate": (32 * n, "d, p, r, q, 1/diag in (FP32); d, r, z out"),
        "kMxP": (12 * n, "z, p in; p out (FP32)"),
        "kPackVel": (56 * n, "velocity in, packed velocity out"),
    }


def _key(name: str) -> str:
    """'void mpsgpu::kFoo<true>(mpsgpu::KArgs, ...)' -> 'kFoo<true>'."""
    s = re.sub(r"^void ", "", name).split("(")[0]
    return s[len("mpsgpu::"):] if s.startswith("mpsgpu::") else s


def _working(durs):
    """Dispatches that did work: at least a quarter of the 90th-percentile
    duration (launches gated off on the device exit in microseconds)."""
    if not durs:
        return []
    srt = sorted(durs)
    p90 = srt[min(len(srt) - 1, int(0.9 * len(srt)))]
    return [d for d in durs if d >= 0.25 * p90]


def kernel_times(db):
    c = sqlite3.connect(db)
    per = defaultdict(list)
    for name, d in c.execute("select name, duration from kernels"):
        per[_key(name)].append(d)
    return per


def counter(db, cname):
    """kernel -> list of (duration ns, value) per dispatch."""
    c = sqlite3.connect(db)
    per = defaultdict(list)
    for name, d, v in c.execute("select kernel_name, duration, value from counters_collection where counter_name=?",
                                (cname,)):
        per[_key(name)].append((d, v))
    return per


def counter_mean(rows):
    """Mean counter value over the working dispatches."""
    if not rows:
        return None
    keep = set(_working([d for d, _ in rows]))
    vals = [v for d, v in rows if d in keep] or [v for _, v in rows]
    return sum(vals) / len(vals)


def sizes_from_log(path):
    for line in reversed(Path(path).read_text().splitlines()):
        line = line.strip()
        if line.startswith("{") and '"sizes"' in line:
            s = json.loads(line)["sizes"]
            return {"n": s["particles"], "L": s["list_entries"], "LS": s["list_entries_small"],
                    "nnz": s["nnz"], "nesc": s["nnz_escaped"], "nlc": s.get("window_escapes", 0)}
    raise ValueError(f"no bench JSON line with sizes in {path}")


def table(args):
    sizes = sizes_from_log(args.sizes)
    mdl = model(sizes)
    times = kernel_times(args.trace)
    total = sum(sum(v) for v in times.values())
    fetch = counter(args.fetch, "FETCH_SIZE") if args.fetch else {}
    write = counter(args.write, "WRITE_SIZE") if args.write else {}
    waves = counter(args.sq, "SQ_WAVES") if args.sq else {}
    valu = counter(args.sq, "SQ_INSTS_VALU") if args.sq else {}
    kf = kw = None
    if args.copy_fetch and args.copy_write:
        cf = counter_mean(counter(args.copy_fetch, "FETCH_SIZE").get("__amd_rocclr_copyBuffer", []))
        cw = counter_mean(counter(args.copy_write, "WRITE_SIZE").get("__amd_rocclr_copyBuffer", []))
        kf = args.copy_bytes / (cf * 1024.0)   # bytes per KiB of counter
        kw = args.copy_bytes / (cw * 1024.0)
    out = []
    out.append(f"Sizes: n = {sizes['n']:,}, L = {sizes['L']:,.0f}, LS = {sizes['LS']:,.0f}, "
               f"nnz = {sizes['nnz']:,.0f}, nesc = {sizes['nesc']:,.0f}")
    if kf:
        out.append(f"Counter scale from the {args.copy_bytes / 2**30:.2f} GiB copy: FETCH_SIZE x {kf:.3f}, "
                   f"WRITE_SIZE x {kw:.3f} (bytes per counted byte)")
    out.append("")
    out.append("| kernel | calls | working | mean us | share | alg. MB | alg. TB/s | alg. / 8 TB/s | HBM MB | "
               "HBM / alg. | HBM TB/s | VALU / wave | VALU busy |")
    out.append("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    rows = sorted(times.items(), key=lambda kv: -sum(kv[1]))[: args.top]
    for k, durs in rows:
        wk = _working(durs)
        mean_ns = sum(wk) / len(wk)
        share = 100.0 * sum(durs) / total
        alg = mdl.get(k, (None, ""))[0]
        cells = [k, str(len(durs)), str(len(wk)), f"{mean_ns / 1e3:.1f}", f"{share:.1f}%"]
        if alg:
            tbs = alg / mean_ns / 1e3
            cells += [f"{alg / 1e6:.0f}", f"{tbs:.2f}", f"{tbs * 1e3 / HBM_PEAK_GBS:.1%}"]
        else:
            cells += ["-", "-", "-"]
        fm = counter_mean(fetch.get(k, []))
        wm = counter_mean(write.get(k, []))
        if fm is not None and wm is not None and kf:
            hbm = fm * 1024.0 * kf + wm * 1024.0 * kw
            cells += [f"{hbm / 1e6:.0f}", f"{hbm / alg:.2f}" if alg else "-", f"{hbm / mean_ns / 1e3:.2f}"]
        else:
            cells += ["-", "-", "-"]
        wv, vi = counter_mean(waves.get(k, [])), counter_mean(valu.get(k, []))
        if wv and vi is not None:
            busy = vi * VALU_CYCLES / SIMDS / (mean_ns * 1e-9 * CLOCK_HZ)
            cells += [f"{vi / wv:,.0f}", f"{busy:.0%}"]
        else:
            cells += ["-", "-"]
        out.append("| " + " | ".join(cells) + " |")
    out.append("")
    out.append(f"Total kernel time {total / 1e6:.1f} ms over {sum(len(v) for v in times.values())} dispatches.")
    out.append("")
    out.append("Algorithmic bytes counted per kernel:")
    out.append("")
    for k, durs in rows:
        if k in mdl and mdl[k][0]:
            out.append(f"* `{k}`: {mdl[k][1]}")
    return "\n".join(out)


def summary(db, top=30):
    """Per-kernel totals of a kernel trace: calls, total and mean time, share,
    VGPRs and scratch bytes per work-item (as the code object declares)."""
    c = sqlite3.connect(db)
    per = defaultdict(lambda: [0, 0.0, 0, 0])
    for name, d, vg, sc in c.execute("select name, duration, vgpr_count, scratch_size from kernels"):
        k = per[_key(name)]
        k[0] += 1
        k[1] += d
        k[2] = max(k[2], vg or 0)
        k[3] = max(k[3], sc or 0)
    total = sum(v[1] for v in per.values())
    out = ["| kernel | calls | total ms | mean us | share | VGPRs | scratch B |", "|---|---:|---:|---:|---:|---:|---:|"]
    for k, (n, t, vg, sc) in sorted(per.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| {k} | {n} | {t / 1e6:.3f} | {t / n / 1e3:.1f} | {100 * t / total:.1f}% | {vg} | {sc} |")
    out.append("")
    out.append(f"Total GPU kernel time: {total / 1e6:.1f} ms")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--sq")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--sizes", help="file with bench.py's JSON line")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--summary", action="store_true", help="only the per-kernel totals of --trace")
    args = ap.parse_args(argv)
    if args.summary:
        print(summary(args.trace, args.top))
        return
    if not args.sizes:
        ap.error("--sizes is required for the roofline table")
    print(table(args))


if __name__ == "__main__":
    main()
