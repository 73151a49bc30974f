"""Per-kernel summaries of rocprofv3 CSV output for this library's kernels.

    python -m gammcor_amd.profsum stats <dir>            # kernel_stats.csv / kernel_trace.csv
    python -m gammcor_amd.profsum pmc <dir> [--calib-first 3 --calib-mib 1024]
                                                        # counter_collection.csv, KiB per dispatch
    python -m gammcor_amd.profsum traffic <fetch dir> <write dir>
                                                        # calibrated bytes/launch per timing name
    python -m gammcor_amd.profsum copy [--mib M] [--reps R]
                                                        # device copy of known size (calibration run)
    python -m gammcor_amd.profsum mfma <dir>             # --pmc MfmaUtil MfmaFlopsF64 MfmaFlopsF32:
                                                        # per kernel MFMA flops, TFLOP/s and utilisation

`<dir>` is the output directory of a `rocprofv3 -f csv -d <dir> -o run ...`
run (the CSV files are found below it).  For `pmc`, a run that starts with
plain device copies of known size (`bench.py --copy-calibration-mib 1024`)
calibrates the counter: the known bytes divided by the copy kernel's reading
give the factor that turns the counter into bytes for the other kernels.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os
import sys


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise SystemExit(f"no {pattern} under {d}")
    return hits[0]


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError(f"none of {names} in {list(row)}")


def kernel_stats(d):
    """[(name, calls, total_ms, avg_us, min_us, max_us)] sorted by total time."""
    try:
        path = _find(d, "*kernel_stats.csv")
        out = []
        with open(path) as f:
            for r in csv.DictReader(f):
                out.append((_col(r, "Name", "KernelName"), int(_col(r, "Calls")),
                            float(_col(r, "TotalDurationNs")) * 1e-6, float(_col(r, "AverageNs")) * 1e-3,
                            float(_col(r, "MinNs")) * 1e-3, float(_col(r, "MaxNs")) * 1e-3))
        return path, sorted(out, key=lambda x: -x[2])
    except SystemExit:
        path = _find(d, "*kernel_trace.csv")
        agg = collections.defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                dur = (int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) * 1e-3
                agg[_col(r, "Kernel_Name")].append(dur)
        out = [(k, len(v), sum(v) * 1e-3, sum(v) / len(v), min(v), max(v)) for k, v in agg.items()]
        return path, sorted(out, key=lambda x: -x[2])


def pmc_per_kernel(d, calib_first=0):
    """{counter: {kernel: [values per dispatch]}} from counter_collection.csv,
    and {counter: [values]} of the first `calib_first` dispatches of a copy
    kernel (the calibration copies run before anything else of that name)."""
    path = _find(d, "*counter_collection.csv")
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(_col(r, "Dispatch_Id")), _col(r, "Counter_Name"), _col(r, "Kernel_Name"),
                         float(_col(r, "Counter_Value"))))
    rows.sort(key=lambda x: x[0])
    res = collections.defaultdict(lambda: collections.defaultdict(list))
    cal = collections.defaultdict(list)
    for _, ctr, k, v in rows:
        if len(cal[ctr]) < calib_first and "copy" in k.lower():
            cal[ctr].append(v)
            continue
        res[ctr][k].append(v)
    return path, res, cal


# rocprof kernel name -> the library's timing name (gmc_prof_get)
_KEYS = (("sytrd_panel_kernel<float", "sytrd_panel_f32"), ("sytrd_panel_kernel<double", "sytrd_panel_f64"),
         ("gemm_kernel<double", "gemm_f64"), ("gemm_kernel<float", "gemm_f32"),
         ("ab_main_kernel", "ab_cas_main"))
_TENSILE_OPS = {"Ailk_Bljk": "nn", "Ailk_Bjlk": "nt", "Alik_Bljk": "tn", "Alik_Bjlk": "tt"}


def timing_key(kernel):
    """The gmc_prof name of a rocprof kernel name, or None."""
    for sub, key in _KEYS:
        if sub in kernel:
            return key
    if kernel.startswith("Cijk_"):
        prec = "f64" if "_DB_" in kernel else "f32" if "_SB_" in kernel else None
        for sub, ab in _TENSILE_OPS.items():
            if sub in kernel and prec:
                return f"blas_gemm_{prec}_{ab}"
    return None


def copy_run(mib, reps):
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy: {reps} x {mib} MiB read + {mib} MiB written (float32 copy_)", flush=True)


def _short(name, width=90):
    return name if len(name) <= width else name[: width - 3] + "..."


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--top", type=int, default=25)
    b = sub.add_parser("pmc")
    b.add_argument("dir")
    b.add_argument("--calib-first", type=int, default=0,
                   help="the first N copy-kernel dispatches are calibration copies of --calib-mib MiB")
    b.add_argument("--calib-mib", type=int, default=1024)
    b.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--calib-first", type=int, default=3)
    t.add_argument("--calib-mib", type=int, default=1024)
    m = sub.add_parser("mfma")
    m.add_argument("dir")
    m.add_argument("--top", type=int, default=25)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)

    if args.cmd == "copy":
        copy_run(args.mib, args.reps)
        return
    if args.cmd == "mfma":
        # per dispatch: MfmaUtil (% of the SIMDs' cycles with the MFMA unit
        # busy), MfmaFlopsF64 / MfmaFlopsF32 (flops issued), and the
        # dispatch's start/end stamps; aggregated per kernel name
        path = _find(args.dir, "*counter_collection.csv")
        disp = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                d = disp.setdefault(int(_col(row, "Dispatch_Id")), {"k": _col(row, "Kernel_Name")})
                d[_col(row, "Counter_Name")] = float(_col(row, "Counter_Value"))
                d["t"] = (int(_col(row, "End_Timestamp")) - int(_col(row, "Start_Timestamp"))) * 1e-9
        per = collections.defaultdict(lambda: [0, 0.0, 0.0, 0.0, 0.0])   # n, seconds, f64, f32, util*seconds
        for d in disp.values():
            a = per[d["k"]]
            a[0] += 1
            a[1] += d["t"]
            a[2] += d.get("MfmaFlopsF64", 0.0)
            a[3] += d.get("MfmaFlopsF32", 0.0)
            a[4] += d.get("MfmaUtil", 0.0) * d["t"]
        rows = sorted(((k, *v) for k, v in per.items() if v[2] + v[3] > 0), key=lambda r: -r[2])
        print(f"# source: {os.path.relpath(path, args.dir)}; kernels that issued FP64 or FP32 MFMA")
        print(f"{'calls':>6} {'ms':>9} {'f64_TF':>8} {'f32_TF':>8} {'TF/s':>8} {'util%':>6}  kernel")
        for k, n, t, f64, f32, ut in rows[: args.top]:
            print(f"{n:6d} {1e3 * t:9.1f} {f64 * 1e-12:8.2f} {f32 * 1e-12:8.2f} {(f64 + f32) * 1e-12 / t:8.1f} "
                  f"{ut / t:6.1f}  {_short(k)}")
        return
    if args.cmd == "traffic":
        # calibrated HBM bytes per launch, FETCH_SIZE + WRITE_SIZE, per timing name
        per = {}
        for ctr_dir in (args.fetch_dir, args.write_dir):
            _, res, cal = pmc_per_kernel(ctr_dir, args.calib_first)
            for ctr, ks in res.items():
                vals = cal.get(ctr, [])
                if not vals or min(vals) <= 0:
                    raise SystemExit(f"no calibration copies for {ctr} in {ctr_dir}")
                f = args.calib_mib * 1024.0 / (sum(vals) / len(vals))
                print(f"# {ctr}: calibration factor {f:.4f} ({len(vals)} copies of {args.calib_mib} MiB)")
                for k, v in ks.items():
                    key = timing_key(k)
                    if key:
                        d = per.setdefault(key, {})
                        d.setdefault(ctr, []).extend(x * f * 1024.0 for x in v)
        print("# TRAFFIC <timing name> <fetch bytes/launch> <write bytes/launch> <launches>")
        for key in sorted(per):
            d = per[key]
            fe = d.get("FETCH_SIZE", [0.0])
            wr = d.get("WRITE_SIZE", [0.0])
            print(f"TRAFFIC {key} {sum(fe) / len(fe):.6e} {sum(wr) / len(wr):.6e} {len(fe)}")
        return
    if args.cmd == "stats":
        path, rows = kernel_stats(args.dir)
        tot = sum(r[2] for r in rows)
        print(f"# source: {os.path.relpath(path, args.dir)}; total kernel time {tot:.1f} ms")
        print(f"{'total_ms':>10} {'share':>6} {'calls':>6} {'avg_us':>10} {'min_us':>10} {'max_us':>10}  kernel")
        for name, n, t, avg, mn, mx in rows[: args.top]:
            print(f"{t:10.1f} {100 * t / tot:5.1f}% {n:6d} {avg:10.1f} {mn:10.1f} {mx:10.1f}  {_short(name)}")
        return
    path, res, cal = pmc_per_kernel(args.dir, args.calib_first)
    factor = {}
    for ctr, vals in cal.items():
        if vals and min(vals) > 0:
            factor[ctr] = args.calib_mib * 1024.0 / (sum(vals) / len(vals))   # true KiB per counted KiB
            print(f"# calibration {ctr}: {len(vals)} copies of {args.calib_mib} MiB counted {sum(vals) / len(vals):.1f} KiB")
    for ctr, ks in res.items():
        f = factor.get(ctr)
        print(f"# {ctr} (KiB per dispatch{'' if f is None else f', calibration factor {f:.4f} from a {args.calib_mib} MiB copy'})")
        rows = sorted(((k, sum(v) / len(v), len(v)) for k, v in ks.items()), key=lambda x: -x[1] * x[2])
        for k, mean, n in rows[: args.top]:
            extra = "" if f is None else f"  calibrated {mean * f / 1024.0:12.1f} MiB"
            print(f"{mean:14.1f} KiB x {n:5d}{extra}  {_short(k)}")


if __name__ == "__main__":
    sys.exit(main())
