"""Per-kernel summary of rocprofv3 CSV output for runs of the fused WAH build.

    python -m indexing_amd.rocprof_summary --trace KT/kt_kernel_trace.csv \\
        --counters SQ1/sq1_counter_collection.csv SQ2/... \\
        --fetch PF/pf_counter_collection.csv --write PW/pw_counter_collection.csv \\
        --calib-fetch CF/cf_counter_collection.csv --calib-write CW/cw_counter_collection.csv \\
        --calib-bytes 1073741824 --n 268435456 --length 536621634 --keys 65536 [--narrow]

Dispatches are grouped into builds (each starts at k_begin) and the radix
launches of a build are numbered in order (launch 0 sorts by the lowest
active digit).  For every kernel launch that does work it prints the mean
duration from the trace, the algorithmic bytes of the byte model (histogram
4n; radix launch 0 reads the values and writes (value, rid) pairs, 12n; later
launches read and write pairs, 16n; tail 8n + 4L + 8K; lengths 4K) with the
GB/s and the fraction of the 8 TB/s HBM peak they give, the bytes the build
moves by design (16-bit keys when the column's values are below 2^16), the
HBM bytes the FETCH_SIZE / WRITE_SIZE counters report, scaled by what they
read for a plain copy of --calib-bytes read and written, and the SQ / TCC
counters per value.
"""
import argparse
import collections
import csv

PEAK_GBS = 8000.0


def _short(name):
    name = name.replace("void ", "")
    return name.split("(")[0]


def _label(names):
    """-> list of labels, one per dispatch in order: kernel name, with the
    launch index for the radix passes."""
    labels = []
    launch = collections.Counter()
    for nm in names:
        s = _short(nm)
        if s.startswith("wah::k_begin"):
            launch.clear()
        if "k_onesweep" in s or "k_wsweep" in s:
            labels.append(f"{s} launch {launch[s]}")
            launch[s] += 1
        else:
            labels.append(s)
    return labels


def load_trace(path):
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    labels = _label([r["Kernel_Name"] for r in rows])
    dur = collections.defaultdict(list)
    for lab, r in zip(labels, rows):
        dur[lab].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    return dur


def load_counters(path):
    """-> {label: {counter: [value per dispatch]}}"""
    per = collections.defaultdict(dict)
    names = {}
    for r in csv.DictReader(open(path)):
        d = int(r["Dispatch_Id"])
        names[d] = r["Kernel_Name"]
        per[d][r["Counter_Name"]] = per[d].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    order = sorted(names)
    labels = _label([names[d] for d in order])
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for lab, d in zip(labels, order):
        for k, v in per[d].items():
            out[lab][k].append(v)
    return out


def _mean(xs):
    return sum(xs) / len(xs) if xs else 0.0


def calib_scale(path, counter, nbytes):
    c = load_counters(path)
    best = max((_mean(v[counter]) for v in c.values() if counter in v), default=0.0)
    return nbytes / (best * 1024.0) if best else None


def byte_model(label, n, L, K, narrow):
    """(algorithmic bytes, bytes moved by design) of one launch, or None."""
    kb = 2.0 if narrow else 4.0
    if "k_histogram" in label:
        return 4.0 * n, 4.0 * n
    if "k_onesweep" in label and "true" not in label:
        i = int(label.rsplit(" ", 1)[1])
        if i == 0:
            return 12.0 * n, (4.0 + kb + 4.0) * n
        return 16.0 * n, 2.0 * (kb + 4.0) * n
    if "k_tail" in label and "true" not in label:
        return 8.0 * n + 4.0 * L + 8.0 * K, (kb + 4.0) * n + 4.0 * L + 8.0 * K
    if "k_finish" in label:
        return 4.0 * K, 4.0 * K
    return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--trace", required=True)
    ap.add_argument("--counters", nargs="*", default=[])
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--calib-fetch")
    ap.add_argument("--calib-write")
    ap.add_argument("--calib-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--n", type=float, default=0.0)
    ap.add_argument("--length", type=float, default=0.0)
    ap.add_argument("--keys", type=float, default=0.0)
    ap.add_argument("--durations-only", action="store_true",
                    help="only the per-launch durations of every wah:: kernel")
    ap.add_argument("--narrow", action="store_true")
    ap.add_argument("--min-us", type=float, default=50.0,
                    help="launches shorter than this are left out (passes with nothing to do)")
    a = ap.parse_args(argv)
    n, L, K = a.n, a.length, a.keys
    dur = load_trace(a.trace)
    if a.durations_only:
        print("| kernel launch | dispatches | mean us | min us | max us |")
        print("|---|---|---|---|---|")
        for lab, d in dur.items():
            if lab.startswith("wah::"):
                print(f"| {lab} | {len(d)} | {_mean(d):.1f} | {min(d):.1f} | {max(d):.1f} |")
        return
    if not (n and L and K):
        ap.error("--n, --length and --keys are needed for the byte model")
    ctr = collections.defaultdict(dict)
    for path in a.counters:
        for lab, cs in load_counters(path).items():
            for k, v in cs.items():
                ctr[lab][k] = _mean(v)
    fs = ws = None
    if a.fetch and a.write and a.calib_fetch and a.calib_write:
        fs = calib_scale(a.calib_fetch, "FETCH_SIZE", a.calib_bytes)
        ws = calib_scale(a.calib_write, "WRITE_SIZE", a.calib_bytes)
        for path, name in ((a.fetch, "FETCH_SIZE"), (a.write, "WRITE_SIZE")):
            for lab, cs in load_counters(path).items():
                ctr[lab][name] = _mean(cs[name])
        print(f"calibration: FETCH_SIZE x {fs:.3f}, WRITE_SIZE x {ws:.3f} "
              f"(a copy of {a.calib_bytes:.0f} bytes read and written)\n")
    # the build's launches (the ranking self-test runs once per context)
    labels = [lab for lab, d in dur.items()
              if _mean(d) >= a.min_us and lab.startswith("wah::") and "selftest" not in lab]
    print("| kernel launch | calls | mean us | alg bytes | alg GB/s | frac of 8 TB/s | design bytes "
          "| HBM bytes (counters) | HBM / design |")
    print("|---|---|---|---|---|---|---|---|---|")
    total_us = total_alg = 0.0
    for lab in labels:
        us = _mean(dur[lab])
        bm = byte_model(lab, n, L, K, a.narrow)
        alg, des = bm if bm else (0.0, 0.0)
        hbm = ""
        ratio = ""
        c = ctr.get(lab, {})
        if fs and "FETCH_SIZE" in c and "WRITE_SIZE" in c:
            h = c["FETCH_SIZE"] * 1024.0 * fs + c["WRITE_SIZE"] * 1024.0 * ws
            hbm = f"{h:.4g}"
            ratio = f"{h / des:.3f}" if des else ""
        gbs = alg / us / 1e3 if alg else 0.0
        total_us += us
        total_alg += alg
        print(f"| {lab} | {len(dur[lab])} | {us:.1f} | {alg:.4g} | {gbs:.0f} | "
              f"{gbs / PEAK_GBS:.3f} | {des:.4g} | {hbm} | {ratio} |")
    print(f"\nlaunches above: {total_us:.1f} us, {total_alg:.4g} algorithmic bytes, "
          f"{total_alg / total_us / 1e3:.0f} GB/s = {total_alg / total_us / 1e3 / PEAK_GBS:.3f} "
          f"of peak\n")
    keys = sorted({k for lab in labels for k in ctr.get(lab, {})
                   if k not in ("FETCH_SIZE", "WRITE_SIZE")})
    if keys:
        print("Counters per dispatch (device totals), and per value where marked:\n")
        print("| kernel launch | " + " | ".join(keys) + " | VALU/value | LDS/value "
              "| WAIT_ANY/WAVE_CYCLES | LDS conflict/active |")
        print("|---|" + "---|" * (len(keys) + 4))
        for lab in labels:
            c = ctr.get(lab, {})
            cells = [f"{c[k]:.3g}" if k in c else "" for k in keys]
            valu = c.get("SQ_INSTS_VALU", 0) * 64 / n
            lds = c.get("SQ_INSTS_LDS", 0) * 64 / n
            wait = c.get("SQ_WAIT_ANY", 0) / c["SQ_WAVE_CYCLES"] if c.get("SQ_WAVE_CYCLES") else 0
            conf = (c.get("SQ_LDS_BANK_CONFLICT", 0) / c["SQ_ACTIVE_INST_LDS"]
                    if c.get("SQ_ACTIVE_INST_LDS") else 0)
            print(f"| {lab} | " + " | ".join(cells) +
                  f" | {valu:.2f} | {lds:.2f} | {wait:.2f} | {conf:.2f} |")
        print("\nVALU/value and LDS/value: wave instructions x 64 / n (lane slots per value).")


if __name__ == "__main__":
    main()
