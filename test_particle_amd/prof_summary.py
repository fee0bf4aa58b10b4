"""Summarise rocprofv3 SQLite outputs (``*_results.db``) as text.

    python -m test_particle_amd.prof_summary trace DB [DB...]
        per-kernel calls, total and average duration
    python -m test_particle_amd.prof_summary pmc DB [--kernel SUBSTR]
        per-kernel average of each collected counter per dispatch
    python -m test_particle_amd.prof_summary headline DIR --out JSON
        the per-launch counters of the PIC-grid step kernels that bench.py's
        roofline reads: DIR/<form>_<steps>_<pass>/ for form f64/f32, steps
        per launch 20/100, pass sq/fetch/write, and DIR/copy_fetch,
        DIR/copy_write for the byte calibration
"""
import argparse
import glob
import json
import os
import sqlite3
import sys
from collections import defaultdict


def trace(dbs):
    rows = defaultdict(lambda: [0, 0.0])
    for db in dbs:
        con = sqlite3.connect(db)
        for name, dur in con.execute("select name, duration from kernels"):
            r = rows[name]
            r[0] += 1
            r[1] += dur
    tot = sum(v[1] for v in rows.values()) or 1.0
    print(f"{'calls':>7} {'total_ms':>10} {'avg_us':>10} {'pct':>6}  kernel")
    for name, (n, d) in sorted(rows.items(), key=lambda kv: -kv[1][1]):
        print(f"{n:7d} {d / 1e6:10.3f} {d / n / 1e3:10.2f} {100 * d / tot:6.2f}  {name[:110]}")


def pmc(dbs, kernel=None):
    acc = defaultdict(lambda: [0, 0.0])
    for db in dbs:
        con = sqlite3.connect(db)
        q = "select kernel_name, counter_name, value, dispatch_id from counters_collection"
        for name, cname, value, _ in con.execute(q):
            if kernel and kernel not in name:
                continue
            a = acc[(name, cname)]
            a[0] += 1
            a[1] += value
    print(f"{'dispatches':>10} {'mean_per_dispatch':>18}  counter  kernel")
    for (name, cname), (n, v) in sorted(acc.items()):
        print(f"{n:10d} {v / n:18.1f}  {cname}  {name[:100]}")


def per_dispatch(dbs, kernel):
    """Mean over the dispatches of kernels whose name contains `kernel` of
    each counter (summed over a dispatch's rows); and the dispatch count."""
    tot = defaultdict(float)
    disp = set()
    for db in dbs:
        con = sqlite3.connect(db)
        q = "select kernel_name, counter_name, value, dispatch_id from counters_collection"
        for name, cname, value, did in con.execute(q):
            if kernel in name:
                tot[cname] += value
                disp.add((db, did))
    nd = max(len(disp), 1)
    return {k: v / nd for k, v in tot.items()}, len(disp)


def headline(root, out, particles):
    dbs = lambda sub: glob.glob(os.path.join(root, sub, "**", "*_results.db"), recursive=True)
    cf, _ = per_dispatch(dbs("copy_fetch"), "copy")
    cw, _ = per_dispatch(dbs("copy_write"), "copy")
    copy_bytes = float(os.environ.get("COPY_BYTES", 1024 * (1 << 20)))
    fscale = copy_bytes / (cf["FETCH_SIZE"] * 1024.0)
    wscale = copy_bytes / (cw["WRITE_SIZE"] * 1024.0)
    res = {}
    for form, tag in (("plain_fp64", "f64"), ("fp32_mixed", "f32")):
        per = {}
        for spl in (20, 100):
            sq, nsq = per_dispatch(dbs(f"{tag}_{spl}_sq"), "k_track<")
            fe, _ = per_dispatch(dbs(f"{tag}_{spl}_fetch"), "k_track<")
            wr, _ = per_dispatch(dbs(f"{tag}_{spl}_write"), "k_track<")
            if not sq or not fe or not wr:
                continue
            waves = sq.get("SQ_WAVES", particles / 64.0)
            per[str(spl)] = {
                "valu_insts": sq["SQ_INSTS_VALU"],
                "valu_insts_per_wave_step": sq["SQ_INSTS_VALU"] / waves / spl,
                "sq_active_inst_valu": sq.get("SQ_ACTIVE_INST_VALU"),
                "vmem_insts": sq.get("SQ_INSTS_VMEM"),
                "wave_cycles": sq.get("SQ_WAVE_CYCLES"),
                "wait_inst_any": sq.get("SQ_WAIT_INST_ANY"),
                "busy_cycles": sq.get("SQ_BUSY_CYCLES"),
                "grbm_gui_active": sq.get("GRBM_GUI_ACTIVE"),
                "waves": waves, "dispatches": nsq,
                "fetch_kib": fe["FETCH_SIZE"], "write_kib": wr["WRITE_SIZE"],
                "hbm_bytes": fe["FETCH_SIZE"] * 1024.0 * fscale +
                             wr["WRITE_SIZE"] * 1024.0 * wscale}
        if per:
            res[form] = {"kernel": "k_track<BORIS, GRID_REF, %s>" %
                         ("FP64" if tag == "f64" else "FP32"),
                         "particles": particles, "per_launch": per}
    res["calibration"] = {"copy_bytes_each_way": copy_bytes, "fetch_scale": fscale,
                          "write_scale": wscale}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["trace", "pmc", "headline"])
    ap.add_argument("dbs", nargs="+")
    ap.add_argument("--kernel")
    ap.add_argument("--out")
    ap.add_argument("--particles", type=int, default=1 << 20)
    a = ap.parse_args()
    if a.what == "trace":
        trace(a.dbs)
    elif a.what == "headline":
        headline(a.dbs[0], a.out, a.particles)
    else:
        pmc(a.dbs, a.kernel)
    sys.stdout.flush()


if __name__ == "__main__":
    main()
