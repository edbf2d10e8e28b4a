"""Text summaries of rocprofv3 result databases (the `*_results.db` files
that `rocprofv3 -d DIR -o NAME` writes).

    python -m walicxe2d_amd.profsum kernels DB [--last N]
        per-kernel calls, total / average duration, share, VGPRs, LDS,
        scratch; with --last, each kernel's average over its last N launches
    python -m walicxe2d_amd.profsum valu DB [DB ...]
        per-wave SQ counters of `--pmc` runs of one command and the VALU
        issue share of each kernel
    python -m walicxe2d_amd.profsum traffic --copy-fetch DB --copy-write DB \
            --fetch DB --write DB [--copy-mib 256]
        HBM bytes per zone-update of the hot-path kernels from two
        `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs, each counter scaled by
        what it reads for a plain device copy of --copy-mib MiB per dispatch.

A hot-path kernel launches one workgroup per leaf block (of 256 threads, or
512 for stage A), so its workgroup count times the 256 interior cells of a
block is the zone-updates of that launch.
"""
from __future__ import annotations

import argparse
import collections
import json
import re
import sqlite3
import sys

HOT = ("k_stageA", "k_stageB", "k_fillI_cool", "k_ens_stageA", "k_ens_stageB", "k_ens_fillI_cool",
       "k_cool_surrogate", "k_mlp_train_step")


def short(name: str) -> str:
    m = re.search(r"wx::(k_\w+(?:<[^>]*>)?)", name)
    return m.group(1) if m else name[:48]


def kernels_last(db: str, last: int) -> str:
    """Average duration of the last `last` launches of each kernel (the
    window a run times with HIP events at its end)."""
    cur = sqlite3.connect(db).cursor()
    rows = collections.defaultdict(list)
    for name, start, end in cur.execute("select name, start, end from kernels order by start"):
        rows[name].append(end - start)
    out = [f"| kernel | launches averaged (the last) | average (us) |", "|---|---|---|"]
    for n, d in sorted(rows.items(), key=lambda kv: -sum(kv[1][-last:])):
        w = d[-last:]
        out.append(f"| `{short(n)}` | {len(w)} | {sum(w) / len(w) / 1e3:.3f} |")
    return "\n".join(out)


def kernels(db: str) -> str:
    cur = sqlite3.connect(db).cursor()
    res = {}
    for name, vg, ag, lds, scr in cur.execute(
            "select name, max(vgpr_count), max(accum_vgpr_count), max(lds_size), max(scratch_size) "
            "from kernels group by name"):
        res[name] = (vg, ag, lds, scr)
    rows = list(cur.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    out = ["| kernel | calls | total (ms) | average (us) | % | VGPR | AGPR | LDS (B) | scratch (B) |",
           "|---|---|---|---|---|---|---|---|---|"]
    for n, c, t, a, p in rows:
        vg, ag, lds, scr = res.get(n, ("", "", "", ""))
        out.append(f"| `{short(n)}` | {c} | {t / 1e3:.1f} | {a:.3f} | {p:.1f} | {vg} | {ag} | {lds} | {scr} |")
    return "\n".join(out)


def valu(dbs, clock_ghz: float = 2.4, simds: int = 1024) -> str:
    """Per-kernel SQ counters per wave from one or more `--pmc` runs of the
    same command (each run's counters divided by that run's SQ_WAVES), and
    the VALU issue share: SQ_INSTS_VALU x 4 cycles (a wave64 instruction
    takes 4 cycles on a 16-lane SIMD) over the SIMD-cycles of the launch
    (duration x clock x SIMDs)."""
    per = collections.defaultdict(dict)
    dur = {}
    for db in dbs:
        cur = sqlite3.connect(db).cursor()
        acc = collections.defaultdict(lambda: collections.defaultdict(float))
        d = collections.defaultdict(lambda: [0.0, set()])
        for name, did, cn, val, du in cur.execute(
                "select kernel_name, dispatch_id, counter_name, value, duration from counters_collection"):
            k = short(name)
            acc[k][cn] += val
            if did not in d[k][1]:
                d[k][1].add(did)
                d[k][0] += du
        for k, c in acc.items():
            w = c.get("SQ_WAVES", 0)
            if not w:
                continue
            for cn, v in c.items():
                if cn != "SQ_WAVES":
                    per[k][cn] = v / w
            per[k]["waves/launch"] = w / len(d[k][1])
            if k not in dur:
                dur[k] = d[k][0] / len(d[k][1])
    out = ["| kernel | avg launch (us) | waves/launch | VALU/wave | VALU issue share | other counters per wave |",
           "|---|---|---|---|---|---|"]
    for k in sorted(per, key=lambda k: -dur.get(k, 0)):
        c = per[k]
        if "SQ_INSTS_VALU" not in c or k.startswith("__"):
            continue
        share = c["SQ_INSTS_VALU"] * c["waves/launch"] * 4 / (simds * dur[k] * 1e-9 * clock_ghz * 1e9)
        rest = ", ".join(f"{n[3:]} {v:.0f}" for n, v in sorted(c.items()) if n not in ("SQ_INSTS_VALU", "waves/launch"))
        out.append(f"| `{k}` | {dur[k] / 1e3:.2f} | {c['waves/launch']:.0f} | {c['SQ_INSTS_VALU']:.0f} | {share:.2f} | {rest} |")
    return "\n".join(out)


def _per_kernel(db: str):
    cur = sqlite3.connect(db).cursor()
    acc = collections.defaultdict(lambda: [0.0, 0, 0])
    for name, val, gs, ws in cur.execute("select kernel_name, value, grid_size, workgroup_size "
                                         "from counters_collection"):
        a = acc[name]
        a[0] += val
        a[1] += 1
        a[2] += gs // max(ws, 1) * 256   # zone-updates: one leaf block per workgroup
    return acc


def traffic(copy_fetch: str, copy_write: str, fetch: str, write: str, copy_mib: int) -> dict:
    out = {"factors": {}, "bytes_per_zone_update": {}}
    for ctr, cdb, kdb in (("FETCH_SIZE", copy_fetch, fetch), ("WRITE_SIZE", copy_write, write)):
        cp = _per_kernel(cdb)
        kib = [v[0] / v[1] for k, v in cp.items() if "copyBuffer" in k or "elementwise" in k.lower()]
        if not kib:
            raise SystemExit(f"no copy kernel in {cdb}")
        fac = copy_mib * 1024 / max(kib)
        out["factors"][ctr] = fac
        for k, (tot, n, gs) in _per_kernel(kdb).items():
            s = short(k).split("<")[0]
            if s in HOT:
                out["bytes_per_zone_update"].setdefault(s, {})[ctr] = tot * 1024 * fac / gs
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--last", type=int, default=0, help="also the average of each kernel's last N launches")
    v = sub.add_parser("valu")
    v.add_argument("dbs", nargs="+")
    t = sub.add_parser("traffic")
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-mib", type=int, default=256)
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        print(kernels(a.db))
        if a.last:
            print()
            print(kernels_last(a.db, a.last))
    elif a.cmd == "valu":
        print(valu(a.dbs))
    else:
        print(json.dumps(traffic(a.copy_fetch, a.copy_write, a.fetch, a.write, a.copy_mib), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
