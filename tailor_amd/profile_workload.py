"""A short, fixed workload for hardware-counter profiles.

    python -m tailor_amd.profile_workload [--ny 200 --nz 100 --solves 4 --gemm-config 5]

Runs on the synthetic case: one assembly of A_bc and B_bc, one FP64
factorization of A_bc - sigma B_bc (sigma = lambda_s for omega_s = 0.02) with
the frontal GEMM tiles forced to one configuration (default 5: the 64x64
tile fed through the LDS ring, XCD-grouped order, the autotuner's choice on
most levels at 200x100; -1 keeps the untuned default),
`--solves` triangular-solve pairs, and one device-to-device copy of exactly
1 GiB (a kernel whose HBM traffic is known, used to scale the FETCH_SIZE /
WRITE_SIZE counters of the other kernels).  Prints the algorithmic bytes or
flops of each kernel family so that counter readings can be compared with
them.

    python -m tailor_amd.profile_workload summarize <kernel_trace.csv>... --command "..." [--out f.md]

    python -m tailor_amd.profile_workload summarize-pmc --fetch <counter_collection.csv>...
        --write <counter_collection.csv>... --command "..." --out-md f.md --out-json f.json

Summarises two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of
this workload: per kernel family the mean KiB per dispatch of each counter
and the HBM bytes per dispatch after scaling both counters by what they
read for the workload's 1 GiB device copy (written as markdown and as the
JSON that bench.py reads for the roofline's `traffic`).

    python -m tailor_amd.profile_workload summarize-solves <kernel_trace.csv> --command "..." [--out f.md]

Splits a kernel trace of this workload into its triangular solves (at each
k_solve_init) and reports the wall time, summed kernel time and idle gaps
per solve, and one solve's launches in order with their grids and
durations.

Summarises a `rocprofv3 --kernel-trace --output-format csv` trace (one file
per traced process, all on the same clock) as a
markdown table of per-kernel launch counts and durations, and separately
the launches inside the window that tailor_profile_factor brackets with
its k_profile_window_begin / k_profile_window_end kernels (the serialized,
event-timed factorization of bench.py's roofline), so the trace's mean
frontal-GEMM duration can be set beside the event-timed one.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def _short_name(name: str) -> str:
    """`void tailor::(anonymous namespace)::k_gemm<double, 2, ...>(int const*, ...)`
    -> `k_gemm<double, 2, ...>`."""
    import re
    m = re.search(r"(k_\w+(?:<[^()]*>)?)\s*\(", name)
    if m:
        return m.group(1)
    return name.split("(")[0].replace("void ", "").strip()


def _table(rows, total_ns, limit=40):
    out = ["| kernel | calls | total ms | avg us | share |", "|---|---|---|---|---|"]
    for name, (calls, ns) in sorted(rows.items(), key=lambda kv: -kv[1][1])[:limit]:
        out.append(f"| {name} | {calls} | {ns / 1e6:.2f} | {ns / calls / 1e3:.2f} | "
                   f"{100.0 * ns / max(total_ns, 1):.1f}% |")
    return out


def summarize(csv_paths, command: str) -> str:
    """Markdown summary of a rocprofv3 kernel trace (see the module docstring)."""
    import csv
    if isinstance(csv_paths, str):
        csv_paths = [csv_paths]
    launches = []
    for path in csv_paths:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                launches.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                                 _short_name(r["Kernel_Name"])))
    launches.sort()
    rows, total = {}, 0
    for t0, t1, name in launches:
        c, ns = rows.get(name, (0, 0))
        rows[name] = (c + 1, ns + (t1 - t0))
        total += t1 - t0
    lines = [f"Command (GPU box, repo root):", "", f"    {command}", "",
             f"{len(csv_paths)} traced processes, {len(launches)} kernel launches, "
             f"{total / 1e6:.1f} ms of summed kernel time.", ""]
    lines += _table(rows, total)
    begin = [t1 for t0, t1, n in launches if n == "k_profile_window_begin"]
    end = [t0 for t0, t1, n in launches if n == "k_profile_window_end"]
    if begin and end:
        lo, hi = begin[-1], end[-1]
        win, wtot = {}, 0
        for t0, t1, name in launches:
            if lo <= t0 and t1 <= hi:
                c, ns = win.get(name, (0, 0))
                win[name] = (c + 1, ns + (t1 - t0))
                wtot += t1 - t0
        gemm = lambda n: n.startswith("k_gemm<") or n.startswith("k_gemm_lds<")
        g_calls = sum(c for n, (c, _) in win.items() if gemm(n))
        g_ns = sum(ns for n, (_, ns) in win.items() if gemm(n))
        lines += ["", "## The event-timed factorization (tailor_profile_factor window)", "",
                  f"Window {(hi - lo) / 1e6:.2f} ms, {wtot / 1e6:.2f} ms of kernel time.  "
                  f"Frontal GEMM (k_gemm and k_gemm_lds, all tile shapes): {g_calls} launches, "
                  f"{g_ns / 1e6:.2f} ms, mean {g_ns / max(g_calls, 1) / 1e3:.2f} us per launch.", ""]
        lines += _table(win, wtot, limit=25)
    return "\n".join(lines) + "\n"


def summarize_solves(csv_path: str, command: str) -> str:
    """Markdown summary of the triangular solves in a kernel trace of this
    workload (see the module docstring)."""
    import csv
    launches = []
    with open(csv_path, newline="") as f:
        for r in csv.DictReader(f):
            grid = tuple(int(r.get(k, 0) or 0) for k in ("Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"))
            launches.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                             _short_name(r["Kernel_Name"]), grid))
    launches.sort()
    starts = [i for i, x in enumerate(launches) if x[2] == "k_solve_init"]
    solves = []
    for a, b in zip(starts, starts[1:]):
        seg = launches[a:b]
        # a solve ends with its last GEMV; later launches (the next workload
        # phase) are not part of it
        last = max(i for i, x in enumerate(seg) if x[2].startswith("k_gemv") or i == 0)
        solves.append(seg[:last + 1])
    solves = solves[1:]  # the warm-up solve
    if not solves:
        return "no solves found\n"
    wall = sorted(s[-1][1] - s[0][0] for s in solves)
    busy = sorted(sum(t1 - t0 for t0, t1, _, _ in s) for s in solves)
    gaps = sorted(sum(max(0, s[i + 1][0] - s[i][1]) for i in range(len(s) - 1)) for s in solves)
    med = lambda v: v[len(v) // 2] / 1e3
    one = solves[len(solves) // 2]
    lines = ["Command (GPU box, repo root):", "", f"    {command}", "",
             f"* solves analysed: {len(solves)} (the warm-up solve excluded)",
             f"* wall per solve (first start to last end): median {med(wall):.1f} us, "
             f"range {wall[0] / 1e3:.1f}-{wall[-1] / 1e3:.1f}",
             f"* summed kernel durations per solve: median {med(busy):.1f} us",
             f"* summed idle gaps between consecutive launches: median {med(gaps):.1f} us", "",
             "One solve (the median one by position), launches in start order:", "",
             "| # | kernel | grid (work-items x, fronts, column tiles) | us |", "|---|---|---|---|"]
    for i, (t0, t1, name, g) in enumerate(one):
        lines.append(f"| {i} | {name} | {g[0]} x {g[1]} x {g[2]} | {(t1 - t0) / 1e3:.1f} |")
    return "\n".join(lines) + "\n"


def _family(name: str) -> str:
    """Counter rows are grouped by kernel family: k_gemm by element type (all
    tile shapes, the FP64 LDS-ring tile k_gemm_lds included), k_gemv by type
    and columns per wave, others by full name."""
    import re
    if name.startswith("k_gemm_lds<"):
        return "k_gemm<double>"
    m = re.match(r"(k_gemm|k_gemv)<([^>]*)>", name)
    if not m:
        return name
    args = [a.strip() for a in m.group(2).split(",")]
    return f"{m.group(1)}<{', '.join(args[:1 if m.group(1) == 'k_gemm' else 2])}>"


def _pmc_rows(paths, counter):
    """{dispatch id: (kernel family, summed counter value)} of one counter's runs."""
    import csv
    out = {}
    for path in paths:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                key = (path, r["Dispatch_Id"])
                fam = _family(_short_name(r["Kernel_Name"]))
                val = float(r["Counter_Value"])
                prev = out.get(key, (fam, 0.0))
                out[key] = (fam, prev[1] + val)
    return out


def summarize_pmc(fetch_paths, write_paths, command: str):
    """(markdown, json dict) of the FETCH_SIZE and WRITE_SIZE runs (module
    docstring); the counters are in KiB."""
    fetch, write = _pmc_rows(fetch_paths, "FETCH_SIZE"), _pmc_rows(write_paths, "WRITE_SIZE")

    def per_family(rows):
        agg = {}
        for fam, v in rows.values():
            n, tot = agg.get(fam, (0, 0.0))
            agg[fam] = (n + 1, tot + v)
        return agg

    fa, wa = per_family(fetch), per_family(write)
    copy = "__amd_rocclr_copyBuffer"
    gib_kib = float(1 << 20)
    fscale = round(gib_kib / fa[copy][1], 1) if copy in fa and fa[copy][1] > 0 else 1.0
    wscale = round(gib_kib / wa[copy][1], 1) if copy in wa and wa[copy][1] > 0 else 1.0
    fams = sorted(set(fa) | set(wa), key=lambda k: -(fa.get(k, (1, 0.0))[1] + wa.get(k, (1, 0.0))[1]))
    lines = ["Commands (GPU box, repo root), one counter per run:", "", f"    {command}", "",
             f"Scale factors from the 1 GiB device copy ({copy}): FETCH_SIZE summed to "
             f"{fa.get(copy, (0, 0.0))[1]:,.0f} KiB (scale {fscale}), WRITE_SIZE to "
             f"{wa.get(copy, (0, 0.0))[1]:,.0f} KiB (scale {wscale}).", "",
             "| kernel | dispatches | FETCH KiB/dispatch | WRITE KiB/dispatch | "
             "HBM bytes/dispatch (scaled) |", "|---|---|---|---|---|"]
    kernels = {}
    for fam in fams:
        nf, tf = fa.get(fam, (0, 0.0))
        nw, tw = wa.get(fam, (0, 0.0))
        n = max(nf, nw)
        fk = tf / nf if nf else 0.0
        wk = tw / nw if nw else 0.0
        traffic = (fk * fscale + wk * wscale) * 1024.0
        lines.append(f"| {fam} | {n} | {fk:.1f} | {wk:.1f} | {traffic / 1e6:.1f} MB |")
        kernels[fam] = {"dispatches": n, "traffic_bytes_per_dispatch": int(round(traffic)),
                        "fetch_kib": round(fk, 1), "write_kib": round(wk, 1)}
    js = {"fetch_scale": fscale, "write_scale": wscale, "kernels": kernels}
    return "\n".join(lines) + "\n", js


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else list(argv)
    if argv and argv[0] == "summarize-pmc":
        import json
        sp = argparse.ArgumentParser(prog="profile_workload summarize-pmc")
        sp.add_argument("--fetch", nargs="+", required=True)
        sp.add_argument("--write", nargs="+", required=True)
        sp.add_argument("--command", required=True)
        sp.add_argument("--out-md", required=True)
        sp.add_argument("--out-json", required=True)
        b = sp.parse_args(argv[1:])
        text, js = summarize_pmc(b.fetch, b.write, b.command)
        with open(b.out_md, "w") as f:
            f.write(text)
        with open(b.out_json, "w") as f:
            json.dump(js, f, indent=1)
        return 0
    if argv and argv[0] == "summarize-solves":
        sp = argparse.ArgumentParser(prog="profile_workload summarize-solves")
        sp.add_argument("csv")
        sp.add_argument("--command", required=True, help="the profiled command, for the record")
        sp.add_argument("--out")
        b = sp.parse_args(argv[1:])
        text = summarize_solves(b.csv, b.command)
        if b.out:
            with open(b.out, "w") as f:
                f.write(text)
        else:
            print(text)
        return 0
    if argv and argv[0] == "summarize":
        sp = argparse.ArgumentParser(prog="profile_workload summarize")
        sp.add_argument("csv", nargs="+")
        sp.add_argument("--command", required=True, help="the profiled command, for the record")
        sp.add_argument("--out")
        b = sp.parse_args(argv[1:])
        text = summarize(b.csv, b.command)
        if b.out:
            with open(b.out, "w") as f:
                f.write(text)
        else:
            print(text)
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--solves", type=int, default=4)
    ap.add_argument("--gemm-config", type=int, default=5)
    a = ap.parse_args(argv)
    import torch
    import tailor_amd as ta
    from tailor_amd import case as cs
    c = cs.synthetic_case(a.ny, a.nz, omega_shift=0.02)
    ctx = ta.Context(0)
    ctx.load_case(c)
    ctx.assemble(diagnostics=False)
    if a.gemm_config >= 0:
        ctx.set_gemm_config(a.gemm_config)
    fd = ctx.factor(c.lambda_shift)
    ms, nbytes = ctx.time_linear_solve(a.solves)
    kms, kbytes = ctx.time_operator(a.solves)
    nodes = a.ny * a.nz
    asm = 16 * 25 * (nodes * 77 + nodes) + 128 * nodes + 3 * 8 * (a.ny * 11 + a.nz * 7)
    print(f"assembly algorithmic bytes {asm}")
    print(f"factorization LU flops {fd['flops']:.6e}, arena bytes {fd['arena_bytes']}")
    print(f"solve algorithmic bytes {nbytes:.6e} per solve, {ms:.4f} ms per solve "
          f"({a.solves + 1} solves incl. warm-up)")
    print(f"SpMV algorithmic bytes {kbytes:.6e} per application, {kms:.4f} ms "
          f"({a.solves + 1} applications incl. warm-up)")
    ctx.close()
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print("device copy: 1073741824 bytes read, 1073741824 bytes written")
    return 0


if __name__ == "__main__":
    sys.exit(main())
