#!/usr/bin/env python3
"""Benchmark: converged eigenpairs/s of the batched Bloch eigensolver on one MI355X.

Workload (512^2 rod, the reference's path-scan configuration): unit cell with a
centred rod (eps = 8.9, r = 0.2, 1/eps sampled at cell corners), 512 x 512
bilinear Nedelec edge grid (524,288 DOFs, 8 multigrid levels), 64 Bloch
parameters arc-length uniform on Gamma-X-M-Gamma, the 10 smallest non-zero
eigenvalues per k with block 12, block PINVIT + Ritz with V(2,2) AFW multigrid
(2 cycles per step) and 5 nodal V-cycles of gradient lifting, stopping at
relative residual |K e - l M e| / (|l| |M e|) <= 1e-8.  Seeded complex Gaussian
start blocks.  All 64 problems are solved in one batch (every kernel launch
covers 768 vectors).

One step = one full solve of the 64 problems to convergence.  Run in mixed
precision (FP32 multigrid, FP64 residuals/Gram/Ritz) -- the headline `value` --
and in plain FP64.  Each eigenvalue of both runs is compared with the oracle's
FP64 eigenvalues at four path points (tests/golden/rod512_path64.json), and
the two runs with each other.

The CPU baseline is the oracle (the FP64 C restatement of the reference's
algorithm: line-ordered AFW smoother, every vector updated every step) solving
one path point to the same tolerance in a child process with its own OpenMP
team (OMP_NUM_THREADS = the cores given to this process, OMP_PROC_BIND=close).

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, N0, LEVELS, NK, NEIG, BLOCK, TOL = 512, 4, 7, 64, 10, 12, 1e-8


def kpath(nk):
    """Arc-length uniform points on Gamma(0,0) -> X(pi,0) -> M(pi,pi) -> Gamma."""
    total = (2.0 + math.sqrt(2.0)) * math.pi
    kx, ky = [], []
    for m in range(nk):
        s = m * total / (nk - 1)
        if s <= math.pi:
            kx.append(s); ky.append(0.0)
        elif s <= 2 * math.pi:
            kx.append(math.pi); ky.append(s - math.pi)
        else:
            t = (s - 2 * math.pi) / math.sqrt(2.0)
            kx.append(max(math.pi - t, 0.0)); ky.append(max(math.pi - t, 0.0))
    kx[0] = ky[0] = 0.0
    kx[-1] = ky[-1] = 0.0
    return np.array(kx), np.array(ky)


def rod_eta(n):
    eta = np.ones((n, n))
    h = 1.0 / n
    for j in range(n):
        for i in range(n):
            x, y = i * h, j * h
            if (x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) <= 0.2 * 0.2:
                eta[j, i] = 1.0 / 8.9
    return eta


def random_field_eta(n, seed, sigma=6.0, eps_hi=8.9):
    """Seeded Gaussian random field on the n x n cells, smoothed with a
    periodic Gaussian kernel of sigma cells (FFT), thresholded at 0: eps =
    eps_hi where positive, 1 elsewhere.  Returns eta = 1/eps per cell."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n, n))
    k = np.fft.fftfreq(n)
    kk = k[:, None] ** 2 + k[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(f) * np.exp(-2.0 * (np.pi * sigma) ** 2 * kk)))
    return np.where(f > 0.0, 1.0 / eps_hi, 1.0)


def rods_eta(n, rods, eps_base=1.0):
    """1/eps at cell corners (i h, j h), h = 1/n, for circular rods given as
    (cx, cy, r, eps) in the unit cell (periodic images included); the
    corner sampling of the reference's pattern rasteriser."""
    eta = np.full((n, n), 1.0 / eps_base)
    x = np.arange(n) / n
    X, Y = np.meshgrid(x, x)  # X[j, i] = i h, Y[j, i] = j h
    for cx, cy, r, eps in rods:
        dx = np.abs(X - cx)
        dy = np.abs(Y - cy)
        dx = np.minimum(dx, 1.0 - dx)
        dy = np.minimum(dy, 1.0 - dy)
        eta[dx * dx + dy * dy <= r * r] = 1.0 / eps
    return eta


def supercell_eta(n, seed):
    """2 x 2 supercell of rods: centres at the sub-cell centres, radii
    seed-uniform in [0.15, 0.25] of the sub-cell side (0.5), eps
    seed-uniform in [6, 12]."""
    rng = np.random.default_rng(seed)
    rods = [((a + 0.5) / 2, (b + 0.5) / 2, 0.5 * rng.uniform(0.15, 0.25), rng.uniform(6.0, 12.0))
            for b in range(2) for a in range(2)]
    return rods_eta(n, rods)


def defect_supercell_eta(n):
    """4 x 4 supercell of rods (eps 8.9, radius 0.2 of the sub-cell side
    0.25) with the rod at sub-cell (2, 2) removed (a point defect)."""
    rods = [((a + 0.5) / 4, (b + 0.5) / 4, 0.2 * 0.25, 8.9) for b in range(4) for a in range(4) if (a, b) != (2, 2)]
    return rods_eta(n, rods)


def bz_grid(nk):
    """The uniform nk x nk Bloch grid k_i = -pi + 2 pi i / (nk - 1) of the
    reference's blochsurface (both ends included)."""
    k = -np.pi + 2.0 * np.pi * np.arange(nk) / (nk - 1)
    KX, KY = np.meshgrid(k, k)
    return KX.ravel(), KY.ravel()


def run_gpu(M, prec, eta, kx, ky, steps, warmup, maxit, graphs, autotune, method, kernels=True, march_rows=192):
    # with autotune the first (warmup) solve also times the smoother tilings
    S = M.Solver(eta, n0=N0, levels=LEVELS, precision=prec, block=BLOCK, max_problems=len(kx), graphs=graphs,
                 autotune=autotune, method=method, march_rows=march_rows)
    for _ in range(warmup):
        S.solve(kx, ky, NEIG, tol=TOL, maxit=maxit, seed=1)
    times, res = [], None
    for k in range(steps):
        t0 = time.perf_counter()
        res = S.solve(kx, ky, NEIG, tol=TOL, maxit=maxit, seed=1)  # synchronises at the end
        times.append(time.perf_counter() - t0)
    st = S.stats()
    st["tiling"] = S.tuning()
    # live kernel timing on the library stream (HIP events) on the final state
    kern = {}
    for kind, name in () if not kernels else ((12, "afw_pass_zero_resid"), (13, "afw_pass_resid"), (14, "afw_pass_prolong"),
                                              (15, "node_pass_zero_resid"), (16, "node_pass_resid"),
                                              (17, "node_pass_prolong"),
                       (10, "afw_march_2sweeps"), (11, "node_march_2sweeps"), (6, "afw_march_sweep"),
                       (7, "node_march_sweep"), (4, "afw_tile_sweep"), (5, "node_tile_sweep"), (9, "gram"),
                       (8, "device_copy"), (1, "edge_vcycle"), (2, "node_vcycle")):
        reps = 4 if kind in (1, 2) else 20
        ms, by = S.time_kernel(kind, reps)
        kern[name] = dict(ms=ms, bytes=by, gbs=by / (ms * 1e-3) / 1e9)
    S.close()
    return times, res, st, kern


CPU_K = 5  # the path point the CPU baseline solves
METHODS = {0: "block PINVIT", 1: "LOBPCG"}


def cpu_cores():
    """Cores this process may use: OMP_NUM_THREADS when set (the GPU box sets
    its CPU share there), else the affinity mask."""
    n = os.environ.get("OMP_NUM_THREADS", "")
    return int(n) if n.isdigit() and int(n) > 0 else len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_solve(nk, p, tol):
    """Child-process body of the CPU baseline: the oracle solves path point p
    of the workload from a seeded start block to relative residual tol (the
    per-k coarse set-up included, as in the GPU solve's timing).  Prints JSON."""
    import oracle as O
    eta = rod_eta(N)
    kx, ky = kpath(nk)
    mg = O.Multigrid(N0, LEVELS, eta, colour_order=0, quirk=True)
    E0 = O.random_block(BLOCK, 2 * N * N, 1)
    t0 = time.perf_counter()
    mg.set_bloch(kx[p], ky[p])
    lam, _, rel, _, steps = mg.block_solve(E0, NEIG, tol=tol, maxit=400)
    dt = time.perf_counter() - t0
    print(json.dumps(dict(seconds=dt, steps=int(steps), max_relres=float(rel[:NEIG].max()),
                          lam=[float(v) for v in lam[:NEIG]])))


def cpu_baseline(nk):
    """The oracle's solve of one path point in a child process whose OpenMP
    team is set before the runtime loads: OMP_NUM_THREADS = cpu_cores(),
    OMP_PROC_BIND=close, OMP_PLACES=cores."""
    import subprocess
    cores = cpu_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-solve", str(CPU_K), "--nk", str(nk)],
                         env=env, check=True, stdout=subprocess.PIPE, text=True).stdout
    r = json.loads(out.strip().splitlines()[-1])
    return dict(value=NEIG / r["seconds"], unit="eigenpairs/s", cores=cores, kind="port",
                cpu_model=cpu_model(), nproc=os.cpu_count(),
                sample="oracle (FP64 C restatement, line-ordered AFW, every vector updated every step) solving "
                       "path point %d of %d at 512^2 to relres %.0e from a seeded start: %d block steps, %.1f s, "
                       "max relres %.2e; OMP_NUM_THREADS=%d OMP_PROC_BIND=close" %
                       (CPU_K, nk, TOL, r["steps"], r["seconds"], r["max_relres"], cores),
                steps=r["steps"], seconds=round(r["seconds"], 2))


def oracle_errors(out, kx, ky):
    """Relative error of each eigenvalue index against the oracle's fixture at
    the fixture's path points (max over those points), for each run."""
    path = os.path.join(ROOT, "tests", "golden", "rod512_path64.json")
    if not os.path.exists(path):
        return None
    g = json.load(open(path))
    ref = g.get("hermitian", g.get("quirk"))
    sel = g["sel"]
    if len(kx) != 64 or not np.allclose(kx[sel], g["kx"]) or not np.allclose(ky[sel], g["ky"]):
        return None
    lam_ref = np.array(ref["lam"])[:, :NEIG]
    res = {"kpoints": sel, "oracle_block": g["oracle_block"], "oracle_max_relres": max(ref["max_relres"])}
    for name, v in out.items():
        rel = np.abs(v["res"]["lam"][sel, :NEIG] - lam_ref) / np.abs(lam_ref)
        res[name] = [float("%.3g" % e) for e in rel.max(axis=0)]
        res[name + "_max"] = float(rel.max())
    return res


def summarize_trace(path, top=40):
    """Summary of a rocprofv3 CSV: a kernel trace (per kernel and launch
    shape: launches, mean and total duration) or a counter collection (per
    kernel, launch shape and counter: mean value per dispatch)."""
    import csv
    rows = list(csv.DictReader(open(path)))
    if not rows:
        return
    counters = "Counter_Name" in rows[0]
    agg = {}
    for r in rows:
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        gx = r.get("Grid_Size_X", r.get("Grid_Size", "?"))
        key = (name, gx, r.get("Grid_Size_Y", "1"))
        if counters:
            key = key + (r["Counter_Name"],)
            v = float(r["Counter_Value"])
        else:
            v = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        a = agg.setdefault(key, [0, 0.0])
        a[0] += 1
        a[1] += v
    items = sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]
    if counters:
        print("%-44s %10s %6s %-12s %8s %14s" % ("kernel", "grid_x", "grid_y", "counter", "dispatch", "mean value"))
        for (n, gx, gy, cn), (k, t) in items:
            print("%-44s %10s %6s %-12s %8d %14.1f" % (n[:44], gx, gy, cn, k, t / k))
    else:
        tot = sum(v[1] for v in agg.values())
        print("%-44s %10s %6s %8s %12s %12s %7s" % ("kernel", "grid_x", "grid_y", "launches", "mean us", "total ms", "share"))
        for (n, gx, gy), (k, t) in items:
            print("%-44s %10s %6s %8d %12.1f %12.2f %6.2f%%" % (n[:44], gx, gy, k, t / k, t * 1e-3, 100 * t / tot))
        print("all kernels: %.3f s" % (tot * 1e-6))
        fam = {}
        for (n, gx, gy), (k, t) in agg.items():
            f = n.split(",")[0] + (">" if "<" in n and "," in n else "")
            a = fam.setdefault(f, [0, 0.0])
            a[0] += k
            a[1] += t
        print("\n%-44s %8s %12s %7s" % ("kernel family (all shapes)", "launches", "total ms", "share"))
        for f, (k, t) in sorted(fam.items(), key=lambda kv: -kv[1][1])[:20]:
            print("%-44s %8d %12.2f %6.2f%%" % (f[:44], k, t * 1e-3, 100 * t / tot))
        # per solve: from a coarse set-up (k_coarse_setup, the first launch of
        # mx_solve) to the last defect kernel before the next set-up
        ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                     r["Kernel_Name"].split("(")[0].replace("void ", "")) for r in rows)
        starts = [i for i, e in enumerate(ev) if e[2].startswith("k_coarse_setup")] + [len(ev)]
        print("\nper solve: wall (first launch to last launch end), kernel time, GPU busy, kernel families")
        for s, (a, b) in enumerate(zip(starts[:-1], starts[1:])):
            last = max([i for i in range(a, b) if ev[i][2].startswith("k_defect")], default=None)
            if last is None:
                continue
            seg = ev[a:last + 1]
            wall = (seg[-1][1] - seg[0][0]) * 1e-9
            busy = sum(e[1] - e[0] for e in seg) * 1e-9
            fam = {}
            for e in seg:
                f = e[2].split("<")[0]
                fam[f] = fam.get(f, 0.0) + (e[1] - e[0]) * 1e-6
            top = sorted(fam.items(), key=lambda kv: -kv[1])[:10]
            print("  solve %d: wall %.3f s, kernels %.3f s, busy %.1f%%" % (s + 1, wall, busy, 100 * busy / wall))
            print("    " + ", ".join("%s %.0f ms (%.1f%%)" % (f, t, 100 * t * 1e-3 / busy) for f, t in top))


ROOF_KERNEL = "k_afw_march<float, 5, 2>"
ROOF_TRAFFIC = os.path.join(ROOT, "profiles", "afw_march_pass5_f32_512_rod_64k_traffic.json")


def counter_traffic(fetch_csv, write_csv, kernel=ROOF_KERNEL):
    """HBM bytes per launch of `kernel` (its finest-level launches over the
    largest batch) from rocprofv3 FETCH_SIZE and WRITE_SIZE runs (KiB per
    dispatch), each scaled by what the same counter reads for bench.py's
    timing copies of the batch's finest-level multigrid vectors (20 equal
    device copies whose bytes each way are known exactly)."""
    import csv
    out = {"kernel": kernel}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        ker = [r for r in rows if r["Kernel_Name"].replace("void ", "").startswith(kernel)]
        gmax = max(int(r["Grid_Size"]) for r in ker)
        ker = [float(r["Counter_Value"]) for r in ker if int(r["Grid_Size"]) == gmax]
        # bench's timing copies: the largest group of (at least 10) equal copies
        cop = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        groups = {}
        for c_ in cop:
            groups.setdefault(round(c_ / 64.0), []).append(c_)
        cop = max((g for g in groups.values() if len(g) >= 10), key=lambda g: g[0])
        out[counter] = {"kernel_kib": sum(ker) / len(ker), "kernel_dispatches": len(ker), "grid_size": gmax,
                        "copy_kib": sum(cop) / len(cop), "copy_dispatches": len(cop)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--maxit", type=int, default=400)
    ap.add_argument("--nk", type=int, default=NK)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", type=int, choices=(0, 1), default=1,
                    help="choose smoother tilings with the autotuner during the warmup solve (1) or by the fixed "
                         "march_min/march_rows rule (0)")
    ap.add_argument("--march-rows", type=int, default=192,
                    help="rows per row-marching workgroup of the fixed tiling rule (used with --autotune 0)")
    ap.add_argument("--graphs", type=int, choices=(0, 1), default=0,
                    help="replay each block step's preconditioner from a HIP graph (1) or launch it eagerly (0)")
    ap.add_argument("--cpu-solve", type=int, metavar="K", default=None,
                    help="(child process of the CPU baseline) solve path point K with the oracle, print JSON, exit")
    ap.add_argument("--method", choices=("lobpcg", "pinvit"), default="lobpcg",
                    help="eigensolver of the headline runs: LOBPCG (Rayleigh-Ritz over the block, its "
                         "preconditioned update and the previous direction) or block PINVIT")
    ap.add_argument("--other-method", type=int, choices=(0, 1), default=1,
                    help="also run the other method in both precisions (at most 3 timed solves each)")
    ap.add_argument("--precisions", default="mixed,fp64",
                    help="comma list of the runs to make (mixed, fp64); the JSON line needs both")
    ap.add_argument("--summarize-trace", metavar="CSV",
                    help="print a summary of a rocprofv3 kernel-trace or counter CSV and exit")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="derive HBM bytes per launch of the roofline kernel from rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE counter CSVs of a `--precisions mixed` run; print JSON and exit")
    ap.add_argument("--traffic-kernel", default=ROOF_KERNEL,
                    help="kernel (name prefix as rocprofv3 reports it) whose launches --traffic evaluates")
    ap.add_argument("--copy-bytes", type=float, default=NK * BLOCK * 2 * N * N * 8.0,
                    help="bytes each way of the largest device copy in the counter runs (default: the batch's "
                         "finest-level FP32 multigrid vectors)")
    args = ap.parse_args()
    if args.cpu_solve is not None:
        cpu_solve(args.nk, args.cpu_solve, TOL)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.traffic:
        t = counter_traffic(*args.traffic, kernel=args.traffic_kernel)
        f, w = t["FETCH_SIZE"], t["WRITE_SIZE"]
        t["copy_bytes_each_way"] = args.copy_bytes
        t["fetch_scale"] = args.copy_bytes / (f["copy_kib"] * 1024.0)
        t["write_scale"] = args.copy_bytes / (w["copy_kib"] * 1024.0)
        t["bytes_per_launch"] = (f["kernel_kib"] * t["fetch_scale"] + w["kernel_kib"] * t["write_scale"]) * 1024.0
        print(json.dumps(t, indent=1))
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the k-points of a batch shard across GPUs "
              "without communication); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import maxwell_amd as M
    eta = rod_eta(N)
    kx, ky = kpath(args.nk)
    out = {}
    runs = [r.strip() for r in args.precisions.split(",") if r.strip()]
    head = M.MX_METHOD_LOBPCG if args.method == "lobpcg" else M.MX_METHOD_PINVIT
    other = M.MX_METHOD_PINVIT if head == M.MX_METHOD_LOBPCG else M.MX_METHOD_LOBPCG
    # the headline method in both precisions with the full step count and
    # the kernel timings; the other method in both precisions beside it, with
    # at most 3 timed solves
    plan = [("mixed", M.MX_PREC_MIXED, head, args.steps, args.warmup, True),
            ("fp64", M.MX_PREC_F64, head, args.steps, args.warmup, True)]
    if args.other_method:
        plan += [("other_mixed", M.MX_PREC_MIXED, other, min(args.steps, 3), min(args.warmup, 1), False),
                 ("other_fp64", M.MX_PREC_F64, other, min(args.steps, 3), min(args.warmup, 1), False)]
    for name, prec, method, steps, warmup, kernels in plan:
        if name.replace("other_", "") not in runs:
            continue
        times, res, st, kern = run_gpu(M, prec, eta, kx, ky, steps, warmup, args.maxit, bool(args.graphs),
                                       bool(args.autotune and warmup > 0), method, kernels, args.march_rows)
        tot = sum(times)
        out[name] = dict(times=times, res=res, stats=st, kern=kern, method=METHODS[method], steps=steps,
                         value=steps * len(kx) * NEIG / tot, ms=1e3 * tot / steps)
        print("[bench] %s (%s): %.3f s/solve, mean steps %.1f, %s" %
              (name, METHODS[method], tot / steps, res["iters"].mean(),
               {k: round(v["gbs"], 1) for k, v in kern.items()}), file=sys.stderr, flush=True)

    if "mixed" not in out or "fp64" not in out:
        print(json.dumps({k: {"value": v["value"], "kernels": v["kern"]} for k, v in out.items()}, default=float))
        return
    others = {k: v for k, v in out.items() if k.startswith("other_")}
    for k in others:
        del out[k]
    mixed, fp64 = out["mixed"], out["fp64"]
    conv_ok = bool(mixed["res"]["converged"].all() and fp64["res"]["converged"].all())
    lam_m, lam_d = mixed["res"]["lam"][:, :NEIG], fp64["res"]["lam"][:, :NEIG]
    rel = np.abs(lam_m - lam_d) / np.abs(lam_d)
    eig_err = float(rel.max())
    # per eigenvalue index: the largest relative difference over the k-points
    eig_err_each = [float("%.3g" % v) for v in rel.max(axis=0)]
    dom = mixed["kern"]["afw_pass_prolong"]
    peak = 8000.0
    # one-pass compulsory bytes of the solve's fused passes (each vector's b,
    # x in and out, the coarse vectors, eta once): what a pass must move
    E, Cc, c, r = 2.0 * N * N, 1.0 * N * N, 8.0, 4.0
    nvb = len(kx) * BLOCK
    onepass = {"afw_pass_zero_resid": nvb * (2.5 * E * c) + Cc * r,
               "afw_pass_resid": nvb * (3.5 * E * c) + Cc * r,
               "afw_pass_prolong": nvb * (3.25 * E * c) + Cc * r}
    passes = {}
    for k, mode in (("afw_pass_zero_resid", 10), ("afw_pass_resid", 8), ("afw_pass_prolong", 5)):
        v = mixed["kern"][k]
        passes[k] = {"kernel": "k_afw_march<float,%d,2>" % mode, "ms": round(v["ms"], 4),
                     "algorithmic_bytes": v["bytes"], "algorithmic_frac": round(v["gbs"] / peak, 4),
                     "onepass_bytes": onepass[k],
                     "onepass_frac": round(onepass[k] / (v["ms"] * 1e-3) / 1e9 / peak, 4)}
    line = {
        "metric": "converged eigenpairs/s at 512^2, mixed vs FP64 GPU; V-cycle HBM % of peak",
        "value": round(mixed["value"], 3),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex fp32 multigrid + complex fp64 residual/Gram/Ritz",
        "data": "synthetic",
        "config": {"workload": "512x512 rod (eps 8.9, r 0.2), 64 k on Gamma-X-M-Gamma, 10 eigs, "
                               "block 12, tol 1e-8, one batch",
                   "grid": N, "levels": LEVELS + 1, "kpoints": len(kx), "neig": NEIG, "block": BLOCK,
                   "method": mixed["method"],
                   "graphs": bool(args.graphs), "autotune": bool(args.autotune and args.warmup > 0)},
        "fp64_value": round(fp64["value"], 3),
        "mixed_over_fp64": round(mixed["value"] / fp64["value"], 3),
        "eig_rel_diff_mixed_vs_fp64": eig_err,
        "eig_rel_diff_mixed_vs_fp64_each": eig_err_each,
        "converged": conv_ok,
        "mean_steps": {"mixed": float(mixed["res"]["iters"].mean()), "fp64": float(fp64["res"]["iters"].mean())},
        "preconditioned_fraction": {k: round(v["stats"]["active_vectors"] / max(1, BLOCK * v["stats"]["block_steps"]), 3)
                                    for k, v in out.items()},
        "max_relres": {"mixed": float(mixed["res"]["relres"][:, :NEIG].max()),
                       "fp64": float(fp64["res"]["relres"][:, :NEIG].max())},
        "roofline": {"bound": "hbm",
                     "kernel": "k_afw_march<float,5,2>: the solve's way-up pass on the finest level (prolongation "
                               "of the coarse correction + two backward AFW sweeps, one pass), all 768 vectors, run "
                               "twice per block step; algorithmic bytes = prolongation (E c / 4 + 2 E c) + 2 sweeps "
                               "x (3 E c + C r) per vector, c = 8, r = 4",
                     "achieved": round(dom["gbs"], 1), "peak": peak, "unit": "GB/s",
                     "frac": round(dom["gbs"] / peak, 4), "traffic": None,
                     "bytes_per_launch": dom["bytes"], "avg_launch_ms": round(dom["ms"], 4),
                     "onepass_bytes_per_launch": onepass["afw_pass_prolong"],
                     "onepass_frac": passes["afw_pass_prolong"]["onepass_frac"],
                     "device_copy_GBps": round(mixed["kern"]["device_copy"]["gbs"], 1)},
        "afw_passes_of_the_solve": passes,
        "tiling": {k: v["stats"]["tiling"] for k, v in out.items()},
        "vcycle_hbm_frac": {k: round(v["kern"]["edge_vcycle"]["gbs"] / peak, 4) for k, v in out.items()},
        "node_vcycle_hbm_frac": {k: round(v["kern"]["node_vcycle"]["gbs"] / peak, 4) for k, v in out.items()},
        "kernels": {k: {kk: {"ms": round(vv["ms"], 4), "GB/s": round(vv["gbs"], 1)} for kk, vv in v["kern"].items()}
                    for k, v in out.items()},
    }
    if os.path.exists(ROOF_TRAFFIC):
        line["roofline"]["traffic"] = json.load(open(ROOF_TRAFFIC)).get("bytes_per_launch")
    if others:
        om, of = others.get("other_mixed"), others.get("other_fp64")
        line["other_method"] = {
            "method": (om or of)["method"], "timed_solves": (om or of)["steps"],
            "mixed_value": round(om["value"], 3) if om else None,
            "fp64_value": round(of["value"], 3) if of else None,
            "mean_steps": {k.replace("other_", ""): float(v["res"]["iters"].mean()) for k, v in others.items()},
            "converged": bool(all(v["res"]["converged"].all() for v in others.values())),
        }
        if of:
            # the headline (mixed) form against the other method in plain FP64
            line["mixed_over_other_fp64"] = round(mixed["value"] / of["value"], 3)
    line["eig_rel_err_vs_oracle"] = oracle_errors(dict(out, **others), kx, ky)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(len(kx))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
