"""IPM linear-algebra iteration on one MI355X: KKT LDL^T factor + refined
solve to ||Kx-b||/||b|| <= 1e-10 + inertia, on the 2D PDE-constrained QP
(Q1 Poisson distributed control, 256 x 256 mesh, n = 198,147).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

One "step" = kkt_ldl_factor(K + Sigma) + kkt_ldl_solve_refined(rhs) +
kkt_ldl_inertia, with K, Sigma and the right-hand side resident in HBM when
the timed region starts.  Prints one JSON line: the headline value (ms per
step of the FP32-factor + FP64-refinement path when available, else FP64),
both precisions' times, per-kernel attribution of one profiled step with the
roofline of the dominant kernel, and the reference's own SparseLDL timed on
the host cores on the same matrix and ordering.
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("IPM iter time (LDLᵀ factor+solve ms), FP32-MFMA+refine vs FP64 "
          "@1e-10 resid")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_MFMA_PEAK_TFS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mesh", type=int, default=256)
    ap.add_argument("--order", default="nd", choices=["nd", "amd"])
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "fp32", "both"])
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run 5 device-to-device copies of MiB MiB (the known-size "
                         "kernel that scales PMC FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--pmc-probe", choices=["fp64", "fp32"], default=None,
                    help="only run one factor + refined solve in this precision (the "
                         "program the PMC traffic passes profile) and exit")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc passes behind roofline.traffic")
    return ap.parse_args()


def large_front_work(s, P, esize=8):
    """Algorithmic work of the large fronts (nominal dimension > 256), from the
    symbolic structure and the pivots of the last factorization:
    bytes: each eliminated column read once and written once as L (esize
    bytes per front entry: 8 for FP64 fronts, 4 for FP32 fronts);
    flops: the rank-1 Schur updates of the lower triangle."""
    sn = s.get_supernodes()
    info = s.node_info()
    ns = np.diff(sn["node_ptr"])
    cs = np.diff(sn["contrib_ptr"])
    N = len(ns)
    din = np.zeros(N, np.int64)
    ch = sn["children_ptr"]
    kids = sn["children"]
    for k in range(N):
        for q in range(ch[k], ch[k + 1]):
            din[k] += info["ndel"][kids[q]]
    m = ns + din + cs
    big = (ns + cs) > 256
    byt = 0.0
    flops = 0.0
    for k in np.nonzero(big)[0]:
        j = np.arange(info["npiv"][k], dtype=np.float64)
        rem = m[k] - j
        byt += float(np.sum(2.0 * esize * rem))
        flops += float(np.sum((rem - 1.0) * rem))
    return byt, flops


def solve_bytes(s, esize=8):
    st = s.stats()
    n = st["n"]
    # L read in each sweep (esize bytes per entry) + 8 B per solution entry
    # read and written per sweep
    return 2.0 * esize * st["nnz_l"] + 2.0 * 2.0 * 8.0 * n


def run_gpu(P, precision, args, torch):
    from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SparseLDL

    order = OrderingType.NESTED_DISSECTION if args.order == "nd" else OrderingType.AMD
    vals = torch.tensor(P.vals, device="cuda")
    diag = torch.tensor(P.diag, device="cuda")
    rhs = torch.tensor(P.rhs, device="cuda")
    x = torch.empty_like(rhs)
    t0 = time.perf_counter()
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, vals), order=order,
                  precision=FP32_REFINE if precision == "fp32" else FP64,
                  refine_tol=1e-10, max_refine=20)
    t_analyze = time.perf_counter() - t0

    def step():
        s.factor(diag)
        x.copy_(rhs)
        st = s.solve(x, refine=True)
        return st, s.get_inertia()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        status, inertia = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stats = s.stats()
    xh = x.cpu().numpy()
    res = float(np.linalg.norm(P.matrix() @ xh - P.rhs) / np.linalg.norm(P.rhs))
    fac_ms, sol_ms = s.last_times()
    # one profiled step for per-kernel attribution
    s.set_profiling(True)
    s.factor(diag)
    kt_f = s.kernel_times()
    x.copy_(rhs)
    s.solve(x, refine=True)
    kt_s = s.kernel_times()
    s.set_profiling(False)
    kt = {k: (kt_f[k][0] + kt_s[k][0], kt_f[k][1] + kt_s[k][1]) for k in kt_f}
    return dict(s=s, ms=ms, status=status, inertia=inertia, residual=res, precision=precision,
                refine_steps=stats["refine_steps"], factor_ms=fac_ms,
                solve_ms=sol_ms, t_analyze_s=t_analyze, stats=stats, kt=kt)


def bench_host_buffers(P, precision, args, reps=5):
    """The same step with host buffers, as the reference's AmigoSolver holds
    them (numpy values, diagonal and right-hand side): values and Sigma cross
    PCIe at factor time, the right-hand side and the solution at solve time.
    Median ms per step over reps steps after one warm-up."""
    from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SparseLDL

    order = OrderingType.NESTED_DISSECTION if args.order == "nd" else OrderingType.AMD
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals.copy()), order=order,
                  precision=FP32_REFINE if precision == "fp32" else FP64,
                  refine_tol=1e-10, max_refine=20)
    x = np.empty(P.n)
    tt = []
    for k in range(reps + 1):
        t0 = time.perf_counter()
        s.factor(P.diag)
        x[:] = P.rhs
        s.solve(x, refine=True)
        s.get_inertia()
        tt.append((time.perf_counter() - t0) * 1e3)
    res = float(np.linalg.norm(P.matrix() @ x - P.rhs) / np.linalg.norm(P.rhs))
    return dict(precision=precision, ms_per_step=round(float(np.median(tt[1:])), 3),
                pcie_bytes_per_step=int(8 * P.nnz + 8 * P.n + 16 * P.n), residual=res)


def bench_multi_rhs(r, P, torch, nrhs=64, reps=3):
    """The post-optimality adjoint shape: nrhs seeded right-hand sides
    (column major, zero on the fixed rows) solved in one refined call on the
    stored factors, against nrhs refined single solves; the worst relative
    residual over the columns."""
    s = r["s"]
    rng = np.random.default_rng(7)
    B = rng.standard_normal((P.n, nrhs))
    B[P.rhs == 0.0, :] = 0.0
    Bt = torch.tensor(B.T.copy(), device="cuda").t()  # column major n x nrhs
    X = torch.empty_like(Bt)
    x1 = torch.empty(P.n, dtype=torch.float64, device="cuda")

    def multi():
        X.copy_(Bt)
        s.solve(X, refine=True)

    def single():
        for c in range(nrhs):
            x1.copy_(Bt[:, c])
            s.solve(x1, refine=True)

    out = {}
    for name, fn in (("multi", multi), ("single", single)):
        fn()
        torch.cuda.synchronize()
        tt = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            tt.append((time.perf_counter() - t0) * 1e3)
        out[name + "_ms"] = round(float(np.median(tt)), 3)
    multi()
    torch.cuda.synchronize()
    K = P.matrix()
    Xh = X.cpu().numpy()
    rel = np.linalg.norm(K @ Xh - B, axis=0) / np.linalg.norm(B, axis=0)
    out.update(nrhs=nrhs, ms_per_rhs=round(out["multi_ms"] / nrhs, 4),
               speedup_vs_single=round(out["single_ms"] / out["multi_ms"], 2),
               worst_residual=float(rel.max()), refine_steps=s.stats()["refine_steps"])
    return out


def roofline(r, P):
    kt = r["kt"]
    dom = max(kt, key=lambda k: kt[k][0])
    t_ms, nl = kt[dom]
    s = r["s"]
    esize = 4 if r["precision"] == "fp32" else 8
    if dom == "update":
        _, flops = large_front_work(s, P)
        ach = flops / (t_ms * 1e-3) / 1e12
        return dict(kernel="k_update", bound="mfma", achieved=ach, peak=FP64_PEAK_TFS,
                    unit="TFLOP/s", frac=ach / FP64_PEAK_TFS, traffic=None,
                    launches=nl, avg_launch_us=t_ms * 1e3 / max(nl, 1),
                    algorithmic_per_launch=flops / max(nl, 1))
    if dom in ("forward", "backward"):
        byt = solve_bytes(s, esize) / 2.0 * (1 + r["refine_steps"])
        ach = byt / (t_ms * 1e-3) / 1e9
        return dict(kernel="k_" + dom, bound="hbm", achieved=ach, peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=ach / HBM_PEAK_GBS, traffic=None, launches=nl,
                    avg_launch_us=t_ms * 1e3 / max(nl, 1),
                    algorithmic_per_launch=byt / max(nl, 1))
    byt, _ = large_front_work(s, P, esize)
    if dom != "panel":
        byt = 2.0 * esize * s.stats()["nnz_l"]
    ach = byt / (t_ms * 1e-3) / 1e9
    # large-front panels: k_panel_rr (the panel's L in LDS) for every level
    # whose fronts fit it -- all of them in the FP32 path of this workload --
    # k_panel above that
    name = {"panel": "k_panel_rr"}.get(dom, "k_" + dom)
    return dict(kernel=name, bound="hbm",
                achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                traffic=None, launches=nl, avg_launch_us=t_ms * 1e3 / max(nl, 1),
                algorithmic_per_launch=byt / max(nl, 1))


def _pmc_pass(counter, child_args, timeout=180):
    """Per-dispatch values of one PMC counter for this script run with
    child_args under rocprofv3 (a child process), as {kernel name: [KiB]}."""
    import csv
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if prof is None:
        raise RuntimeError("rocprofv3 not found")
    d = tempfile.mkdtemp(prefix="kkt_pmc_")
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    subprocess.run(cmd, check=True, timeout=timeout, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, env=dict(os.environ, TMPDIR="/tmp"), cwd="/tmp")
    out = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    shutil.rmtree(d, ignore_errors=True)
    return out


def pmc_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE) in one
    factor + refined solve of the headline precision, each counter scaled by
    what it reads for device copies of known size (5 x 256 MiB)."""
    copy_bytes = 256 << 20
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = _pmc_pass(counter, ["--copy-calibration", "256"])
        cv = [v for k, vs in cal.items() if "copyBuffer" in k for v in vs]
        if not cv:
            raise RuntimeError("no copy kernel in the calibration pass")
        scale = copy_bytes / 1024.0 / (sum(cv) / len(cv))
        run = _pmc_pass(counter, ["--pmc-probe", precision, "--order", "nd"])
        pat = re.compile(r"\b" + re.escape(kernel) + r"[<(]")
        kv = [v for k, vs in run.items() if pat.search(k) for v in vs]
        if not kv:
            raise RuntimeError(f"no {kernel} dispatch in the PMC pass")
        res[counter] = dict(kib_per_launch=sum(kv) / len(kv), scale=scale, launches=len(kv))
    byt = sum(1024.0 * v["kib_per_launch"] * v["scale"] for v in res.values())
    return byt, res


def bench_assembly(args, torch, reps=50, cpu=True):
    """KKT value assembly of the same KKT from its Q1 elements (N^2 elements
    x 12 variables, 144 contributions each) with the boundary rows fixed:
    device time of k_gather_sum over reps launches (HIP events on torch's
    current stream, the one the launch uses), and the reference's own host
    assembly (CSRMat::add_row per element column + zero_rows_and_columns)
    on the same contributions."""
    from amigo_amd import workloads as W
    from amigo_amd.assembly import KKTAssembler, element_contributions, element_values

    E = W.poisson2d_elements(N=args.mesh)
    A = KKTAssembler.from_elements(E["n"], E["rowp"], E["cols"], [E["idx"]], E["fixed"])
    cvh = element_values(E["H"])
    cv = torch.tensor(cvh, device="cuda")
    out = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
    for _ in range(3):
        A.values(cv, out=out)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        A.values(cv, out=out)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    kept = A.ncontrib - A.ndropped
    # per kept contribution: 8 B value + 4 B gather index; per CSR entry:
    # 4 B segment pointer + 1 B mode + 8 B value written
    byt = 12.0 * kept + 13.0 * A.nnz + 4.0
    ach = byt / (ms * 1e-3) / 1e9
    res = dict(kernel="k_gather_sum", ms_per_launch=round(ms, 5), launches=reps,
               ncontrib=int(A.ncontrib), ndropped=int(A.ndropped), nnz=int(A.nnz),
               algorithmic_bytes=byt, achieved_GBs=round(ach, 1),
               frac=round(ach / HBM_PEAK_GBS, 4),
               bitwise_vs_snapshot=bool(np.array_equal(out.cpu().numpy(), E["vals"])))
    if cpu:
        from oracle import pyoracle as po
        if po.ref_available():
            r, c = element_contributions(E["idx"])
            po.ref_assemble(E["n"], E["rowp"], E["cols"], r, c, cvh, E["fixed"], ncomp=12)
            tc = []
            for _ in range(3):
                t = time.perf_counter()
                po.ref_assemble(E["n"], E["rowp"], E["cols"], r, c, cvh, E["fixed"], ncomp=12)
                tc.append((time.perf_counter() - t) * 1e3)
            res["cpu_reference_ms"] = round(float(np.median(tc)), 3)
            res["cpu_reference_sample"] = ("reference host assembly (serial add_row + fixed "
                                           "rows), median of 3, 1 core")
    return res


def bench_condense(args, torch, reps=50, cpu=True):
    """The IPM condensation around the solve on the same problem's iterate
    (mesh 256: 198,147 entries, 131,074 primals of which the 66,049 controls
    are bounded, 65,025 constraints): the barrier
    diagonal Sigma (k_sigma) and the condensed right-hand side with its norm
    (k_residual + final reduction), device time per call over reps calls
    (HIP events on torch's current stream, the one the launches use), and
    the oracle's numpy restatement of the reference's loops on the host."""
    from amigo_amd import workloads as W
    from amigo_amd.ipm import InteriorPointKernels

    S = W.poisson2d_ipm_state(N=args.mesh)
    K = InteriorPointKernels(S["n"], S["pidx"], S["lbx"], S["ubx"], S["cidx"], S["lbh"])
    d = {k: torch.tensor(S[k], device="cuda") for k in ("x", "zl", "zu", "grad")}
    st = (d["x"], d["zl"], d["zu"])
    diag = torch.empty(S["n"], dtype=torch.float64, device="cuda")
    res = torch.empty_like(diag)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for _ in range(3):
        K.compute_diagonal(st, diag)
        K.compute_residual(S["mu"], st, d["grad"], res)
    ev[0].record()
    for _ in range(reps):
        K.compute_diagonal(st, diag)
    ev[1].record()
    ev[2].record()
    for _ in range(reps):
        K.compute_residual(S["mu"], st, d["grad"], res)
    ev[3].record()
    torch.cuda.synchronize()
    ms_sig = ev[0].elapsed_time(ev[1]) / reps
    ms_res = ev[2].elapsed_time(ev[3]) / reps
    n, npr = S["n"], len(S["pidx"])
    # per entry: the two index maps (8 B) and the output (8 B); per primal:
    # x, lb, ub, zl, zu (40 B); the residual also reads grad (8 B / entry)
    # and the constraint targets (8 B / constraint)
    b_sig = 16.0 * n + 40.0 * npr
    b_res = b_sig + 8.0 * n + 8.0 * len(S["cidx"])
    out = dict(kernels=["k_sigma", "k_residual"], ms_sigma=round(ms_sig, 5),
               ms_residual=round(ms_res, 5), calls=reps, n=n, primals=npr,
               algorithmic_bytes_sigma=b_sig, algorithmic_bytes_residual=b_res,
               sigma_GBs=round(b_sig / (ms_sig * 1e-3) / 1e9, 1),
               residual_GBs_incl_sync=round(b_res / (ms_res * 1e-3) / 1e9, 1))
    if cpu:
        from oracle import ipm_condense as O
        tc = []
        for _ in range(3):
            t = time.perf_counter()
            O.compute_diagonal(n, S["pidx"], S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"])
            O.compute_residual(n, S["pidx"], S["lbx"], S["ubx"], S["cidx"], S["lbh"], S["mu"],
                               S["x"], S["zl"], S["zu"], S["grad"])
            tc.append((time.perf_counter() - t) * 1e3)
        out["cpu_oracle_ms"] = round(float(np.median(tc)), 3)
        out["cpu_oracle_sample"] = "numpy restatement of Sigma + residual, median of 3, 1 core"
    return out


def cpu_baseline(P, perm, reps):
    """The reference's SparseLDL (oracle/_ref, compiled from its own headers)
    on the same KKT with the same elimination order (the matrix permuted by
    the engine's ordering, analysed with NATURAL), else the oracle."""
    import scipy.sparse as sp
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import pyoracle as po

    K = sp.csr_matrix((P.vals, P.cols, P.rowp), shape=(P.n, P.n))
    Kp = K[perm][:, perm].tocsr()
    Kp.sort_indices()
    dp = P.diag[perm]
    bp = P.rhs[perm]
    kind = "reference" if po.ref_available() else "port"
    if kind == "reference":
        solver = po.RefLDL(P.n, Kp.indptr, Kp.indices, order=po.NATURAL)
    else:
        solver = po.OracleLDL(P.n, Kp.indptr, Kp.indices, order=po.NATURAL)
    solver.factor(Kp.data, dp)
    solver.solve(bp)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        solver.factor(Kp.data, dp)
        solver.solve(bp)
        times.append((time.perf_counter() - t0) * 1e3)
    inertia = solver.inertia()
    return dict(value=float(np.median(times)), unit="ms", cores=threads, kind=kind,
                sample=(f"{reps} factor+solve runs (median) after 1 warm-up of the same "
                        f"n={P.n} KKT and elimination order, OMP threads={threads}"),
                inertia=list(inertia))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the factorization is single-GPU by design)", file=sys.stderr)
        sys.exit(2)
    import torch

    from amigo_amd import workloads as W

    assert torch.cuda.is_available(), "needs an MI355X (cuda:0)"
    if args.copy_calibration:
        a = torch.ones(args.copy_calibration << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 8, "copies": 5}))
        return
    if args.pmc_probe:
        from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SparseLDL
        P = W.poisson2d(N=args.mesh)
        order = OrderingType.NESTED_DISSECTION if args.order == "nd" else OrderingType.AMD
        s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, torch.tensor(P.vals, device="cuda")),
                      order=order, precision=FP32_REFINE if args.pmc_probe == "fp32" else FP64)
        s.factor(torch.tensor(P.diag, device="cuda"))
        x = torch.tensor(P.rhs, device="cuda")
        s.solve(x, refine=True)
        torch.cuda.synchronize()
        return
    P = W.poisson2d(N=args.mesh)
    precisions = {"auto": ["fp64", "fp32"], "both": ["fp64", "fp32"],
                  "fp64": ["fp64"], "fp32": ["fp32"]}[args.precision]
    results = {}
    for prec in precisions:
        try:
            results[prec] = run_gpu(P, prec, args, torch)
        except RuntimeError as e:
            if args.precision == "auto" and prec == "fp32":
                print(f"fp32 path unavailable: {e}", file=sys.stderr)
                continue
            raise
    head = "fp32" if "fp32" in results else "fp64"
    r = results[head]
    asm = bench_assembly(args, torch, cpu=not args.no_cpu_baseline)
    cond = bench_condense(args, torch, cpu=not args.no_cpu_baseline)
    mrhs = {p: bench_multi_rhs(v, P, torch) for p, v in results.items()}
    host = bench_host_buffers(P, head, args)
    rl = roofline(r, P)
    if not args.no_traffic:
        try:
            t, detail = pmc_traffic(rl["kernel"], head)
            rl["traffic"] = t
            rl["traffic_over_algorithmic"] = t / rl["algorithmic_per_launch"]
            rl["traffic_pmc"] = detail
        except Exception as e:  # the counters are a diagnostic, not the measurement
            print(f"PMC traffic unavailable: {e}", file=sys.stderr)
    perm = r["s"].get_perm()[0]
    cpu = None if args.no_cpu_baseline else cpu_baseline(P, perm, args.cpu_reps)
    out = {
        "metric": METRIC,
        "value": round(r["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refine" if head == "fp32" else "f64",
        "data": "synthetic (seeded KKT snapshot: y_d = 8 Gaussians, late-IPM Sigma)",
        "config": {
            "workload": "2D Poisson distributed-control QP, 256x256 Q1 mesh, KKT "
                        "n=198147, nested dissection, alpha=1e-4, -1<=u<=1",
            "n": P.n, "nnz": P.nnz, "order": args.order,
            "num_nodes": r["stats"]["num_nodes"], "num_levels": r["stats"]["num_levels"],
            "max_front": r["stats"]["max_front"],
            "flops_factor": r["stats"]["flops_factor"], "nnz_l": r["stats"]["nnz_l"],
        },
        "precisions": {
            p: {"ms_per_step": round(v["ms"], 3), "factor_ms": round(v["factor_ms"], 3),
                "solve_refined_ms": round(v["solve_ms"], 3),
                "refine_steps": v["refine_steps"], "residual": v["residual"],
                "inertia": list(v["inertia"]), "status": v["status"],
                "analyze_s": round(v["t_analyze_s"], 3)}
            for p, v in results.items()
        },
        "kernels": {k: {"ms": round(t, 3), "launches": n} for k, (t, n) in r["kt"].items()},
        "roofline": rl,
        "assembly": asm,
        "condensation": cond,
        "multi_rhs_64": mrhs,
        "host_buffers": host,
        "cpu_baseline": cpu,
    }
    if cpu and cpu.get("inertia"):
        out["inertia_matches_reference"] = cpu["inertia"] == list(r["inertia"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
