"""Headline benchmark of the matrix-free IGA Poisson operator on one GPU.

Headline (the JSON ``value``): the FP32 sum-factorised apply y = A x on the
quarter annulus r in [1,2], exact NURBS, p = 3, 1024 x 1024 elements, 4 x 4
Gauss points (1,050,625 interior unknowns), x ~ N(0,1) (splitmix64 seed 0),
operator and vectors resident in HBM; a step is one apply (inverse-LM gather
assembly).  Alongside, in the same JSON line:

  solve     CG to ||f - A u|| / ||f|| <= 1e-8, f = 1: plain FP64 PCG against
            FP32 PCG with FP64 reliable updates (defect correction), and the
            same with the inner FP32 operator on FP16-stored G_q, Jacobi,
            both assemblies, the three iteration forms (separate p.Ap kernel,
            p.Ap inside the apply, single-reduction iteration);
  block     multi-RHS apply, quarter annulus p = 4, 512 x 512, k = 32
            (autotuned over k in {8, 16, 32}) and the block solve on the load
            vectors of the k Fourier right-hand sides;
  c5        unit square p = 5, 2048 x 2048, perturbed interior points and
            per-function weights U[0.8, 1.2]: FP32 SF and FP64 EBE applies and
            the GPU FEM-preconditioner assembly, its indices compared bit for
            bit with the CPU's;
  solve_fem_mg  the same solve preconditioned by the reference's Q1 FEM
            matrix, applied as one multigrid V-cycle;
  c1        the plain method (p = 2, 32^2, FP64 EBE + FEM preconditioner);
  fempc_iterations  the FEM-preconditioned outer iteration counts at p = 3,
            32^2 / 64^2 / 128^2 beside the reference's logged ones;
  c4        the kappa adjoint (p = 3, 256^2, Fourier kappa of seed 1, u_obs
            of seed 2 plus noise) with the gradient against central
            differences along 16 directions;
  cpu_baseline  the oracle's OpenMP restatement of the reference apply on the
            host cores (all cores and one core), the C1 CPU solve and the C5
            CPU FEM assembly.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-solve] ...
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-computing_amd"))

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, vendor figure
FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 (vector with packed math / MFMA)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_apply_f32_annulus_p3_1024.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nelem", type=int, default=1024)
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-block", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the C1 and C4 checks")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-applies", type=int, default=10)
    ap.add_argument("--assembly", choices=["gather", "atomic"], default="gather")
    ap.add_argument("--kernel", choices=["sf", "ebe"], default="sf")
    ap.add_argument("--variant-f32", type=int, default=-1,
                    help="FP32 apply variant of the operator (-1: the library's default)")
    ap.add_argument("--detail", action="store_true", help="also time every apply variant")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 256 MiB device copy first (PMC byte-counter calibration)")
    ap.add_argument("--only-apply", action="store_true",
                    help="headline apply only (for profiler runs)")
    ap.add_argument("--also-time", default="",
                    help="with --only-apply: also run the element kernel of these forms "
                         "back to back (comma-separated: g16, f64), for profiler runs")
    ap.add_argument("--solve-configs", default="",
                    help="comma-separated solve keys to run (e.g. fp64_atomic_fused1); "
                         "default all")
    ap.add_argument("--solve-reps", type=int, default=3, help="timed solves per config")
    ap.add_argument("--no-block-solve", action="store_true",
                    help="block section: the applies only (for profiler runs)")
    ap.add_argument("--only", default="",
                    help="after the headline, run only these sections (comma list of "
                         "variants, solve, mg, cpu, block, c5, extra)")
    return ap.parse_args()


def median(v):
    v = sorted(v)
    return v[len(v) // 2]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ev_time(fn, steps, warmup):
    """Median and mean of per-call HIP-event times (ms) on the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    per = [a.elapsed_time(b) for a, b in ev]
    return median(per), sum(per) / len(per)


def settle(fn, seconds=0.3):
    """Untimed calls of fn for about `seconds`, so that the timing that
    follows sees the device's steady clocks: after seconds of host-side work
    (mesh generation, set-up) the device has idled and its clocks have
    dropped, and the first few hundred calls of a hot loop run several per
    cent slower."""
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            fn()
        torch.cuda.synchronize()


def stream_rates():
    """Practical HBM rates of plain torch kernels on this device, as context
    for the apply's fraction of the published peak: a sum over 1 GiB of FP32
    (read-only) and a 512 MiB device copy (read + write), GB/s, best of 10
    event-timed calls."""
    import torch
    x = torch.ones(256 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty(128 << 20, dtype=torch.float32, device="cuda")
    out = {}
    for name, fn, nbytes in (("read_gbs", lambda: x.sum(), x.numel() * 4),
                             ("copy_gbs", lambda: y.copy_(x[:y.numel()]), 2 * y.numel() * 4)):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(10):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            best = min(best, a.elapsed_time(b))
        out[name] = round(nbytes / (best * 1e-3) / 1e9, 1)
    del x, y
    torch.cuda.empty_cache()
    return out


def kernel_avg_ms(op, x, y, steps, kern, asm):
    """Average duration of the apply kernel itself: HIP events recorded by
    the library around each k_apply launch on its stream."""
    import torch
    op.profile(True, steps)
    for _ in range(steps):
        op.apply(x, y, kernel=kern, assembly=asm)
    torch.cuda.synchronize()
    cnt, tot, kmin, kmax = op.profile_read()
    op.profile(False)
    return tot / max(cnt, 1), kmin


def headline(a, out):
    import torch
    import igamf
    p, n = a.p, a.nelem
    t0 = time.perf_counter()
    mesh = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, p, n, r0=1.0, r1=2.0)
    op = igamf.Operator(mesh, f16_geometry=True, variant_f32=a.variant_f32)
    setup_s = time.perf_counter() - t0
    info = op.info
    nfunc = op.nfunc
    n_int = (info.nfx - 2) * (info.nfy - 2)
    E, nqp = op.nelem, op.nqp
    kern = igamf.SF if a.kernel == "sf" else igamf.EBE
    asm = igamf.INVLM_GATHER if a.assembly == "gather" else igamf.IEN_ATOMIC
    x64 = torch.from_numpy(igamf.rng_normal(0, nfunc)).cuda()
    x32 = x64.float()
    if a.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    y32 = torch.empty_like(x32)
    # 0.3 s of untimed applies before the W warmup steps, so that the timed
    # steps see the device's steady clocks (the first 200 applies of a fresh
    # process otherwise ran about 7 % slower per step)
    settle(lambda: op.apply(x32, y32, kernel=kern, assembly=asm))
    # the contract's timing: W untimed steps, then K steps between two
    # synchronisations, wall clock
    for _ in range(a.warmup):
        op.apply(x32, y32, kernel=kern, assembly=asm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        op.apply(x32, y32, kernel=kern, assembly=asm)
    torch.cuda.synchronize()
    wall32 = (time.perf_counter() - t0) * 1e3 / a.steps
    med32, mean32 = ev_time(lambda: op.apply(x32, y32, kernel=kern, assembly=asm),
                            a.steps, a.warmup)
    k_avg, k_min = kernel_avg_ms(op, x32, y32, a.steps, kern, asm)
    # the element kernel alone, back to back between two events (no event
    # per launch): the figure the roofline uses
    k_b2b = op.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm)
    op.apply(x32, y32, kernel=kern, assembly=asm)
    alg_bytes32 = 4 * (2 * nfunc) + 4 * 3 * E * nqp
    achieved = alg_bytes32 / (k_b2b * 1e-3) / 1e9
    gdofs = n_int / (wall32 * 1e-3) / 1e9
    # (not in byte-counter runs: its copy would be taken for the calibration copy)
    practical = None if a.calibrate_copy else stream_rates()
    variant = info.variant_f32
    tx, ty = info.tile_x_f32, info.tile_y_f32
    if variant == 1 and a.kernel == "sf":
        kname = f"k_apply_pf<float, {p}, {p + 1}, {tx}, {ty}>"
    elif variant == 5 and a.kernel == "sf":
        kname = f"k_apply_lds<float, {p}, {p + 1}, float>"
    else:
        split = 2 if (variant == 3 and a.kernel == "sf") else 1
        # the split-element kernel on thread-major geometry (16x8, even point
        # count) is built for three waves per SIMD
        minw = 4 if (variant == 2 and a.kernel == "sf") else (
            3 if (split == 2 and (tx, ty) == (16, 8) and (p + 1) % 2 == 0) else 1)
        ebe = "true" if a.kernel == "ebe" else "false"
        # last argument: G_q loaded non-temporally (large factors, split kernels)
        ntg = "true" if (split == 2 and info.stream_g_f32) else "false"
        kname = (f"k_apply<float, {p}, {p + 1}, {tx}, {ty}, {ebe}, false, {minw}, {split}, "
                 f"float, {ntg}>")
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if tr.get("kernel") == kname and tr.get("nelem") == n and tr.get("p") == p:
            traffic = tr.get("bytes_per_launch")
    out.update({
        "metric": "MF apply GDoF/s & CG solve-to-1e-8 time: FP32+refine vs FP64, p=3 1024² elems",
        "value": round(gdofs, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall32, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: x ~ N(0,1) splitmix64 seed 0; exact NURBS quarter annulus generated "
                "on the host (no checkpoint or dataset)",
        "config": {"workload": f"quarter annulus r in [1,2], p={p}, {n}x{n} elements, "
                               f"{p + 1}x{p + 1} Gauss, FP32 sum-factorised apply, "
                               f"{a.assembly} assembly",
                   "p": p, "nelem": n, "interior_dofs": n_int, "control_points": nfunc,
                   "kernel": a.kernel, "assembly": a.assembly,
                   "tile_f32": [info.tile_x_f32, info.tile_y_f32], "variant_f32": variant,
                   "tile_f64": [info.tile_x_f64, info.tile_y_f64],
                   "variant_f64": info.variant_f64},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": kname,
                     "kernel_avg_ms": round(k_b2b, 5),
                     "kernel_avg_ms_event_per_launch": round(k_avg, 5),
                     "kernel_min_ms_event_per_launch": round(k_min, 5),
                     "algorithmic_bytes": alg_bytes32,
                     "practical_torch": practical},
        "apply": {"f32_wall_ms_per_step": round(wall32, 5),
                  "f32_event_median_ms": round(med32, 5), "f32_event_mean_ms": round(mean32, 5),
                  "setup_s": round(setup_s, 3), "op_setup_ms": round(info.setup_ms, 2),
                  "device_bytes": info.device_bytes},
    })
    log(f"headline f32 apply {wall32 * 1e3:.1f} us/step, kernel {k_b2b * 1e3:.1f} us "
        f"({k_avg * 1e3:.1f} us with an event pair per launch), "
        f"{achieved / HBM_PEAK_GBS:.3f} of HBM" +
        (f"; torch sum {practical['read_gbs']:.0f} GB/s, copy {practical['copy_gbs']:.0f} GB/s"
         if practical else ""))
    return mesh, op, x64, x32, kern, asm, n_int


def fp64_and_variants(a, out, op, x64, x32, kern, asm, n_int, nfunc, E, nqp):
    import torch
    import igamf
    y64 = torch.empty_like(x64)
    med64, _ = ev_time(lambda: op.apply(x64, y64, kernel=kern, assembly=asm), a.steps, a.warmup)
    alg_bytes64 = 8 * (2 * nfunc) + 8 * 3 * E * nqp
    out["apply"]["f64_median_ms"] = round(med64, 5)
    out["apply"]["f64_gdofs"] = round(n_int / (med64 * 1e-3) / 1e9, 4)
    out["apply"]["f64_hbm_frac"] = round(alg_bytes64 / (med64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    k64 = op.time_kernel(x64, torch.empty_like(x64), a.steps, kern, asm)
    out["apply"]["f64_kernel_ms"] = round(k64, 5)
    out["apply"]["f64_kernel_hbm_frac"] = round(alg_bytes64 / (k64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    # FP32 arithmetic on the FP16 copy of G_q (the inner operator of the
    # FP16-geometry refinement solve)
    if kern == igamf.SF and op.info.f16_geometry:
        y16 = torch.empty_like(x32)
        med16, _ = ev_time(lambda: op.apply(x32, y16, kernel=kern, assembly=asm, g16=True),
                           a.steps, a.warmup)
        k16 = op.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm, g16=True)
        alg_bytes16 = 4 * (2 * nfunc) + 2 * 3 * E * nqp
        y64r = op.apply(x64, kernel=kern, assembly=asm)
        op.apply(x32, y16, kernel=kern, assembly=asm, g16=True)
        out["apply"]["g16"] = {
            "wall_event_median_ms": round(med16, 5), "kernel_ms": round(k16, 5),
            "gdofs": round(n_int / (med16 * 1e-3) / 1e9, 4),
            "algorithmic_bytes": alg_bytes16,
            "kernel_hbm_frac": round(alg_bytes16 / (k16 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "max_rel_diff_vs_f64": ((y16.double() - y64r).abs().max()
                                    / y64r.abs().max()).item(),
            "f16_scale": op.info.f16_scale}
        log(f"g16 apply kernel {k16 * 1e3:.1f} us")
    # the matrix-core form of the FP32 kernel (apply variant 4, degree 3)
    if kern == igamf.SF and op.p == 3 and op.info.rational != 2:
        op4 = igamf.Operator(op.mesh, variant_f32=4, f16_geometry=True)
        k4 = op4.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm)
        k4g = op4.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm, g16=True)
        alg_bytes32 = 4 * (2 * nfunc) + 4 * 3 * E * nqp
        out["apply"]["matrix_core"] = {
            "kernel_ms": round(k4, 5),
            "kernel_hbm_frac": round(alg_bytes32 / (k4 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "g16_kernel_ms": round(k4g, 5)}
        del op4
        log(f"matrix-core apply kernel {k4 * 1e3:.1f} us (FP16 G_q {k4g * 1e3:.1f} us)")
    # the LDS-staged persistent form (apply variant 5, degrees 1-3), unless
    # it is the headline's own
    if kern == igamf.SF and op.p <= 3 and op.info.rational != 2 and op.info.variant_f32 != 5:
        op5 = igamf.Operator(op.mesh, variant_f32=5, f16_geometry=op.p == 3)
        k5 = op5.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm)
        alg_bytes32 = 4 * (2 * nfunc) + 4 * 3 * E * nqp
        out["apply"]["lds_staged"] = {
            "kernel_ms": round(k5, 5),
            "kernel_hbm_frac": round(alg_bytes32 / (k5 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if op.p == 3:
            out["apply"]["lds_staged"]["g16_kernel_ms"] = round(
                op5.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm, g16=True), 5)
        del op5
        log(f"LDS-staged apply kernel {k5 * 1e3:.1f} us")
    if a.detail:
        var = {}
        for kn, kv in (("sf", igamf.SF), ("ebe", igamf.EBE)):
            for an, av in (("gather", igamf.INVLM_GATHER), ("atomic", igamf.IEN_ATOMIC)):
                for dn, xx in (("f32", x32), ("f64", x64)):
                    yy = torch.empty_like(xx)
                    m, _ = ev_time(lambda: op.apply(xx, yy, kernel=kv, assembly=av),
                                   max(10, a.steps // 4), 3)
                    var[f"{kn}_{an}_{dn}_ms"] = round(m, 5)
        out["apply"]["variants"] = var


def solves(out, op, only=(), reps=3):
    """Plain FP64 PCG against FP32 PCG + FP64 reliable updates, Jacobi, f = 1,
    to 1e-8 relative residual (FP64 residual check), for both assemblies and
    the three iteration forms (fused = 0: separate p.Ap kernel; 1: p.Ap inside
    the apply kernel; 2: single-reduction iteration)."""
    import torch
    import igamf
    F = op.load(igamf.LOAD_CONST, 1.0)
    res = {}
    sols = {}
    cases = [("fp64", igamf.SOLVE_PCG_F64, 0), ("fp32_refine", igamf.SOLVE_PCG_F32_REFINE, 0)]
    if op.info.f16_geometry:
        cases.append(("fp32g16_refine", igamf.SOLVE_PCG_F32_REFINE, 1))
    for an, av in (("gather", igamf.INVLM_GATHER), ("atomic", igamf.IEN_ATOMIC)):
        for fused in (0, 1, 2):
            for name, kind, g16 in cases:
                key = f"{name}_{an}_fused{fused}"
                if only and key not in only:
                    continue
                s = igamf.Solver(op, None, kind=kind, pc=igamf.PC_JACOBI, rtol=1e-8,
                                 assembly=av, maxit=200000, fused=fused, inner_f16_geometry=g16)
                s.solve(F)  # first call captures the CG chunk graph
                ts = []
                for _ in range(reps):
                    u, st = s.solve(F)
                    ts.append(st.solve_ms)
                torch.cuda.synchronize()
                res[key] = {"solve_ms": round(median(ts), 2), "iterations": st.iterations,
                            "reliable_updates": st.outer_iterations,
                            "us_per_iteration": round(median(ts) * 1e3 / st.iterations, 1),
                            "rel_residual": st.rel_residual, "status": st.status}
                sols[key] = u
                log(f"solve {key}: {median(ts):.1f} ms, {st.iterations} its")
    if only:
        out["solve"] = res
        return
    ref = sols["fp64_gather_fused0"]
    for k in sols:
        if k != "fp64_gather_fused0":
            res[k]["solution_rel_diff"] = (sols[k] - ref).norm().item() / ref.norm().item()
    best = {}
    for name, _, _ in cases:
        k = min((k for k in res if k.startswith(name + "_")), key=lambda k: res[k]["solve_ms"])
        best[name] = k
        res[f"best_{name}"] = k
    classic64 = min(res["fp64_gather_fused0"]["solve_ms"], res["fp64_atomic_fused0"]["solve_ms"])
    for name, _, _ in cases[1:]:
        res[f"speedup_best_{name}_vs_best_fp64"] = round(
            res[best["fp64"]]["solve_ms"] / res[best[name]]["solve_ms"], 3)
        # against the iteration with a separate p.Ap kernel (the classic form)
        res[f"speedup_best_{name}_vs_classic_fp64"] = round(
            classic64 / res[best[name]]["solve_ms"], 3)
    res["preconditioner"] = "jacobi"
    res["inner_rtol"] = "library default (automatic: 0.1 with Jacobi)"
    out["solve"] = res


def c3_coeffs(j):
    """a_mn ~ N(0,1) of right-hand side j (splitmix64 seed j), a[m-1][n-1]."""
    import igamf
    return igamf.rng_normal(j, 64).reshape(8, 8)


def c3_greville_columns(p, ne, k):
    """f_j = sum_{m,n<=8} a_mn sin(m pi xi) sin(n pi eta) at the Greville
    abscissae of the open uniform knots (row eta, column xi)."""
    import numpy as np
    nf = ne + p
    knots = np.concatenate([np.zeros(p), np.linspace(0, 1, ne + 1), np.ones(p)])
    gr = np.array([knots[i + 1:i + p + 1].mean() for i in range(nf)])
    S = np.sin(np.pi * np.outer(gr, np.arange(1, 9)))
    return np.stack([(S @ c3_coeffs(j) @ S.T).T.reshape(-1) for j in range(k)])


def c3_qp_values(p, ne, j):
    """f_j at the quadrature points' parametric coordinates, [E][nqp] in the
    reference's element and point order (the input of iga_op_load_qp)."""
    import numpy as np
    nq = p + 1
    g = (np.polynomial.legendre.leggauss(nq)[0] + 1) / 2
    xi = ((np.arange(ne)[:, None] + g[None, :]) / ne).reshape(-1)
    S = np.sin(np.pi * np.outer(xi, np.arange(1, 9)))
    T = S @ c3_coeffs(j) @ S.T
    return T.reshape(ne, nq, ne, nq).transpose(2, 0, 3, 1).reshape(-1)


def solves_mg(out, mesh, op, reps=3):
    """The same C2 solve (f = 1, rtol 1e-8) with the reference's
    preconditioner, the Q1 FEM matrix on the reference's node grid mapped
    onto the annulus, applied as one multigrid V-cycle (IGA_PC_FEM_MG): plain
    FP64 PCG against FP32 PCG + FP64 reliable updates, both assemblies, p.Ap
    by its own kernel (fused 0) or inside the apply (fused 1)."""
    import torch
    import igamf
    t0 = time.perf_counter()
    nodes = mesh.fem_nodes()
    fem = igamf.FemPC(op.info.nfx, op.info.nfy, nodes)
    torch.cuda.synchronize()
    t_fem = time.perf_counter() - t0
    F = op.load(igamf.LOAD_CONST, 1.0)
    res, sols = {}, {}
    for an, av in (("gather", igamf.INVLM_GATHER), ("atomic", igamf.IEN_ATOMIC)):
        for fused in (0, 1):
            for name, kind in (("fp64", igamf.SOLVE_PCG_F64),
                               ("fp32_refine", igamf.SOLVE_PCG_F32_REFINE)):
                key = f"{name}_{an}_fused{fused}"
                t0 = time.perf_counter()
                s = igamf.Solver(op, fem, kind=kind, pc=igamf.PC_FEM_MG, rtol=1e-8,
                                 assembly=av, maxit=200000, fused=fused)
                t_setup = time.perf_counter() - t0
                s.solve(F)  # captures the iteration graph
                ts = []
                for _ in range(reps):
                    u, st = s.solve(F)
                    ts.append(st.solve_ms)
                torch.cuda.synchronize()
                res[key] = {"solve_ms": round(median(ts), 3), "iterations": st.iterations,
                            "reliable_updates": st.outer_iterations,
                            "us_per_iteration": round(median(ts) * 1e3 / st.iterations, 1),
                            "rel_residual": st.rel_residual, "status": st.status,
                            "setup_s": round(t_setup, 3)}
                sols[key] = u
                log(f"solve mg {key}: {median(ts):.2f} ms, {st.iterations} its")
                del s
    ref = sols["fp64_gather_fused0"]
    for k in sols:
        if k != "fp64_gather_fused0":
            res[k]["solution_rel_diff"] = (sols[k] - ref).norm().item() / ref.norm().item()
    best = {}
    for name in ("fp64", "fp32_refine"):
        best[name] = min((k for k in res if k.startswith(name + "_")),
                         key=lambda k: res[k]["solve_ms"])
        res[f"best_{name}"] = best[name]
    res["speedup_best_fp32_refine_vs_best_fp64"] = round(
        res[best["fp64"]]["solve_ms"] / res[best["fp32_refine"]]["solve_ms"], 3)
    res["preconditioner"] = ("fem_mg: Q1 FEM matrix on the reference's node grid mapped onto the "
                             "annulus, one V-cycle (Chebyshev-Jacobi degree 2, Galerkin)")
    res["inner_rtol"] = "library default (automatic: 1e-4 with a FEM preconditioner)"
    res["fem_assembly_s"] = round(t_fem, 3)
    if "solve" in out and "best_fp64" in out["solve"]:
        jb = out["solve"]
        res["speedup_vs_best_jacobi_fp64"] = round(
            jb[jb["best_fp64"]]["solve_ms"] / res[best["fp32_refine"]]["solve_ms"], 2)
        res["speedup_fp64_vs_jacobi_fp64"] = round(
            jb[jb["best_fp64"]]["solve_ms"] / res[best["fp64"]]["solve_ms"], 2)
    out["solve_fem_mg"] = res


def fempc_iterations(out):
    """Outer iteration counts of the FEM-preconditioned solve on the unit
    square, p = 3, f = x(1-x)y(1-y), beside the reference's logged counts
    (matrix_free/data/fempc_mf.txt:12,24,36: 35, 33, 28).  First the
    reference's own solver semantics, driven through this library's
    MATSHELL and PCSHELL entry points: PETSc's default outer KSPGMRES
    (igamf.ksp_gmres: left preconditioning, preconditioned norm, rtol =
    abstol = 1e-10) around Operator.apply, with Solver.apply_pc running the
    inner CG + Jacobi on the Q1 matrix to rtol = abstol = 1e-10.  Then this
    library's solver (stop on the unpreconditioned relative residual 1e-10):
    the reference's preconditioner (flexible outer CG), one V-cycle, and
    Jacobi."""
    import time as _t
    import torch
    import igamf
    ref = {32: 35, 64: 33, 128: 28}
    rows = {}
    for n in (32, 64, 128):
        p = 3
        mesh = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, n)
        op = igamf.Operator(mesh)
        fem = igamf.FemPC(n + p, n + p, igamf.fem_reference_nodes(p, p, n, n))
        F = op.load(igamf.LOAD_REF_POLY, 0.0)
        row = {"reference_logged": ref[n]}
        spc = igamf.Solver(op, fem, pc=igamf.PC_FEM_CG, pc_inner_rtol=1e-10,
                           pc_inner_abstol=1e-10, pc_inner_maxit=400)
        torch.cuda.synchronize()
        t0 = _t.perf_counter()
        u, its = igamf.ksp_gmres(lambda v: op.apply(v.contiguous()), F,
                                 lambda v: spc.apply_pc(v.contiguous()))
        torch.cuda.synchronize()
        r = F - op.apply(u)
        row["ksp_gmres_reference_semantics"] = {
            "iterations": its, "solve_ms": round((_t.perf_counter() - t0) * 1e3, 2),
            "rel_residual": (r.norm() / F.norm()).item()}
        # the inner CG is unrolled into pc_inner_maxit iterations per outer
        # one (each returning at once after convergence): launched eagerly,
        # not captured, at this length
        for name, kw in (("inner_cg_1e-10", dict(pc=igamf.PC_FEM_CG, pc_inner_rtol=1e-10,
                                                pc_inner_maxit=1500, use_graph=0)),
                         ("mg_vcycle", dict(pc=igamf.PC_FEM_MG)),
                         ("jacobi", dict(pc=igamf.PC_JACOBI))):
            s = igamf.Solver(op, fem, rtol=1e-10, **kw)
            s.solve(F)
            _, st = s.solve(F)
            row[name] = {"iterations": st.iterations, "solve_ms": round(st.solve_ms, 2),
                         "rel_residual": st.rel_residual}
        rows[f"{n}x{n}"] = row
    out["fempc_iterations"] = rows
    log("fem-pc iterations (reference KSP semantics / ours inner-CG / V-cycle / jacobi vs "
        "reference log): " + ", ".join(
            f"{k}: {v['ksp_gmres_reference_semantics']['iterations']}/"
            f"{v['inner_cg_1e-10']['iterations']}/{v['mg_vcycle']['iterations']}/"
            f"{v['jacobi']['iterations']} vs {v['reference_logged']}" for k, v in rows.items()))


def block(out, with_solve=True):
    """Multi-RHS apply, quarter annulus p = 4, 512^2, 5 x 5 Gauss, k = 32:
    the apply on f_j sampled at the Greville abscissae, the block solve on the
    load vectors of f_j (a_mn ~ N(0,1) from seed j)."""
    import numpy as np
    import torch
    import igamf
    p, ne, k = 4, 512, 32
    mesh = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, p, ne, r0=1.0, r1=2.0)
    op = igamf.Operator(mesh)
    n = op.nfunc
    X = torch.from_numpy(c3_greville_columns(p, ne, k)).cuda().float()
    Y = torch.empty_like(X)
    x0 = X[0].contiguous()
    t_single, _ = ev_time(lambda: op.apply(x0), 20, 3)
    cfgs = igamf.block_configs()
    tuned = {}
    for kk in (8, 16, 32):
        cfg, ms = op.tune_block(X[:kk], Y[:kk], reps=5)
        tuned[kk] = {"cfg": list(cfgs[cfg]), "tuned_ms": round(ms, 4)}
    best = op.block_config(k)
    # every matrix-core configuration at k, against the tuned choice
    mfma = {}
    for ci, c in enumerate(cfgs):
        if not igamf.block_config_matrix_core(ci):
            continue
        op.set_block_config(k, ci)
        tm, _ = ev_time(lambda: op.apply_block(X, Y), 10, 2)
        mfma[f"{c[0]}x{c[1]}x{c[2]}"] = {
            "ms": round(tm, 4),
            "tflops": round((16 * (p + 1) ** 3 + 6 * (p + 1) ** 2) * op.nelem * k
                            / (tm * 1e-3) / 1e12, 2)}
    op.set_block_config(k, best)
    t_block, _ = ev_time(lambda: op.apply_block(X, Y), 20, 3)
    single = torch.stack([op.apply(X[j].contiguous(), assembly=igamf.IEN_ATOMIC)
                          for j in range(k)])
    rel = ((Y - single).abs().max() / single.abs().max()).item()
    flops = (16 * (p + 1) ** 3 + 6 * (p + 1) ** 2) * op.nelem * k
    tf = flops / (t_block * 1e-3) / 1e12
    out["block"] = {
        "workload": f"quarter annulus p={p}, {ne}x{ne}, {p + 1}x{p + 1} Gauss, k={k}, FP32",
        "block_ms": round(t_block, 4), "per_rhs_us": round(t_block * 1e3 / k, 2),
        "single_apply_us": round(t_single * 1e3, 2),
        "speedup_vs_single": round(t_single * k / t_block, 3),
        "tflops": round(tf, 2), "frac_fp32_peak": round(tf / FP32_PEAK_TFLOPS, 4),
        "flops_per_call": flops, "tuned": tuned,
        "config_used": list(cfgs[op.block_config(k)]),
        "config_used_matrix_core": igamf.block_config_matrix_core(op.block_config(k)),
        "matrix_core_configs": mfma,
        "max_rel_diff_vs_single": rel,
        "note": "tuned over the VALU (non-packed FP32) and matrix-core "
                "(v_mfma_f32_16x16x4f32, Kronecker form) kernels; tflops counts the "
                "sum-factorised 16(p+1)^3 + 6(p+1)^2 per element and rhs for both"}
    log(f"block k={k}: {t_block * 1e3 / k:.1f} us/rhs, {tf:.1f} TFLOP/s")
    if not with_solve:
        return
    # block PCG: the k load vectors F_j of f_j in one block solve, against k
    # single FP32 + refinement solves and k plain FP64 solves, Jacobi, rtol 1e-8
    F = torch.stack([op.load_qp(torch.from_numpy(c3_qp_values(p, ne, j)).cuda())
                     for j in range(k)]).contiguous()
    sb = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE, pc=igamf.PC_JACOBI, rtol=1e-8,
                      maxit=200000)
    sb.solve_block(F)  # captures the block chunk graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Ub, stb = sb.solve_block(F)
    torch.cuda.synchronize()
    t_bs = (time.perf_counter() - t0) * 1e3
    res = {}
    for name, kind in (("fp32_refine", igamf.SOLVE_PCG_F32_REFINE), ("fp64", igamf.SOLVE_PCG_F64)):
        ss = igamf.Solver(op, None, kind=kind, pc=igamf.PC_JACOBI, rtol=1e-8, maxit=200000)
        ss.solve(F[0].contiguous())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Us, its = [], []
        for j in range(k):
            u, st = ss.solve(F[j].contiguous())
            Us.append(u)
            its.append(st.iterations)
        torch.cuda.synchronize()
        res[name] = ((time.perf_counter() - t0) * 1e3, torch.stack(Us), its)
    ref = res["fp64"][1]
    out["block"]["solve"] = {
        "k": k, "rtol": 1e-8, "preconditioner": "jacobi",
        "rhs": "F_j = load vector of f_j (iga_op_load_qp at the 5x5 Gauss points)",
        "block_fp32_refine_ms": round(t_bs, 2),
        "single_fp32_refine_ms_total": round(res["fp32_refine"][0], 2),
        "single_fp64_ms_total": round(res["fp64"][0], 2),
        "speedup_block_vs_single_fp32_refine": round(res["fp32_refine"][0] / t_bs, 3),
        "speedup_block_vs_single_fp64": round(res["fp64"][0] / t_bs, 3),
        "block_iterations_max": max(st.iterations for st in stb),
        "single_fp64_iterations_max": max(res["fp64"][2]),
        "block_rel_residual_max": max(st.rel_residual for st in stb),
        "max_rel_diff_vs_fp64": ((Ub - ref).abs().max() / ref.abs().max()).item()}
    log(f"block solve k={k}: {t_bs:.1f} ms against {res['fp32_refine'][0]:.1f} ms (FP32 "
        f"singles) and {res['fp64'][0]:.1f} ms (FP64 singles)")


def c5(out):
    """p = 5, 2048^2 unit square, interior points moved U[-0.1h, 0.1h] (seed
    3), weights U[0.8, 1.2]: FP32 SF and FP64 EBE applies, FEM assembly.
    Returns the assembled FEM matrix for the CPU check of its indices."""
    import torch
    import igamf
    p, ne = 5, 2048
    t0 = time.perf_counter()
    mesh = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, p, ne, perturb=0.1, w_lo=0.8,
                            w_hi=1.2, seed=3)
    op = igamf.Operator(mesh)
    setup = time.perf_counter() - t0
    info = op.info
    n_int = (info.nfx - 2) * (info.nfy - 2)
    x = torch.from_numpy(igamf.rng_normal(0, op.nfunc)).cuda()
    x32 = x.float()
    y32 = torch.empty_like(x32)
    y64 = torch.empty_like(x)
    settle(lambda: op.apply(x32, y32))  # after seconds of host-side mesh set-up
    t32, _ = ev_time(lambda: op.apply(x32, y32), 20, 2)
    k32 = op.time_kernel(x32, torch.empty_like(x32), 30)
    t64e, _ = ev_time(lambda: op.apply(x, y64, kernel=igamf.EBE), 5, 1)
    nc = 5 if info.rational == 2 else 3
    b32 = 4 * (2 * op.nfunc) + 4 * nc * op.nelem * op.nqp + 4 * op.nfunc
    del op
    nodes = igamf.fem_reference_nodes(p, p, ne, ne)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fem = igamf.FemPC(info.nfx, info.nfy, nodes)
    torch.cuda.synchronize()
    t_fem = time.perf_counter() - t0
    out["c5"] = {
        "workload": f"unit square p={p}, {ne}x{ne}, perturbed, weights U[0.8,1.2] "
                    "(general rational kernels)",
        "interior_dofs": n_int, "rational_mode": info.rational,
        "f32_streamed_geometry_loads": bool(info.stream_g_f32),
        "f32_sf_ms": round(t32, 4), "f32_sf_gdofs": round(n_int / (t32 * 1e-3) / 1e9, 4),
        "f32_kernel_ms": round(k32, 4),
        "f32_hbm_frac": round(b32 / (k32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "f32_algorithmic_bytes": b32,
        "f64_ebe_ms": round(t64e, 4), "speedup_f32_sf_vs_f64_ebe": round(t64e / t32, 3),
        "fem_assembly_ms": round(t_fem * 1e3, 2), "fem_nnz": fem.view.nnz,
        "setup_s": round(setup, 2)}
    log(f"c5 f32 sf {t32 * 1e3:.0f} us, f64 ebe {t64e * 1e3:.0f} us")
    return fem, nodes, p, ne


def c1(out):
    """C1: FP64 EBE apply (100 applies, x ~ N(0,1) seed 0) and the solve to
    rtol 1e-10 inside CG with the Q1 FEM preconditioner: ILU(0) (the
    configuration's) and the reference GPU driver's inner CG + Jacobi."""
    import numpy as np
    import torch
    import igamf
    p, ne = 2, 32
    mesh = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, ne)
    op = igamf.Operator(mesh)
    x = torch.from_numpy(igamf.rng_normal(0, op.nfunc)).cuda()
    y = torch.empty_like(x)
    t100, _ = ev_time(lambda: op.apply(x, y, kernel=igamf.EBE), 100, 10)
    fem = igamf.FemPC(ne + p, ne + p, igamf.fem_reference_nodes(p, p, ne, ne))
    F = op.load(igamf.LOAD_SINSIN, 2 * np.pi ** 2)
    a = mesh.arrays()
    cp = a["CP"].reshape(-1, 2)
    out["c1"] = {"workload": "unit square p=2 32x32, FP64 EBE apply, rtol 1e-10, preconditioned "
                             "by the Q1 FEM matrix: ILU(0) (the configuration's) and the "
                             "reference GPU driver's inner CG + Jacobi",
                 "apply_ms": round(t100, 5)}
    for name, kw in (("ilu", dict(pc=igamf.PC_FEM_ILU)),
                     ("fem_cg", dict(pc=igamf.PC_FEM_CG, pc_inner_maxit=400,
                                     pc_inner_rtol=1e-10))):
        s = igamf.Solver(op, fem, kernel=igamf.EBE, rtol=1e-10, **kw)
        s.solve(F)
        u, st = s.solve(F)
        # reference sign convention: u_h ~ -sin(pi x) sin(pi y) (Greville interpolant)
        err = float(np.abs(u.cpu().numpy() + np.sin(np.pi * cp[:, 0]) *
                           np.sin(np.pi * cp[:, 1])).max())
        out["c1"][name] = {"solve_ms": round(st.solve_ms, 2), "iterations": st.iterations,
                           "rel_residual": st.rel_residual,
                           "max_err_vs_minus_sinsin_at_cp": err}
    log(f"c1 solve ilu {out['c1']['ilu']['iterations']} its, fem-cg "
        f"{out['c1']['fem_cg']['iterations']} its")


def c4_kappa(op, seed):
    """kappa_q = exp(0.5 s(x_q)), s = sum_{a,b<=4} c_ab sin(a pi x) sin(b pi y),
    c_ab ~ N(0, 1/(a^2+b^2)) (splitmix64 normals of the seed, a-major), at
    the physical quadrature points; float64 CUDA [E * nqp]."""
    import numpy as np
    import torch
    import igamf
    xq = op.qp_coords().reshape(-1, 2)
    z = igamf.rng_normal(seed, 16)
    s = torch.zeros(xq.shape[0], dtype=torch.float64, device=xq.device)
    for a in range(1, 5):
        for b in range(1, 5):
            c = z[(a - 1) * 4 + b - 1] / np.sqrt(a * a + b * b)
            s += c * torch.sin(a * np.pi * xq[:, 0]) * torch.sin(b * np.pi * xq[:, 1])
    return torch.exp(0.5 * s).contiguous()


def c4(out):
    """C4: -div(kappa grad u) = f on the unit square, p = 3, 256^2, 4 x 4
    Gauss, f = 1; kappa from seed 1, u_obs = the solution for the seed-2 kappa
    plus N(0, 1e-3) noise (seed 9) on the interior; J = 1/2 |u - u_obs|^2.
    Forward solve, adjoint solve and the grad_kappa kernel, timed; the
    gradient against FP64 central differences (step 1e-6) along 16 random
    directions (seeds 300..315)."""
    import torch
    import igamf
    from igamf import autograd as IA
    p, ne = 3, 256
    mesh = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, ne)
    op = igamf.Operator(mesh)
    f = op.load(igamf.LOAD_CONST, 1.0)
    s = igamf.Solver(op, None, pc=igamf.PC_JACOBI, rtol=1e-13, maxit=200000)
    k1, k2 = c4_kappa(op, 1), c4_kappa(op, 2)
    interior = IA.interior_mask(op)
    with torch.no_grad():
        u_obs = IA.solve_kappa(s, f, k2)
        u_obs = u_obs + 1e-3 * torch.from_numpy(igamf.rng_normal(9, op.nfunc)).cuda() * interior
    IA.solve_kappa(s, f, k1)  # graph capture outside the timed region
    torch.cuda.synchronize()
    # three timed passes, the median reported: the first also pays torch's
    # first use of its autograd and elementwise kernels
    t_runs = []
    for _ in range(3):
        k = k1.clone().requires_grad_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u = IA.solve_kappa(s, f, k)
        J = 0.5 * ((u - u_obs) ** 2).sum()
        J.backward()
        torch.cuda.synchronize()
        t_runs.append(time.perf_counter() - t0)
    t_grad = median(t_runs)
    grad = k.grad.clone()
    h = 1e-6
    errs = []
    with torch.no_grad():
        for d in range(16):
            dk = torch.from_numpy(igamf.rng_normal(300 + d, k1.numel())).cuda()
            jp = 0.5 * ((IA.solve_kappa(s, f, k1 + h * dk) - u_obs) ** 2).sum().item()
            jm = 0.5 * ((IA.solve_kappa(s, f, k1 - h * dk) - u_obs) ** 2).sum().item()
            fd = (jp - jm) / (2 * h)
            an = torch.dot(grad, dk).item()
            errs.append(abs(fd - an) / max(abs(fd), 1e-300))
    out["c4"] = {"workload": "unit square p=3 256x256, 4x4 Gauss, kappa = exp(0.5 s(x_q)) with "
                             "16 Fourier modes (seed 1), u_obs from the seed-2 kappa + N(0,1e-3) "
                             "noise, J = 0.5|u - u_obs|^2, f = 1, Jacobi PCG to rtol 1e-13",
                 "J": J.item(),
                 "forward_adjoint_grad_s": round(t_grad, 4),
                 "fd_directions": 16, "fd_step": h,
                 "fd_rel_err_max": max(errs), "fd_rel_err_median": median(errs)}
    log(f"c4 forward+adjoint+gradient {t_grad * 1e3:.1f} ms; FD rel err max {max(errs):.2e} "
        f"over 16 directions")


def _cpu_pcg_ilu(O, m, F, rowptr, col, lu, diag, rtol):
    """CG on the oracle's apply with the oracle's ILU(0) of the Q1 matrix (the
    reference CPU driver's KSPCG + PCILU; zero initial guess, ||f - A u|| /
    ||f|| <= rtol)."""
    import numpy as np
    u = np.zeros_like(F)
    r = F.copy()
    z = O.ilu0_solve(rowptr, col, lu, diag, r)
    pv = z.copy()
    rz = r @ z
    nf = np.linalg.norm(F)
    for it in range(1, 10000):
        Ap = O.apply_mf(m, pv)
        a = rz / (pv @ Ap)
        u += a * pv
        r -= a * Ap
        if np.linalg.norm(r) <= rtol * nf:
            return u, it
        z = O.ilu0_solve(rowptr, col, lu, diag, r)
        rz, rz_old = r @ z, rz
        pv = z + (rz / rz_old) * pv
    return u, it


def cpu_baseline(a, out, mesh, x64, op, kern, asm, n_int):
    """The oracle's restatement of the reference CPU path on the host cores:
    C2 applies on all cores (median of --cpu-applies), one apply on one core
    of a 1024 x 128 element strip of the same annulus, C1 applies (100) and
    the C1 solve (CG + ILU(0) of the Q1 matrix, rtol 1e-10)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    arr = mesh.arrays()
    om = O.Mesh(arr["p"], arr["q"], arr["nex"], arr["ney"], arr["CP"], arr["W"], arr["IEN"],
                arr["ID"], arr["Dir"], arr["C1"], arr["C2"], arr["h1"], arr["h2"])
    xh = x64.cpu().numpy()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ts = []
    used = threads
    yc = None
    for _ in range(a.cpu_applies):
        t0 = time.perf_counter()
        yc, used = O.apply_mf_colored(om, xh, threads)
        ts.append(time.perf_counter() - t0)
    tc = median(ts)
    out["cpu_baseline"] = {"value": round(n_int / tc / 1e9, 6), "unit": "GDoF/s",
                           "cores": used, "kind": "port",
                           "sample": f"{a.cpu_applies} full FP64 applies of the same "
                                     f"{a.nelem}x{a.nelem} p={a.p} mesh (median {tc:.3f} s), "
                                     f"OpenMP over a (p+1)^2 element colouring"}
    y64 = op.apply(x64, kernel=kern, assembly=asm).cpu().numpy()
    out["cpu_baseline"]["gpu_f64_rel_diff"] = float(np.abs(y64 - yc).max() / np.abs(yc).max())
    log(f"cpu baseline {tc:.2f} s per apply on {used} threads")
    # one core, bounded sample: a strip of 1024 x 128 elements of the annulus
    sm = O.mesh_quarter_annulus(a.p, a.nelem, a.nelem // 8)
    sx = O.rng_normal(0, sm.nfunc)
    t0 = time.perf_counter()
    O.apply_mf_colored(sm, sx, 1)
    t1c = time.perf_counter() - t0
    s_int = (a.nelem + a.p - 2) * (a.nelem // 8 + a.p - 2)
    out["cpu_baseline"]["one_core"] = {
        "value": round(s_int / t1c / 1e9, 6), "unit": "GDoF/s", "cores": 1,
        "sample": f"one FP64 apply of a {a.nelem}x{a.nelem // 8} element strip of the same "
                  f"annulus, p={a.p} ({t1c:.2f} s)"}
    # C1: 100 applies on one core and on all, the CPU solve
    m1 = O.mesh_unit_square(2, 32)
    x1 = O.rng_normal(0, m1.nfunc)
    c1 = {}
    for nt in (1, threads):
        t0 = time.perf_counter()
        for _ in range(100):
            O.apply_mf_colored(m1, x1, nt)
        c1[f"apply_ms_{nt}_cores"] = round((time.perf_counter() - t0) * 1e3 / 100, 4)
    F1 = O.load(m1, O.LOAD_SINSIN, 2 * np.pi ** 2)
    rows, cols, _, _, val = O.fem_assemble(34, 34, O.fem_nodes(2, 2, 32, 32))
    rowptr = np.searchsorted(rows, np.arange(34 * 34 + 1))
    t0 = time.perf_counter()
    lu, diag = O.ilu0(rowptr, cols, val)
    u1, its = _cpu_pcg_ilu(O, m1, F1, rowptr, cols, lu, diag, 1e-10)
    c1["solve_s"] = round(time.perf_counter() - t0, 4)
    c1["solve_iterations"] = its
    c1["solve"] = "CG on the oracle apply + ILU(0) of the Q1 matrix (Python triangular solves)"
    out["cpu_baseline"]["c1"] = c1
    log(f"cpu one core {t1c:.2f} s on the strip; c1 cpu solve {its} its")


def cpu_check_fem(out, fem, nodes, p, ne):
    """C5 FEM assembly on the CPU (the oracle's restatement of elem2coo /
    dir2coo / the Q1 values), timed, and the GPU indices compared bit for
    bit with it."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nf = ne + p
    t0 = time.perf_counter()
    rows, cols, e2c, d2c, val = O.fem_assemble(nf, nf, nodes)
    t_cpu = time.perf_counter() - t0
    t = fem.to_torch()
    same = (np.array_equal(t["coo_row"].cpu().numpy(), rows) and
            np.array_equal(t["col"].cpu().numpy(), cols) and
            np.array_equal(t["elem2coo"].cpu().numpy().reshape(-1), e2c) and
            np.array_equal(t["dir2coo"].cpu().numpy(), d2c))
    vd = float(np.abs(t["val"].cpu().numpy() - val).max() / np.abs(val).max())
    out["c5"]["fem_indices_match_cpu"] = bool(same)
    out["c5"]["fem_values_max_rel_diff_cpu"] = vd
    out["cpu_baseline"]["fem_assembly_c5_s"] = round(t_cpu, 3)
    log(f"c5 fem indices match the CPU: {same}; CPU assembly {t_cpu:.1f} s")


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {a.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    out = {}
    mesh, op, x64, x32, kern, asm, n_int = headline(a, out)
    if a.only_apply and a.also_time:
        import torch
        for form in a.also_time.split(","):
            if form == "g16":
                op.time_kernel(x32, torch.empty_like(x32), a.steps, kern, asm, g16=True)
            elif form == "f64":
                op.time_kernel(x64, torch.empty_like(x64), a.steps, kern, asm)
    only = set(x for x in a.only.split(",") if x)

    def want(sec, skip):
        return sec in only if only else not skip

    if not a.only_apply:
        if want("variants", False):
            fp64_and_variants(a, out, op, x64, x32, kern, asm, n_int, op.nfunc, op.nelem,
                              op.nqp)
        if want("solve", a.no_solve):
            solves(out, op, tuple(x for x in a.solve_configs.split(",") if x), a.solve_reps)
        if want("mg", a.no_solve or bool(a.solve_configs)):
            solves_mg(out, mesh, op, a.solve_reps)
        if want("cpu", a.no_cpu):
            cpu_baseline(a, out, mesh, x64, op, kern, asm, n_int)
        del op
        if want("block", a.no_block):
            block(out, not a.no_block_solve)
        if want("c5", a.no_c5):
            fem, nodes, p5, ne5 = c5(out)
            if want("cpu", a.no_cpu):
                cpu_check_fem(out, fem, nodes, p5, ne5)
            del fem
        if want("extra", a.no_extra):
            c1(out)
            fempc_iterations(out)
            c4(out)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
