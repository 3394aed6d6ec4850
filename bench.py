"""Headline benchmark: matrix-free IGA Poisson apply (GDoF/s) and the
CG solve-to-1e-8 time, FP32 apply + FP64 refinement vs plain FP64, on the
quarter annulus r in [1,2], exact NURBS, p = 3, 1024 x 1024 elements, 4 x 4
Gauss points (1,050,625 interior unknowns), synthetic x ~ N(0,1) (seed 0).

A step is one FP32 sum-factorised apply y = A x (inverse-LM gather
assembly) with x and the operator resident in HBM.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-computing_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nelem", type=int, default=1024)
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-applies", type=int, default=3)
    ap.add_argument("--assembly", choices=["gather", "atomic"], default="gather")
    ap.add_argument("--kernel", choices=["sf", "ebe"], default="sf")
    ap.add_argument("--detail", action="store_true", help="also time every apply variant")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 256 MiB device copy first (PMC byte-counter calibration)")
    return ap.parse_args()


def median(v):
    v = sorted(v)
    return v[len(v) // 2]


def time_applies(op, x, steps, warmup, kernel, assembly):
    import torch
    y = torch.empty_like(x)
    for _ in range(warmup):
        op.apply(x, y, kernel=kernel, assembly=assembly)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        op.apply(x, y, kernel=kernel, assembly=assembly)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) for a, b in ev]
    return median(per), sum(per) / len(per), wall


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {a.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    import igamf

    p, n = a.p, a.nelem
    t_setup0 = time.perf_counter()
    mesh = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, p, n, r0=1.0, r1=2.0)
    op = igamf.Operator(mesh)
    setup_s = time.perf_counter() - t_setup0
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

    # headline: FP32 SF apply
    med32, mean32, wall32 = time_applies(op, x32, a.steps, a.warmup, kern, asm)
    # the dominant kernel alone (HIP events around k_apply on its stream)
    op.profile(True, a.steps)
    y = torch.empty_like(x32)
    for _ in range(a.steps):
        op.apply(x32, y, kernel=kern, assembly=asm)
    torch.cuda.synchronize()
    cnt, tot, kmin, kmax = op.profile_read()
    op.profile(False)
    k_avg_ms = tot / max(cnt, 1)
    alg_bytes32 = 4 * (2 * nfunc) + 4 * 3 * E * nqp
    achieved = alg_bytes32 / (k_avg_ms * 1e-3) / 1e9
    gdofs = n_int / (med32 * 1e-3) / 1e9

    out = {
        "metric": "MF apply GDoF/s & CG solve-to-1e-8 time: FP32+refine vs FP64, p=3 1024² elems",
        "value": round(gdofs, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(med32, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: x ~ N(0,1) splitmix64 seed 0; geometry generated (exact NURBS quarter annulus)",
        "config": {"workload": f"quarter annulus r in [1,2], p={p}, {n}x{n} elements, "
                               f"{p + 1}x{p + 1} Gauss, FP32 sum-factorised apply, "
                               f"{a.assembly} assembly",
                   "p": p, "nelem": n, "interior_dofs": n_int, "control_points": nfunc,
                   "kernel": a.kernel, "assembly": a.assembly,
                   "tile_f32": [info.tile_x_f32, info.tile_y_f32],
                   "tile_f64": [info.tile_x_f64, info.tile_y_f64]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_apply<float,3,4,...>",
                     "kernel_avg_ms": round(k_avg_ms, 5), "kernel_min_ms": round(kmin, 5),
                     "algorithmic_bytes": alg_bytes32},
        "apply": {"f32_median_ms": round(med32, 5), "f32_mean_ms": round(mean32, 5),
                  "setup_s": round(setup_s, 3), "op_setup_ms": round(info.setup_ms, 2),
                  "device_bytes": info.device_bytes},
    }

    # FP64 apply
    med64, mean64, _ = time_applies(op, x64, a.steps, a.warmup, kern, asm)
    alg_bytes64 = 8 * (2 * nfunc) + 8 * 3 * E * nqp
    out["apply"]["f64_median_ms"] = round(med64, 5)
    out["apply"]["f64_gdofs"] = round(n_int / (med64 * 1e-3) / 1e9, 4)
    out["apply"]["f64_hbm_frac_wall"] = round(alg_bytes64 / (med64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if a.detail:
        var = {}
        for kn, kv in (("sf", igamf.SF), ("ebe", igamf.EBE)):
            for an, av in (("gather", igamf.INVLM_GATHER), ("atomic", igamf.IEN_ATOMIC)):
                for dn, xx in (("f32", x32), ("f64", x64)):
                    m, _, _ = time_applies(op, xx, max(10, a.steps // 4), 3, kv, av)
                    var[f"{kn}_{an}_{dn}_ms"] = round(m, 5)
        out["apply"]["variants"] = var

    # solves: plain FP64 PCG vs FP32 PCG under FP64 defect correction
    if not a.no_solve:
        F = op.load(igamf.LOAD_CONST, 1.0)
        res = {}
        sols = {}
        for name, kind in (("fp64", igamf.SOLVE_PCG_F64), ("fp32_refine", igamf.SOLVE_PCG_F32_REFINE)):
            s = igamf.Solver(op, None, kind=kind, pc=igamf.PC_JACOBI, rtol=1e-8,
                             inner_rtol=1e-4, maxit=200000)
            u, st = s.solve(F)  # warm (graph capture)
            u, st = s.solve(F)
            torch.cuda.synchronize()
            res[name] = {"solve_ms": round(st.solve_ms, 2), "iterations": st.iterations,
                         "outer": st.outer_iterations, "rel_residual": st.rel_residual,
                         "status": st.status}
            sols[name] = u
        d = (sols["fp32_refine"] - sols["fp64"]).norm().item() / sols["fp64"].norm().item()
        res["solution_rel_diff"] = d
        res["speedup_fp32_refine_vs_fp64"] = round(res["fp64"]["solve_ms"] /
                                                   res["fp32_refine"]["solve_ms"], 3)
        res["preconditioner"] = "jacobi"
        out["solve"] = res

    # CPU baseline: the oracle's restatement of the reference apply on host cores
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        arr = mesh.arrays()
        om = O.Mesh(arr["p"], arr["q"], arr["nex"], arr["ney"], arr["CP"], arr["W"], arr["IEN"],
                    arr["ID"], arr["Dir"], arr["C1"], arr["C2"], arr["h1"], arr["h2"])
        xh = x64.cpu().numpy()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ts = []
        for _ in range(a.cpu_applies):
            t0 = time.perf_counter()
            yc, used = O.apply_mf_colored(om, xh, threads)
            ts.append(time.perf_counter() - t0)
        tc = median(ts)
        out["cpu_baseline"] = {"value": round(n_int / tc / 1e9, 6), "unit": "GDoF/s",
                               "cores": used, "kind": "port",
                               "sample": f"{a.cpu_applies} full FP64 applies of the same "
                                         f"{n}x{n} p={p} mesh (median {tc:.3f} s), OpenMP "
                                         f"over a (p+1)^2 element colouring"}
        # the GPU FP64 apply agrees with it
        y64 = op.apply(x64, kernel=kern, assembly=asm).cpu().numpy()
        import numpy as np
        out["cpu_baseline"]["gpu_f64_rel_diff"] = float(np.abs(y64 - yc).max() /
                                                        np.abs(yc).max())
    print(json.dumps(out))


if __name__ == "__main__":
    main()
