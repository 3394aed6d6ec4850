"""Write the golden vectors under tests/golden/ from the reference's own CPU
units (oracle/_ref/libref_mf.so and libref_gpu.so, compiled from the reference
sources by oracle/Makefile; no reference source enters the repository).

Run in the development container only (the reference tree is not on the GPU
machines):  make -C oracle ref && python tests/golden/make_golden.py

Each .npz holds inputs and the reference's outputs; SHA256SUMS lists the
files.  Inputs: Greville control points (the reference's Eigen control-point
solve is not buildable here; Greville points are its exact solution),
x ~ N(0,1) from splitmix64/Box-Muller (seed in the file).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def mf_case(name, p, n, seed):
    R = O.ref_mf()
    S = np.empty(n + 2 * p + 1)
    R.ref_open_knots(p, n, 1.0, S)
    Cx = np.empty(n * (p + 1) ** 2)
    R.ref_extraction(p, n, 1.0, Cx)
    nf = n + p
    IEN = np.empty(n * n * (p + 1) ** 2, np.int32)
    ID = np.empty(nf * nf, np.int32)
    Dir = np.empty(nf * nf, np.int32)
    nd = R.ref_topology(p, p, n, n, IEN, ID, Dir)
    Dir = Dir[:nd].copy()
    g = np.empty(nf)
    R.ref_greville(p, n, 1.0, g)
    CP = np.stack(np.meshgrid(g, g, indexing="xy"), axis=-1).reshape(-1)
    h = np.full(n, 1.0 / n)
    m = O.Mesh(p, p, n, n, CP, None, IEN, ID, Dir, Cx, Cx, h, h)
    x = O.rng_normal(seed, m.nfunc)
    qx = np.empty(p + 1); qw = np.empty(p + 1)
    R.ref_quadrature(p + 1, qx, qw)
    B = np.empty((p + 1) ** 2); dB = np.empty((p + 1) ** 2)
    R.ref_bernstein_table(p, p + 1, B, dB)
    # one interior and one corner element's per-qp data
    eq = {}
    for tag, (ex, ey) in (("corner", (0, 0)), ("interior", (n // 2, n // 2 - 1))):
        nloc = (p + 1) ** 2
        e = ey * n + ex
        ien = IEN[e * nloc:(e + 1) * nloc]
        eCP = np.ascontiguousarray(CP.reshape(-1, 2)[ien].reshape(-1))
        Rr = np.empty(nloc * nloc); dx = np.empty(nloc * nloc); dy = np.empty(nloc * nloc)
        J = np.empty(nloc)
        C1e = np.ascontiguousarray(Cx.reshape(n, -1)[ex])
        C2e = np.ascontiguousarray(Cx.reshape(n, -1)[ey])
        R.ref_element_qp(p, p, C1e, C2e, 1.0 / n, 1.0 / n, eCP, Rr, dx, dy, J)
        eq.update({f"{tag}_elem": np.array([ex, ey], np.int32), f"{tag}_R": Rr,
                   f"{tag}_dRdx": dx, f"{tag}_dRdy": dy, f"{tag}_JxW": J})
    out = dict(p=p, nelem=n, seed=seed, knots=S, extraction=Cx, IEN=IEN, ID=ID, Dir=Dir,
               greville=g, CP=CP, quad_x=qx, quad_w=qw, bern_B=B, bern_dB=dB, x=x,
               y_mf=O.ref_apply_mf(m, x), y_sf=O.ref_apply_sf(m, x), F_ref=O.ref_load(m), **eq)
    np.savez_compressed(os.path.join(HERE, name), **out)


def fem_case(name, p, n):
    G = O.ref_gpu()
    nf = n + p
    cap = 9 * nf * nf
    rows = np.empty(cap, np.int32); cols = np.empty(cap, np.int32)
    e2c = np.empty((nf - 1) ** 2 * 16, np.int32)
    d2c = np.empty(nf * nf, np.int32)
    val = np.empty(cap)
    import ctypes
    nd = ctypes.c_int()
    nnz = G.ref_fem_assemble(p, p, n, n, 1.0, 1.0, cap, rows, cols, e2c, d2c, val,
                             ctypes.byref(nd))
    CPf = np.empty(2 * nf * nf)
    G.ref_fem_nodes(p, p, n, n, 1.0, 1.0, CPf)
    np.savez_compressed(os.path.join(HERE, name), p=p, nelem=n, nodes=CPf,
                        rows=rows[:nnz], cols=cols[:nnz], elem2coo=e2c,
                        dir2coo=d2c[:nd.value], val=val[:nnz])


def invlm_case(name, p, n):
    G = O.ref_gpu()
    IEN, ID, Dir = O.topology(p, p, n, n)
    nf = (n + p) ** 2
    nloc = (p + 1) ** 2
    en = np.empty(nf, np.int32); off = np.empty(nf, np.int32)
    ei = np.empty(n * n * nloc, np.int32); bi = np.empty(n * n * nloc, np.int32)
    G.ref_invlm(nloc, n * n, nf, IEN, en, off, ei, bi)
    np.savez_compressed(os.path.join(HERE, name), p=p, nelem=n, IEN=IEN, elemNum=en,
                        offset=off, elemIdx=ei, baseIdx=bi)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    mf_case("mf_p2_n32.npz", 2, 32, 0)
    mf_case("mf_p3_n16.npz", 3, 16, 0)
    fem_case("fem_p3_n16.npz", 3, 16)
    fem_case("fem_p4_n12.npz", 4, 12)
    invlm_case("invlm_p2_n6.npz", 2, 6)
    with open(os.path.join(HERE, "SHA256SUMS"), "w") as f:
        for fn in sorted(os.listdir(HERE)):
            if fn.endswith(".npz"):
                h = hashlib.sha256(open(os.path.join(HERE, fn), "rb").read()).hexdigest()
                f.write(f"{h}  {fn}\n")


if __name__ == "__main__":
    main()
