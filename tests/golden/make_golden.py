"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in a checkout where the reference sources were available when
oracle/Makefile built oracle/_ref/libfeecref.so:

    python tests/golden/make_golden.py

Each fixture holds inputs that are cheap to regenerate (the matrix comes from
simplefem_amd.systems by name, the right-hand side from a seed) and the
outputs of the compiled reference at one OpenMP thread: colourings
(F, B, R), iteration counts, final residual norms and, for the small
systems, x and residual in full.  For the full-size systems only
SHA-256 digests of F/B/R, the colour count and the iteration counts are kept.
The reference's refined unit square / unit cube and its P1 stiffness
matrices are kept in full for small levels (mesh_*.npz) and as SHA-256
digests of the CSR structure at the full-size levels (2D level 7, 3D level 6).
Block-system solves of HodgeConjugateResidualSolverCSR_SSOR keep K, x and
residual in full (hodge_*.npz, and hodge_{diagonal,textbook}_*.npz for
HodgeConjugateResidualSolverCSR_diagonal and _textbook); `--hodge-only` adds just those to an existing
golden.json.  The reference's lowest-order Whitney 1-form curl-curl and mass
matrices on the refined unit cube (FEEC inclusion, broken exterior derivative
and broken mass matrices, tangential Dirichlet edges) are kept in full with the
edges' vertices (whitney1_L*.npz); `--whitney-only` writes just those.
The degree-r (1, 2, 3) Lagrange stiffness in FEEC's Sullivan basis on the
refined unit square is kept in full at levels 1-3 (sullivan2d_L*_r*.npz) and
as structure digests with sampled values at levels 5 and 6
(sullivan2d_digests.json); `--sullivan-only` writes just those.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.ref import RefRunner  # noqa: E402
from simplefem_amd import systems as S  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
VARIANTS = ("cg", "jacobi", "ssor", "ssor_eisenstat", "rainbow", "eisenstat_rainbow")

SMALL = {
    "fd2d_8": lambda: S.fd_laplacian_2d(8),
    "p1_2d_L3": lambda: S.poisson2d_p1(3),
    "p1_2d_L4": lambda: S.poisson2d_p1(4),
    "kuhn_4": lambda: S.poisson3d_p1_kuhn(4),
    "spd_120": lambda: S.random_spd(120, 0.05, 3),
}
FULL = {
    "p1_2d_L7": lambda: S.poisson2d_p1(7),
    "kuhn_64": lambda: S.poisson3d_p1_kuhn(64),
    # the reference's own vertex numbering (uniform refinement order)
    "p1_2d_L7_refnum": lambda: S.poisson2d_p1(7, numbering="reference"),
    "kuhn_64_refnum": lambda: S.poisson3d_p1_kuhn(64, numbering="reference"),
}
# (dimension, refinement levels) of the meshes whose full data are kept
MESH_SMALL = [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)]
# full-size meshes: SHA-256 digests of the reference's CSR structure
MESH_FULL = {"p1_2d_L7_refnum": (2, 7), "kuhn_64_refnum": (3, 6)}


def aux_for(variant, d):
    if variant == "cg":
        return None
    if variant == "jacobi":
        return np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    return d


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


def mesh_fixtures(ref, meta):
    """The reference's refined meshes and their P1 stiffness matrices:
    small ones in full (mesh_<d>d_L<l>.npz), the full-size ones as digests."""
    meta["mesh"] = {}
    for dim, lv in MESH_SMALL:
        m = ref.mesh(dim, lv)
        bA, bC, bV = m["structural"]
        aA, aC, aV = m["compressed"]
        np.savez_compressed(os.path.join(OUT, f"mesh_{dim}d_L{lv}.npz"), coords=m["coords"],
                            simplices=m["simplices"], dirichlet=m["dirichlet"], sA=bA, sC=bC,
                            sV=bV, cA=aA, cC=aC, cV=aV)
        meta["mesh"][f"{dim}d_L{lv}"] = {"n": int(len(bA) - 1), "nnz": int(bA[-1]),
                                         "compressed_nnz": int(aA[-1])}
        print("mesh", dim, lv, flush=True)
    for name, (dim, lv) in MESH_FULL.items():
        m = ref.mesh(dim, lv)
        bA, bC, _ = m["structural"]
        aA, aC, _ = m["compressed"]
        meta["mesh"][name] = {"dim": dim, "levels": lv, "n": int(len(bA) - 1),
                              "nnz": int(bA[-1]), "sha256_A": sha(bA), "sha256_C": sha(bC),
                              "compressed_nnz": int(aA[-1]), "sha256_compressed_A": sha(aA),
                              "sha256_compressed_C": sha(aC),
                              "dirichlet_rows": int(m["dirichlet"].sum())}
        print("mesh", name, meta["mesh"][name]["nnz"], flush=True)


# HodgeConjugateResidualSolverCSR_SSOR cases: schur_blocks(m, coarse, c_weight, seed)
HODGE = {
    "schur_m16_w0": (16, 2, 0.0, 3),
    "schur_m16_w0.5": (16, 2, 0.5, 3),
    "schur_m32_w0": (32, 2, 0.0, 3),
    "schur_m40_w1": (40, 2, 1.0, 3),
}
HODGE_RHS_SEED = 5


def hodge_fixtures(ref, meta):
    """Block-system solves (HodgeConjugateResidualSolverCSR_SSOR) at
    precision 1e-10, inner precision 1e-12: K, x and residual in full."""
    meta["hodge"] = {}
    for name, (m, coarse, cw, seed) in HODGE.items():
        bs = S.schur_blocks(m, coarse, cw, seed)
        b = S.uniform_pm1(HODGE_RHS_SEED, bs.N)
        x, r, K = ref.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b, precision=1e-10,
                                    inner_precision=1e-12)
        np.savez_compressed(os.path.join(OUT, "hodge_" + name + ".npz"), x=x, res=r)
        # the _diagonal and _textbook variants on the same system
        ks = {}
        for inner in ("diagonal", "textbook"):
            xv, rv, ks[inner] = ref.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b,
                                                  precision=1e-10, inner_precision=1e-12,
                                                  inner=inner)
            np.savez_compressed(os.path.join(OUT, f"hodge_{inner}_{name}.npz"), x=xv, res=rv)
        meta["hodge"][name] = {"m": m, "coarse": coarse, "c_weight": cw, "seed": seed,
                               "rhs_seed": HODGE_RHS_SEED, "precision": 1e-10,
                               "inner_precision": 1e-12, "iterations": int(K),
                               "iterations_diagonal": int(ks["diagonal"]),
                               "iterations_textbook": int(ks["textbook"])}
        print("hodge", name, K, ks, flush=True)


# refinement levels of UnitCube3D whose Whitney 1-form matrices are kept
WHITNEY_LEVELS = (1, 2)


def whitney_fixtures(ref):
    """The reference's Whitney 1-form curl-curl and mass matrices as its FEEC
    assembly forms them, CSR as Conjugation leaves it (duplicates kept)."""
    for lv in WHITNEY_LEVELS:
        w = ref.whitney1(lv)
        kA, kC, kV = w["curlcurl"]
        mA, mC, mV = w["mass"]
        np.savez_compressed(os.path.join(OUT, f"whitney1_L{lv}.npz"), edges=w["edges"],
                            dirichlet=w["dirichlet"], kA=kA, kC=kC, kV=kV, mA=mA, mC=mC, mV=mV)
        print("whitney1", lv, len(w["edges"]), flush=True)


# (level, degrees) of UnitSquare2D_simple whose Sullivan-basis stiffness is
# kept in full; SULLIVAN_DIGEST_LEVELS keep SHA-256 digests of the structure
SULLIVAN_LEVELS = (1, 2, 3)
SULLIVAN_DIGEST_LEVELS = (5, 6)


def sullivan_fixtures(ref):
    """The reference's degree-r (1, 2, 3) stiffness in the Sullivan basis on the
    refined unit square, composed as tests/sullivan2D/poissondirichlet2D.cpp
    composes it: in full (sullivan2d_L*_r*.npz, with the mesh's edges and the
    Dirichlet flags) at small levels; structure digests, a sample of values
    and the largest magnitude at larger ones (sullivan2d_digests.json)."""
    for lv in SULLIVAN_LEVELS:
        for r in (1, 2, 3):
            o = ref.sullivan2d(lv, r)
            A, C, V = o["stiffness"]
            np.savez_compressed(os.path.join(OUT, f"sullivan2d_L{lv}_r{r}.npz"), A=A, C=C, V=V,
                                edges=o["edges"], vertex_dirichlet=o["vertex_dirichlet"],
                                edge_dirichlet=o["edge_dirichlet"], triangles=o["triangles"],
                                triangle_edges=o["triangle_edges"])
            print("sullivan2d", lv, r, len(A) - 1, len(V), flush=True)
    digests = {}
    for lv in SULLIVAN_DIGEST_LEVELS:
        for r in (1, 2, 3):
            A, C, V = ref.sullivan2d(lv, r)["stiffness"]
            pick = np.linspace(0, len(V) - 1, 64).astype(np.int64)
            digests[f"L{lv}_r{r}"] = {"n": len(A) - 1, "nnz": len(V), "sha256_A": sha(A),
                                      "sha256_C": sha(C), "sample_index": pick.tolist(),
                                      "sample_value": V[pick].tolist(),
                                      "max_abs": float(np.abs(V).max())}
            print("sullivan2d digest", lv, r, flush=True)
    with open(os.path.join(OUT, "sullivan2d_digests.json"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)


def main():
    if "--whitney-only" in sys.argv:
        with RefRunner(1) as ref:
            whitney_fixtures(ref)
        return
    if "--sullivan-only" in sys.argv:
        with RefRunner(1) as ref:
            sullivan_fixtures(ref)
        return
    if "--hodge-only" in sys.argv:  # adds the block-system fixtures to golden.json
        path = os.path.join(OUT, "golden.json")
        meta = json.load(open(path))
        with RefRunner(1) as ref:
            hodge_fixtures(ref, meta)
        with open(path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        return
    with RefRunner(1) as ref:
        meta = {"desired_precision": ref.info()["desired_precision"], "small": {}, "full": {}}
        mesh_fixtures(ref, meta)
        whitney_fixtures(ref)
        sullivan_fixtures(ref)
        for name, make in SMALL.items():
            s = make()
            b = S.rhs_uniform(s, 11)
            d = ref.csr_diagonal(s.A, s.C, s.V)
            col = ref.rainbow(s.A, s.C, s.V)
            arrays = {"F": col[1], "B": col[2], "R": col[3], "diag": d,
                      "apply": ref.csr_apply(s.A, s.C, s.V, S.uniform_pm1(5, s.n), -0.37)}
            ks = {}
            for omega in (1.0, 1.2):
                for v in VARIANTS:
                    if v in ("cg", "jacobi") and omega != 1.0:
                        continue
                    x, r, K, _ = ref.solve(v, s.A, s.C, s.V, b, aux=aux_for(v, d), omega=omega,
                                           colouring=col)
                    tag = f"{v}_w{omega}"
                    arrays["x_" + tag] = x
                    arrays["res_" + tag] = r
                    ks[tag] = K
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
            meta["small"][name] = {"n": s.n, "nnz": s.nnz, "num_colors": int(col[0]),
                                   "rhs_seed": 11, "apply_seed": 5, "apply_scaling": -0.37,
                                   "iterations": ks}
            print(name, ks, flush=True)
        for name, make in FULL.items():
            s = make()
            b = S.rhs_uniform(s, 2026)
            d = ref.csr_diagonal(s.A, s.C, s.V)
            col = ref.rainbow(s.A, s.C, s.V)
            entry = {"n": s.n, "nnz": s.nnz, "num_colors": int(col[0]), "sha256_F": sha(col[1]),
                     "sha256_B": sha(col[2]), "sha256_R": sha(col[3]),
                     "colour_sizes": np.diff(col[2]).tolist(), "rhs_seed": 2026, "iterations": {}}
            for v in ("rainbow", "ssor"):
                for omega in (1.0, 1.2):
                    x, r, K, t = ref.solve(v, s.A, s.C, s.V, b, aux=d, omega=omega,
                                           colouring=col)
                    entry["iterations"][f"{v}_w{omega}"] = K
                    entry.setdefault("x_norm", {})[f"{v}_w{omega}"] = float(np.linalg.norm(x))
                    print(name, v, omega, K, f"{t[0]:.2f}s", flush=True)
            meta["full"][name] = entry
        hodge_fixtures(ref, meta)
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
