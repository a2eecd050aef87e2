"""Write the golden fixture of the 2D Laplace ellipse workload (1,024 GL
nodes, leaf 64, ID tol 1e-10, 64 proxy points at 1.5x) from the oracle.

Run from the repository root:  python tests/golden/make_golden.py

The fixture holds inputs (a hash of the geometry and right-hand side) and
outputs: the tree (level, index, leaf flag, original DOFs of every box),
every compressed box's skeleton, the top-level size, the density mu and the
field at the 100 check points.  The CPU tests check that the oracle still
reproduces it; the GPU tests check the HIP path against it.
"""
import hashlib
import importlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
OUT = Path(__file__).resolve().parent / "c1_laplace2d_ellipse.npz"


def input_hash(pb):
    h = hashlib.sha256()
    b = pb.boundary
    for a in (b.points, b.normals, b.weights, b.curvatures, pb.rhs):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def compute():
    import ki_oracle as O
    G = importlib.import_module("kern-interp-dev_amd.geometry")
    pb = G.config(0)
    b = pb.boundary
    prm = O.Params(id_tol=pb.id_tol, max_leaf_dofs=pb.max_leaf_dofs,
                   proxy_points=pb.proxy_points, proxy_ratio=pb.proxy_ratio)
    kern, tree, fac, mu = O.factor_and_solve(b, O.LAPLACE, prm, pb.rhs)
    box, orig, orig_off, skel, skel_off = [], [], [0], [], [0]
    for lv, nodes in enumerate(tree.levels):
        for i, nd in enumerate(nodes):
            box.append((lv, i, int(nd.is_leaf), int(nd.compressed)))
            orig.extend(nd.original)
            orig_off.append(len(orig))
            skel.extend(nd.skel if nd.compressed else [])
            skel_off.append(len(skel))
    field = kern.apply_forward(pb.check_points, mu)
    return dict(input_sha256=np.array(input_hash(pb)),
                box=np.array(box, dtype=np.int32),
                original=np.array(orig, dtype=np.int32),
                original_off=np.array(orig_off, dtype=np.int64),
                skel=np.array(skel, dtype=np.int32),
                skel_off=np.array(skel_off, dtype=np.int64),
                top_size=np.array(len(fac.allskel), dtype=np.int64),
                mu=np.asarray(mu, dtype=np.float64),
                field=np.asarray(field, dtype=np.float64).ravel(),
                exact=np.asarray(pb.exact, dtype=np.float64).ravel())


if __name__ == "__main__":
    np.savez(OUT, **compute())
    print("wrote", OUT)
