"""The HIP path against the committed golden fixture of the 2D Laplace
ellipse workload (oracle outputs): tree bit-exact, skeleton sets equal,
top-level size equal, density within 1e-10 relative after FP64 GMRES, and
the interior field within 1e-10 of the fixture."""
import numpy as np
import pytest

from conftest import geometry, pkg
from test_golden import load_fixture

pytestmark = pytest.mark.gpu

ki = pkg()
G = geometry()


def test_hip_path_matches_golden_fixture():
    gold = load_fixture()
    pb = G.config(0)
    print("[case] golden c1_laplace2d_ellipse", flush=True)
    ctx = ki.Context()
    ctx.set_geometry(pb.boundary, ki.KI_LAPLACE)
    ctx.build_tree(pb.max_leaf_dofs)
    box, oo, so = gold["box"], gold["original_off"], gold["skel_off"]
    assert ctx.num_levels() == box[:, 0].max() + 1
    for b, (lv, i, leaf, comp) in enumerate(box):
        assert ctx.level_size(lv) > i
        got = ctx.box_list(lv, i, "original")
        assert np.array_equal(got, gold["original"][oo[b]:oo[b + 1]]), (lv, i)
        assert bool(ctx.box(lv, i)["is_leaf"]) == bool(leaf)
    ctx.skeletonize(ki.default_params(id_tol=pb.id_tol,
                                      proxy_points=pb.proxy_points,
                                      proxy_ratio=pb.proxy_ratio))
    for b, (lv, i, leaf, comp) in enumerate(box):
        info = ctx.box(lv, i)
        assert info["compressed"] == bool(comp), (lv, i)
        if comp:
            got = np.sort(ctx.box_list(lv, i, "skel"))
            want = np.sort(gold["skel"][so[b]:so[b + 1]])
            assert np.array_equal(got, want), (lv, i)
    assert ctx.stats()["top_size"] == int(gold["top_size"])
    mu, it, rel = ctx.gmres(pb.rhs, tol=1e-12)
    assert np.linalg.norm(mu - gold["mu"]) <= 1e-10 * np.linalg.norm(
        gold["mu"])
    field = ctx.apply_forward(pb.check_points, mu)
    assert np.max(np.abs(field - gold["field"])) <= 1e-10
