"""Tree files through annf_tree_import, and the evaluation CLI end to end
(GPU): a restored tree answers kneighbors exactly as the fitted one (and as
the reference's golden vectors); a damaged tree is refused before upload;
a field written by the CLI re-evaluates to the same numbers."""
import json

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _fitted(name):
    from annfmp_amd import KDTree
    g = golden(name)
    t = KDTree(n_neighbors=int(g["K"]), leaf_size=2, tree_depth=int(g["depth"]), float_type="float",
               splitting_type="longest")
    return g, t.fit(g["X"])


@pytest.mark.parametrize("name", ["tree_gauss.npz", "tree_dups.npz"])
def test_saved_tree_answers_like_the_fitted_one(name, tmp_path, gpu_available):
    from annfmp_amd import KDTree
    g, t = _fitted(name)
    path = str(tmp_path / "tree.npz")
    t.save(path)
    u = KDTree.load(path)
    a, b = t.tree_arrays(), u.tree_arrays()
    for k in ("axis", "leaves", "xi_orig", "pli"):
        assert np.array_equal(a[k], b[k])
    assert np.array_equal(bits(a["split"]), bits(b["split"]))
    assert np.array_equal(bits(a["xi_pts"]), bits(b["xi_pts"]))
    assert np.array_equal(bits(u.X), bits(g["X"].astype(np.float32)))
    d, idx = u.kneighbors(g["Q"])
    assert np.array_equal(idx, g["neighbors_idx"])
    assert np.array_equal(bits(d), bits(g["neighbors_dist"]))
    assert u.get_params() == t.get_params()


def test_damaged_tree_is_refused(tmp_path, gpu_available):
    from annfmp_amd import KDTree, formats
    from annfmp_amd._lib import ANNFError
    g, t = _fitted("tree_gauss.npz")
    arr = t.tree_arrays()
    path = str(tmp_path / "bad.npz")
    pli = arr["pli"].copy()
    pli[5] = (pli[5] + 1) % (1 << arr["depth"])
    leaves = arr["leaves"].copy()
    leaves[1] += 1
    leaves[2] += 1
    for change in (dict(pli=pli), dict(leaves=leaves)):
        formats.save_tree(path, dict(arr, **change), t.get_params())
        with pytest.raises(ANNFError):
            KDTree.load(path)


@pytest.mark.parametrize("method,extra", [
    ("plain", ["--psize", "7", "--dim", "10", "--leaf", "32", "--neighbors", "3"]),
    ("pca-bf16", ["--psize", "7", "--dim", "10", "--leaf", "32", "--neighbors", "3"]),
    ("learned-bf16", ["--psize", "11", "--dim", "16", "--leaf", "64", "--neighbors", "3"]),
])
def test_cli_field_file_reevaluates_the_same(method, extra, tmp_path, capsys, gpu_available):
    from annfmp_amd import evaluate as E, formats, synth
    out = str(tmp_path / "field.npz")
    common = ["--synthetic", "256x256-shift", "--seed", "0", "--sample", "2048"]
    assert E.main(common + ["--method", method, "--out", out] + extra) == 0
    r1 = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert E.main(common + ["--field", out, "--k", "3"]) == 0
    r2 = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    for k in ("recall@3", "distance_ratio", "score", "sample"):
        assert r1[k] == r2[k]
    assert 0.0 <= r1["recall@3"] <= 1.0 and r1["distance_ratio"] >= 1.0
    # the field's distances are the full-patch L2 of its matches, as the host recomputes them
    f = formats.load_field(out)
    a, b = synth.make_pair("256x256-shift", seed=0)
    assert np.array_equal(bits(f["dist"]), bits(E.full_patch_l2(a, b, f["psize"], f["nn"])))
    assert f["meta"]["method"] == method
