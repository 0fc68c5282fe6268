"""Tree and field files (annfmp_amd.formats) and the evaluation CLI's host
parts: round trips, the checks that refuse a damaged file, and the mean-L2
score against the oracle's norm_array (the reference's distance,
annfmp/openmp/kdtree/src/base.c norm_array)."""
import json
import os

import numpy as np
import pytest

from conftest import golden, GOLDEN
from annfmp_amd import formats
from annfmp_amd.formats import FormatError
from oracle import oracle as O


def _tree_arrays(name="tree_gauss.npz"):
    g = golden(name)
    return {"axis": g["axis"], "split": g["split"], "leaves": g["leaves"], "xi_orig": g["orig"],
            "xi_pts": g["pts"], "pli": g["pli"], "depth": int(g["depth"])}


def _rewrite(src, dst, **changes):
    with np.load(src, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d.update(changes)
    for k in [k for k, v in changes.items() if v is None]:
        del d[k]
    np.savez(dst, **d)


@pytest.mark.parametrize("name", ["tree_gauss.npz", "tree_ties.npz", "tree_1d.npz"])
def test_tree_round_trip(name, tmp_path):
    arr = _tree_arrays(name)
    path = str(tmp_path / "t.npz")
    formats.save_tree(path, arr, {"leaf_size": 16, "splitting_type": "longest"})
    back = formats.load_tree(path)
    for k in ("axis", "split", "leaves", "xi_orig", "xi_pts", "pli"):
        assert back[k].dtype == np.asarray(arr[k]).dtype
        assert np.array_equal(back[k].view(np.uint32) if k in ("split", "xi_pts") else back[k],
                              np.asarray(arr[k]).view(np.uint32) if k in ("split", "xi_pts") else arr[k])
    assert back["depth"] == arr["depth"]
    assert (back["n"], back["d"]) == arr["xi_pts"].shape
    assert back["params"] == {"leaf_size": 16, "splitting_type": "longest"}


def test_tree_file_checks(tmp_path):
    arr = _tree_arrays()
    good = str(tmp_path / "good.npz")
    formats.save_tree(good, arr)
    bad = str(tmp_path / "bad.npz")
    cases = {
        "version": dict(version=np.array(2, np.int64)),
        "format": dict(format=np.array("annfmp_amd.field")),
        "perm": dict(xi_orig=np.concatenate([arr["xi_orig"][:-1], arr["xi_orig"][:1]]).astype(np.int32)),
        "axis": dict(axis=np.full_like(arr["axis"], arr["xi_pts"].shape[1])),
        "dtype": dict(split=arr["split"].astype(np.float64)),
        "shape": dict(pli=arr["pli"][:-1]),
        "depth": dict(depth=np.array(40, np.int64)),
        "missing": dict(pli=None),
        "untagged": dict(format=None),
    }
    for what, change in cases.items():
        _rewrite(good, bad, **change)
        with pytest.raises(FormatError):
            formats.load_tree(bad)


def test_object_arrays_are_refused(tmp_path):
    """A file holding a pickled object is rejected, never unpickled."""
    path = str(tmp_path / "obj.npz")
    np.savez(path, format=np.array("annfmp_amd.kdtree"), version=np.array(1, np.int64),
             params=np.array([{"x": 1}], dtype=object))
    with pytest.raises(FormatError):
        formats.load_tree(path)


def test_field_round_trip_and_checks(tmp_path):
    rng = np.random.default_rng(3)
    nr, nc, br, bc, K = 7, 9, 6, 11, 3
    nn = rng.integers(0, br * bc, nr * nc).astype(np.int32)
    knn = rng.integers(0, br * bc, (nr * nc, K)).astype(np.int32)
    dist = rng.random(nr * nc).astype(np.float32)
    path = str(tmp_path / "f.npz")
    formats.save_field(path, nn, nr, nc, 5, (br, bc), knn=knn, dist=dist, meta={"method": "plain"})
    f = formats.load_field(path)
    assert np.array_equal(f["nn"], nn) and np.array_equal(f["knn"], knn)
    assert np.array_equal(f["dist"].view(np.uint32), dist.view(np.uint32))
    assert (f["n_rows"], f["n_cols"], f["psize"], f["b_rows"], f["b_cols"]) == (nr, nc, 5, br, bc)
    assert f["meta"] == {"method": "plain"}
    # optional arrays absent
    formats.save_field(path, nn, nr, nc, 5, (br, bc))
    f = formats.load_field(path)
    assert f["knn"] is None and f["dist"] is None
    # an index outside B's grid, a wrong grid, a wrong list shape
    bad = str(tmp_path / "bad.npz")
    formats.save_field(path, nn, nr, nc, 5, (br, bc), knn=knn)
    for change in (dict(nn=np.where(nn == nn.max(), br * bc, nn).astype(np.int32)),
                   dict(n_rows=np.array(nr + 1, np.int64)),
                   dict(knn=knn[:-1]),
                   dict(knn=np.full_like(knn, -2)),
                   dict(version=np.array(0, np.int64))):
        _rewrite(path, bad, **change)
        with pytest.raises(FormatError):
            formats.load_field(bad)
    with pytest.raises(FormatError):
        formats.save_field(path, nn[:-1], nr, nc, 5, (br, bc))


def test_score_distance_matches_norm_array():
    """The CLI's full-patch L2 is the reference's norm_array, bit for bit at
    p = 8 (sums of squares stay below 2^24, so float accumulation is exact)."""
    from annfmp_amd.evaluate import full_patch_l2
    g = golden("example_crop.npz")
    a, b, p = g["a"], g["b"], 8
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    rng = np.random.default_rng(0)
    nn = rng.integers(0, pb.shape[0], pa.shape[0]).astype(np.int32)
    sel = np.arange(0, pa.shape[0], 37)
    got = full_patch_l2(a, b, p, nn, chunk=4096)[sel]
    ref = np.array([O.norm_array(pa[i], pb[nn[i]]) for i in sel], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_cli_arguments_and_images(tmp_path):
    from annfmp_amd import evaluate as E
    with pytest.raises(SystemExit):
        E.main(["only_one.png"])
    with pytest.raises(SystemExit):
        E.main(["a.png", "b.png", "--synthetic", "256x256-shift"])
    g = golden("example_crop.npz")
    npy = str(tmp_path / "a.npy")
    np.save(npy, g["a"])
    assert np.array_equal(E.load_image(npy), g["a"])
    png = str(tmp_path / "a.png")
    from PIL import Image
    Image.fromarray(g["a"]).save(png)
    assert np.array_equal(E.load_image(png), g["a"])
    np.save(npy, g["a"][:, :, 0])
    with pytest.raises(ValueError):
        E.load_image(npy)
    assert E.default_encoder(11, 16).endswith("encoder_p11_d16.npz")
    assert E.default_encoder(5, 16) is None


def test_example_crop_fixture():
    """The real-image pair: two frames of the reference's example, B, G, R."""
    g = golden("example_crop.npz")
    assert g["a"].shape == g["b"].shape == (160, 240, 3) and g["a"].dtype == np.uint8
    assert os.path.exists(os.path.join(GOLDEN, "make_example_crop.py"))
