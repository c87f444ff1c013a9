"""PLY I/O in the reference's layout (scene/beta_model.py:286-404).

No reference PLY file, and no plyfile, is available here, so the layout is
pinned by the PLY specification and the reference's attribute list: the
exact header text plyfile writes for float properties, the body as a
little-endian float32 table in construct_list_of_attributes order, lookup
by name with families ordered by numeric suffix (so shuffled properties,
big-endian and ASCII bodies read back the same), and save/load round trips.
"""
import struct

import numpy as np
import pytest

from gsplat import ply, scenes


def _leaves(D, N=5, seed=0):
    g = np.random.Generator(np.random.PCG64(seed))
    M = D * (D - 1) // 2
    return dict(xyz=g.normal(size=(N, 3)), mean_rest=g.normal(size=(N, D - 3)), scale_raw=g.normal(size=(N, D)),
                l_triangle=g.normal(size=(N, M)), opacity_raw=g.normal(size=(N, 1)), beta_raw=g.normal(size=(N, D - 2)),
                rgb=g.uniform(size=(N, 3)))


def test_attribute_names_match_reference_order():
    # construct_list_of_attributes, scene/beta_model.py:286-296, for D = 6
    want = (["x", "y", "z", "red", "green", "blue", "opacity"] + [f"beta_{i}" for i in range(4)]
            + [f"mean_{i}" for i in range(3)] + [f"scale_{i}" for i in range(6)]
            + [f"l_triangle_{i}" for i in range(15)])
    assert ply.attribute_names(6) == want
    assert len(ply.attribute_names(7)) == 7 + 5 + 4 + 7 + 21


@pytest.mark.parametrize("D", [6, 7])
def test_header_and_body_layout(tmp_path, D):
    lv = _leaves(D, N=3)
    path = str(tmp_path / "s.ply")
    ply.save_ply(path, **lv)
    raw = open(path, "rb").read()
    names = ply.attribute_names(D)
    header = ("ply\nformat binary_little_endian 1.0\nelement vertex 3\n"
              + "".join(f"property float {n}\n" for n in names) + "end_header\n").encode()
    assert raw[:len(header)] == header
    body = np.frombuffer(raw[len(header):], dtype="<f4").reshape(3, len(names))
    order = ["xyz", "rgb", "opacity_raw", "beta_raw", "mean_rest", "scale_raw", "l_triangle"]
    want = np.concatenate([lv[k].astype(np.float32) for k in order], axis=1)
    assert np.array_equal(body, want)
    # the first record, field by field
    assert struct.unpack("<3f", raw[len(header):len(header) + 12]) == tuple(np.float32(lv["xyz"][0]))


@pytest.mark.parametrize("D", [6, 7])
def test_round_trip(tmp_path, D):
    lv = _leaves(D, N=64, seed=D)
    path = str(tmp_path / "a" / "b.ply")  # parent folders are created, as mkdir_p does
    ply.save_ply(path, **lv)
    got = ply.load_ply(path)
    for k, v in lv.items():
        assert got[k].dtype == np.float32
        assert np.array_equal(got[k], v.astype(np.float32)), k


def _write(path, fmt, names, table, types=None):
    types = types or ["float"] * len(names)
    hdr = f"ply\nformat {fmt} 1.0\ncomment made by a test\nelement vertex {table.shape[0]}\n"
    hdr += "".join(f"property {t} {n}\n" for t, n in zip(types, names)) + "end_header\n"
    with open(path, "wb") as f:
        f.write(hdr.encode())
        if fmt == "ascii":
            for row in table:
                f.write((" ".join(repr(float(x)) for x in row) + "\n").encode())
        else:
            order = "<" if fmt == "binary_little_endian" else ">"
            dt = np.dtype([(n, order + ply._PLY_TYPES[t]) for n, t in zip(names, types)])
            rec = np.empty(table.shape[0], dt)
            for i, n in enumerate(names):
                rec[n] = table[:, i]
            f.write(rec.tobytes())


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
def test_shuffled_properties_and_formats(tmp_path, fmt):
    D = 6
    lv = _leaves(D, N=7, seed=3)
    names = ply.attribute_names(D)
    order = ["xyz", "rgb", "opacity_raw", "beta_raw", "mean_rest", "scale_raw", "l_triangle"]
    table = np.concatenate([lv[k] for k in order], axis=1).astype(np.float32)
    perm = np.random.Generator(np.random.PCG64(1)).permutation(len(names))
    # scale_10 style suffixes sort numerically, not as text: 15 l_triangle columns cover it
    types = ["double" if i % 5 == 0 else "float" for i in range(len(names))]
    path = str(tmp_path / "p.ply")
    _write(path, fmt, [names[i] for i in perm], table[:, perm], [types[i] for i in perm])
    got = ply.load_ply(path)
    for k, v in lv.items():
        np.testing.assert_array_equal(got[k], v.astype(np.float32), err_msg=k)


def test_rejects_wrong_column_counts(tmp_path):
    lv = _leaves(6, N=2)
    lv["beta_raw"] = lv["beta_raw"][:, :3]
    with pytest.raises(ValueError):
        ply.save_ply(str(tmp_path / "x.ply"), **lv)
    names = [n for n in ply.attribute_names(6) if n != "l_triangle_14"]
    _write(str(tmp_path / "y.ply"), "binary_little_endian", names, np.zeros((2, len(names)), np.float32))
    with pytest.raises(ValueError):
        ply.load_ply(str(tmp_path / "y.ply"))


def test_benchmark_scene_round_trip(tmp_path):
    sc = scenes.make_scene(1000, 7, seed=2)
    path = str(tmp_path / "scene.ply")
    ply.save_ply(path, sc.xyz, sc.mean_rest, sc.scale_raw, sc.l_triangle, sc.opacity_raw, sc.beta_raw, sc.rgb)
    got = ply.load_ply(path)
    assert np.array_equal(got["scale_raw"], sc.scale_raw) and np.array_equal(got["l_triangle"], sc.l_triangle)
