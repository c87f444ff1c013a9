"""Beta scenes on disk in the reference's PLY layout.

BetaModel.save_ply / load_ply (scene/beta_model.py:286-404) write one
``vertex`` element whose float properties are, in this order,

    x, y, z, red, green, blue, opacity, beta_0..beta_{D-3},
    mean_0..mean_{D-4}, scale_0..scale_{D-1}, l_triangle_0..l_triangle_{M-1}

(M = D(D-1)/2), all pre-activation (the model's raw parameters), through
plyfile's binary little-endian writer.  The loader looks the properties up
by name and orders each family by its numeric suffix, so any property order
reads back the same.  This module reads and writes that layout with numpy
only: binary little/big-endian and ASCII bodies, any numeric property type
(cast to float32, as the reference's ``torch.tensor(..., dtype=float)``).

The benchmark scenes of this package rotate with quaternions, which the
reference layout has no property for; ``save_ply`` stores l_triangle as
given and drops the quaternions, as the reference model has none.
"""
from __future__ import annotations

import os
import re
from typing import Dict, List, Tuple

import numpy as np

# PLY scalar types -> numpy (the spec's names and their sized aliases)
_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def attribute_names(D: int) -> List[str]:
    """Property names of a D-dim scene in the order BetaModel writes them
    (construct_list_of_attributes, scene/beta_model.py:286-296)."""
    names = ["x", "y", "z", "red", "green", "blue", "opacity"]
    names += [f"beta_{i}" for i in range(D - 2)]
    names += [f"mean_{i}" for i in range(D - 3)]
    names += [f"scale_{i}" for i in range(D)]
    names += [f"l_triangle_{i}" for i in range(D * (D - 1) // 2)]
    return names


def _as_np(a) -> np.ndarray:
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float32)


def save_ply(path: str, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, rgb) -> None:
    """Write a scene (leaf parameters, pre-activation; numpy arrays or
    tensors) in the reference's PLY layout: binary little-endian, one float
    property per column (save_ply, scene/beta_model.py:298-320)."""
    cols = [_as_np(a) for a in (xyz, rgb, opacity_raw, beta_raw, mean_rest, scale_raw, l_triangle)]
    N = cols[0].shape[0]
    D = cols[5].shape[1]
    widths = (3, 3, 1, D - 2, D - 3, D, D * (D - 1) // 2)
    for c, w in zip(cols, widths):
        if c.shape != (N, w):
            raise ValueError(f"leaf of shape {c.shape} does not fit a {D}-D scene of {N} primitives ({w} columns)")
    names = attribute_names(D)
    body = np.concatenate(cols, axis=1).astype("<f4", copy=False)
    header = "ply\nformat binary_little_endian 1.0\n" f"element vertex {N}\n"
    header += "".join(f"property float {n}\n" for n in names) + "end_header\n"
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(np.ascontiguousarray(body).tobytes())


def _read_header(f) -> Tuple[str, List[Tuple[str, int, List[Tuple[str, str]]]]]:
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt, elements = None, []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        words = line.decode("ascii").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "format":
            fmt = words[1]
        elif words[0] == "element":
            elements.append((words[1], int(words[2]), []))
        elif words[0] == "property":
            if words[1] == "list":
                raise ValueError("list properties are not part of the Beta scene layout")
            elements[-1][2].append((words[2], _PLY_TYPES[words[1]]))
        elif words[0] == "end_header":
            break
    if fmt not in ("binary_little_endian", "binary_big_endian", "ascii"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, elements


def read_vertices(path: str) -> Dict[str, np.ndarray]:
    """All properties of the first element (the vertices) by name."""
    with open(path, "rb") as f:
        fmt, elements = _read_header(f)
        if not elements:
            raise ValueError("PLY file has no element")
        name, count, props = elements[0]
        if fmt == "ascii":
            rows = [f.readline().split() for _ in range(count)]
            table = np.asarray(rows, dtype=np.float64).reshape(count, len(props))
            return {p: table[:, i].astype(t) for i, (p, t) in enumerate(props)}
        order = "<" if fmt == "binary_little_endian" else ">"
        dt = np.dtype([(p, order + t) for p, t in props])
        data = np.frombuffer(f.read(dt.itemsize * count), dtype=dt, count=count)
        return {p: np.asarray(data[p]) for p, _ in props}


def _family(v: Dict[str, np.ndarray], prefix: str, N: int) -> np.ndarray:
    names = [n for n in v if n.startswith(prefix) and re.fullmatch(r"\d+", n[len(prefix):])]
    names.sort(key=lambda n: int(n[len(prefix):]))
    out = np.zeros((N, len(names)), np.float32)
    for i, n in enumerate(names):
        out[:, i] = v[n]
    return out


def load_ply(path: str) -> Dict[str, np.ndarray]:
    """Leaf parameters (float32, pre-activation) from a PLY in the
    reference's layout, looked up by name as load_ply does
    (scene/beta_model.py:322-404): xyz [N,3], rgb [N,3], opacity_raw [N,1],
    beta_raw [N,D-2], mean_rest [N,D-3], scale_raw [N,D], l_triangle [N,M]."""
    v = read_vertices(path)
    N = len(v["x"])
    out = {
        "xyz": np.stack([v["x"], v["y"], v["z"]], axis=1).astype(np.float32),
        "rgb": np.stack([v["red"], v["green"], v["blue"]], axis=1).astype(np.float32),
        "opacity_raw": np.asarray(v["opacity"], np.float32)[:, None],
        "beta_raw": _family(v, "beta_", N),
        "mean_rest": _family(v, "mean_", N),
        "scale_raw": _family(v, "scale_", N),
        "l_triangle": _family(v, "l_triangle_", N),
    }
    D = out["scale_raw"].shape[1]
    want = {"beta_raw": D - 2, "mean_rest": D - 3, "l_triangle": D * (D - 1) // 2}
    for k, w in want.items():
        if out[k].shape[1] != w:
            raise ValueError(f"{k} has {out[k].shape[1]} columns, a {D}-D scene needs {w}")
    return out
