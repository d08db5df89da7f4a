"""Triangle-mesh input and point-cloud output (engine/mesh_io.py).

``load_mesh`` reads a PLY file (ascii, binary little- or big-endian) with a
small parser of its own in place of the plyfile package, and returns the
(n, 9) float32 triangle array ``MPMSolver.add_mesh`` takes: per face, the
three vertices' x * scale + offset (engine/mesh_io.py:5-24).  Faces must be
triangles, as the reference asserts.  ``write_point_cloud`` writes the
binary PLY of engine/mesh_io.py:27-44.
"""
import numpy as np

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _parse_header(f):
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []  # [name, count, [(prop, type) or (prop, ("list", count_type, item_type))]]
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        words = line.decode("ascii", "replace").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "end_header":
            break
        if words[0] == "format":
            fmt = words[1]
        elif words[0] == "element":
            elements.append([words[1], int(words[2]), []])
        elif words[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if words[1] == "list":
                elements[-1][2].append((words[4], ("list", _PLY_TYPES[words[2]],
                                                   _PLY_TYPES[words[3]])))
            else:
                elements[-1][2].append((words[2], _PLY_TYPES[words[1]]))
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, elements


def _read_binary(buf, pos, elements, endian):
    out = {}
    for name, count, props in elements:
        if all(not isinstance(t, tuple) for _, t in props):
            dt = np.dtype([(p, endian + t) for p, t in props])
            arr = np.frombuffer(buf, dt, count, pos)
            pos += dt.itemsize * count
            out[name] = {p: arr[p] for p, _ in props}
            continue
        cols = {p: [] for p, _ in props}
        for _ in range(count):
            for p, t in props:
                if isinstance(t, tuple):
                    ct = np.dtype(endian + t[1])
                    k = int(np.frombuffer(buf, ct, 1, pos)[0])
                    pos += ct.itemsize
                    it = np.dtype(endian + t[2])
                    cols[p].append(np.frombuffer(buf, it, k, pos))
                    pos += it.itemsize * k
                else:
                    dt = np.dtype(endian + t)
                    cols[p].append(np.frombuffer(buf, dt, 1, pos)[0])
                    pos += dt.itemsize
        out[name] = {p: (v if isinstance(t, tuple) else np.asarray(v, t))
                     for (p, t), v in zip(props, cols.values())}
    return out


def _read_ascii(lines, elements):
    out = {}
    at = 0
    for name, count, props in elements:
        cols = {p: [] for p, _ in props}
        for _ in range(count):
            words = lines[at].split()
            at += 1
            i = 0
            for p, t in props:
                if isinstance(t, tuple):
                    k = int(words[i])
                    cols[p].append(np.asarray(words[i + 1:i + 1 + k], t[2]))
                    i += 1 + k
                else:
                    cols[p].append(words[i])
                    i += 1
        out[name] = {p: (v if isinstance(t, tuple) else np.asarray(v, np.float64).astype(t))
                     for (p, t), v in zip(props, cols.values())}
    return out


def read_ply(fn):
    """{element: {property: array or list of arrays}} of a PLY file."""
    with open(fn, "rb") as f:
        fmt, elements = _parse_header(f)
        body = f.read()
    if fmt == "ascii":
        return _read_ascii(body.decode("ascii").splitlines(), elements)
    return _read_binary(body, 0, elements, "<" if fmt == "binary_little_endian" else ">")


def load_mesh(fn, scale=1, offset=(0, 0, 0)):
    """(n, 9) float32 triangles of a PLY mesh, x * scale + offset per axis
    (engine/mesh_io.py:5-24; the products and sums are taken in float64 and
    stored in float32)."""
    if isinstance(scale, (int, float)):
        scale = (scale, scale, scale)
    data = read_ply(fn)
    vert, faces = data["vertex"], data["face"]
    key = "vertex_indices" if "vertex_indices" in faces else "vertex_index"
    xyz = np.stack([np.asarray(vert[a], np.float64) for a in "xyz"], 1)
    idx = faces[key]
    for face in idx:
        assert len(face) == 3
    idx = np.asarray([np.asarray(f, np.int64) for f in idx]).reshape(-1, 3)
    s = np.asarray(scale, np.float64)
    o = np.asarray(offset, np.float64)
    return (xyz[idx] * s + o).reshape(-1, 9).astype(np.float32)


def write_mesh_ply(fn, vertices, faces):
    """Binary little-endian PLY of a triangle mesh (float32 vertices, int32
    indices): the format ``load_mesh`` reads."""
    vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    faces = np.asarray(faces, np.int32).reshape(-1, 3)
    header = ("ply\nformat binary_little_endian 1.0\n"
              f"element vertex {len(vertices)}\nproperty float x\nproperty float y\n"
              f"property float z\nelement face {len(faces)}\n"
              "property list uchar int vertex_indices\nend_header\n")
    rec = np.zeros(len(faces), np.dtype([("k", "u1"), ("i", "<i4", (3,))]))
    rec["k"] = 3
    rec["i"] = faces
    with open(fn, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(vertices.astype("<f4").tobytes())
        f.write(rec.tobytes())


def write_point_cloud(fn, pos_and_color):
    """engine/mesh_io.py:27-44: binary PLY of (n, 4) float32 rows whose last
    float carries the R, G, B bytes and a placeholder byte."""
    num_particles = len(pos_and_color)
    with open(fn, "wb") as f:
        header = ("ply\nformat binary_little_endian 1.0\ncomment Created by taichi\n"
                  f"element vertex {num_particles}\nproperty float x\nproperty float y\n"
                  "property float z\nproperty uchar red\nproperty uchar green\n"
                  "property uchar blue\nproperty uchar placeholder\nend_header\n")
        f.write(header.encode("ascii"))
        f.write(np.ascontiguousarray(pos_and_color).tobytes())
