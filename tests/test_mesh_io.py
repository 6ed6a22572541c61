"""Frame writers (reference: ChysX/src/io/obj_io.{h,cpp},
ChysX/src/io/bgeo_writer.{h,cpp}).  The .bgeo output is decoded by an
independent reader of the binary-JSON token stream written here from the
format's token table; OBJ round-trips; the box mesh is checked for closure
and outward winding."""
import struct

import numpy as np
import pytest


def _io():
    import chysx

    return chysx.io


# ---- independent binary-JSON reader ---------------------------------------
class _Reader:
    ARRAY, ARRAY_END, MAP, MAP_END, STRING, INT32, REAL32, MAGIC, UNIFORM = (
        0x5B, 0x5D, 0x7B, 0x7D, 0x27, 0x13, 0x19, 0x7F, 0x40)

    def __init__(self, data: bytes):
        self.d, self.i = data, 0

    def _u8(self):
        v = self.d[self.i]
        self.i += 1
        return v

    def _take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.d, self.i)
        self.i += struct.calcsize("<" + fmt)
        return v[0]

    def _length(self):
        b = self._u8()
        if b < 0xF1:
            return b
        return {0xF2: lambda: self._take("H"), 0xF4: lambda: self._take("I"), 0xF8: lambda: self._take("q")}[b]()

    def value(self):
        t = self._u8()
        if t == self.ARRAY:
            out = []
            while self.d[self.i] != self.ARRAY_END:
                out.append(self.value())
            self.i += 1
            return out
        if t == self.MAP:
            out = {}
            while self.d[self.i] != self.MAP_END:
                k = self.value()
                out[k] = self.value()
            self.i += 1
            return out
        if t == self.STRING:
            n = self._length()
            s = self.d[self.i:self.i + n].decode()
            self.i += n
            return s
        if t == self.INT32:
            return self._take("i")
        if t == self.UNIFORM:
            kind = self._u8()
            n = self._length()
            dt = {self.REAL32: np.float32, self.INT32: np.int32}[kind]
            a = np.frombuffer(self.d, dt, n, self.i).copy()
            self.i += 4 * n
            return a
        raise ValueError(f"unexpected token 0x{t:02x} at {self.i - 1}")

    def document(self):
        assert self._u8() == self.MAGIC
        assert self._take("I") == 0x624A534E
        doc = self.value()
        assert self.i == len(self.d)
        return doc


def _pairs(flat):
    return {flat[k]: flat[k + 1] for k in range(0, len(flat), 2)}


def _attribute(attr):
    definition, data = attr
    d, v = _pairs(definition), _pairs(data)
    values = _pairs(v["values"])
    return d, v, values


def _decode(path):
    with open(path, "rb") as f:
        return _pairs(_Reader(f.read()).document())


def _mesh(n_side=5, seed=0):
    rng = np.random.default_rng(seed)
    n = n_side * n_side
    pos = rng.normal(size=(n, 3)).astype(np.float32)
    tris = []
    for y in range(n_side - 1):
        for x in range(n_side - 1):
            v0 = y * n_side + x
            tris += [(v0, v0 + 1, v0 + n_side), (v0 + 1, v0 + n_side + 1, v0 + n_side)]
    return pos, np.asarray(tris, np.int32)


def test_bgeo_document(tmp_path):
    pos, tris = _mesh()
    pc = np.random.default_rng(1).random((len(pos), 3)).astype(np.float32)
    fc = np.random.default_rng(2).random((len(tris), 3)).astype(np.float32)
    f = str(tmp_path / "m.bgeo")
    _io().write_bgeo(f, pos, tris, pc, fc)
    g = _decode(f)
    assert g["pointcount"] == len(pos) and g["vertexcount"] == 3 * len(tris) and g["primitivecount"] == len(tris)
    topo = _pairs(g["topology"])
    assert np.array_equal(_pairs(topo["pointref"])["indices"], tris.ravel())
    attrs = _pairs(g["attributes"])
    P, Cd = attrs["pointattributes"]
    d, v, vals = _attribute(P)
    assert d["name"] == "P" and d["scope"] == "public" and d["type"] == "numeric"
    assert d["options"] == {"type": {"type": "string", "value": "point"}}
    assert v["size"] == 3 and v["storage"] == "fpreal32" and vals["pagesize"] == 1024
    assert np.array_equal(vals["rawpagedata"], pos.ravel())
    d, _, vals = _attribute(Cd)
    assert d["name"] == "Cd" and d["options"] == {} and np.array_equal(vals["rawpagedata"], pc.ravel())
    (pcd,) = attrs["primitiveattributes"]
    assert np.array_equal(_attribute(pcd)[2]["rawpagedata"], fc.ravel())
    ((kind, run),) = g["primitives"]
    assert _pairs(kind) == {"type": "p_r"}
    run = _pairs(run)
    assert run["s_v"] == 0 and run["n_p"] == len(tris) and np.all(run["n_v"] == 3)


def test_bgeo_long_arrays_and_multi(tmp_path):
    """Arrays past 0xFFFF elements take the 32-bit length form; pieces merge
    with offset indices and per-piece point colours."""
    io = _io()
    pos, tris = _mesh(n_side=160)            # 25,600 points -> 76,800 floats
    bp, bt = io.box_mesh(np.array([0.0, 0.0, 0.5], np.float32), np.array([0.2, 0.2, 0.05], np.float32))
    f = str(tmp_path / "multi.bgeo")
    io.write_bgeo_multi(f, [(pos, tris, (1.0, 0.0, 0.0)), (bp, bt, (0.0, 0.5, 1.0))])
    g = _decode(f)
    assert g["pointcount"] == len(pos) + 8 and g["primitivecount"] == len(tris) + 12
    idx = _pairs(_pairs(g["topology"])["pointref"])["indices"]
    assert np.array_equal(idx, np.concatenate([tris.ravel(), bt.ravel() + len(pos)]))
    P, Cd = _pairs(g["attributes"])["pointattributes"]
    assert np.array_equal(_attribute(P)[2]["rawpagedata"], np.concatenate([pos.ravel(), bp.ravel()]))
    cd = _attribute(Cd)[2]["rawpagedata"].reshape(-1, 3)
    assert np.all(cd[:len(pos)] == [1, 0, 0]) and np.all(cd[len(pos):] == np.float32([0, 0.5, 1]))
    assert "primitiveattributes" not in _pairs(g["attributes"])


def test_obj_round_trip(tmp_path):
    io = _io()
    pos, tris = _mesh(seed=3)
    f = str(tmp_path / "m.obj")
    io.save_obj(f, pos, tris)
    lines = open(f).read().splitlines()
    assert lines[0].startswith("#") and sum(l.startswith("v ") for l in lines) == len(pos)
    p2, t2 = io.load_obj(f)
    assert np.array_equal(t2, tris)
    assert np.allclose(p2, pos, rtol=1e-5, atol=1e-6)  # text with 6 significant digits


def test_obj_reader_face_forms(tmp_path):
    f = tmp_path / "faces.obj"
    f.write_text("# c\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f 1/1/1 2/1/1 3/1/1\nf 2//1 4//1 3//1\nf 1 2 4 3\n")
    p, t = _io().load_obj(str(f))
    assert p.shape == (4, 3)
    assert t.tolist() == [[0, 1, 2], [1, 3, 2], [0, 1, 3]]  # a polygon line gives its first three
    with pytest.raises(RuntimeError):
        _io().load_obj(str(tmp_path / "missing.obj"))


def test_box_mesh_is_closed_and_outward():
    c = np.array([0.3, -0.2, 0.5], np.float32)
    h = np.array([0.4, 0.1, 0.25], np.float32)
    p, t = _io().box_mesh(c, h)
    assert np.allclose(np.sort(np.abs(p - c), axis=0), np.tile(h, (8, 1)))
    area = 0.0
    edges = {}
    for a, b, cc in t:
        n = np.cross(p[b] - p[a], p[cc] - p[a])
        assert np.dot(n, (p[a] + p[b] + p[cc]) / 3 - c) > 0   # outward winding
        area += 0.5 * np.linalg.norm(n)
        for e in ((a, b), (b, cc), (cc, a)):
            edges[e] = edges.get(e, 0) + 1
    assert area == pytest.approx(8 * (h[0] * h[1] + h[1] * h[2] + h[0] * h[2]), rel=1e-6)
    # closed, consistently oriented: every directed edge appears once, and
    # its reverse once
    assert all(v == 1 for v in edges.values()) and all((b, a) in edges for a, b in edges)
