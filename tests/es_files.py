"""Independent readers of libMesh EquationSystems files for the tests.

ASCII (.xda): one datum per line, "value<TAB> # comment"; value streams are
space-separated on one line that ends with the vector's comment.  XDR (.xdr):
RFC 4506 -- big-endian 4-byte unsigned/int (unsigned short widened to 4
bytes), 8-byte IEEE doubles, strings as a 4-byte length plus the bytes padded
to a multiple of 4, vectors as a 4-byte length plus elements, value streams
without a length, no comments.  Layout (both): version; system count; per
system: name, type, variable count, per variable name, subdomain list, order,
family; vector count, per vector name, projection flag, parallel type; then per
system the solution and each vector as a length plus node-major values
(src/systems/equation_systems_io.C:335-560, src/systems/system_io.C:1117-1300,
1518-1580, 1968-2010)."""
import struct

import numpy as np


def _val(line):
    return line.split("\t")[0]


def parse_xda(path):
    lines = open(path).read().split("\n")
    pos = [0]

    def line():
        s = lines[pos[0]]
        pos[0] += 1
        return s

    out = {"version": _val(line()), "systems": []}
    for _ in range(int(_val(line()))):
        nl = line()
        s = {"name": _val(nl), "number": int(nl.rsplit("No. ", 1)[1]), "type": _val(line()), "vars": [],
             "vectors": []}
        for _ in range(int(_val(line()))):
            name = _val(line())
            nsub = int(_val(line()))
            subs = [int(t) for t in line().split("#")[0].split()]
            assert len(subs) == nsub
            s["vars"].append((name, subs, int(_val(line())), int(_val(line()))))
        for _ in range(int(_val(line()))):
            s["vectors"].append((_val(line()), int(_val(line())), int(_val(line()))))
        out["systems"].append(s)
    for s in out["systems"]:
        s["data"] = []
        for k in range(1 + len(s["vectors"])):
            n = int(_val(line()))
            data_line = line()
            toks = data_line.split("\t")[0].split()
            assert len(toks) == n, (s["name"], k, n, len(toks))
            s["data"].append(np.array([float(t) for t in toks]))
            s.setdefault("comments", []).append(data_line.split("\t", 1)[1])
    assert all(not x for x in lines[pos[0]:]), "trailing content"
    return out


class _XdrReader:
    def __init__(self, b):
        self.b, self.p = b, 0

    def u32(self):
        (v,) = struct.unpack_from(">I", self.b, self.p)
        self.p += 4
        return v

    def i32(self):
        (v,) = struct.unpack_from(">i", self.b, self.p)
        self.p += 4
        return v

    def string(self):
        n = self.u32()
        s = self.b[self.p:self.p + n].decode()
        pad = (4 - n % 4) % 4
        assert self.b[self.p + n:self.p + n + pad] == b"\0" * pad
        self.p += n + pad
        return s

    def doubles(self, n):
        v = np.frombuffer(self.b, dtype=">f8", count=n, offset=self.p).astype(np.float64)
        self.p += 8 * n
        return v


def parse_xdr(path):
    r = _XdrReader(open(path, "rb").read())
    out = {"version": r.string(), "systems": []}
    for i in range(r.u32()):
        s = {"name": r.string(), "number": i, "type": r.string(), "vars": [], "vectors": []}
        for _ in range(r.u32()):
            name = r.string()
            subs = [r.u32() for _ in range(r.u32())]
            s["vars"].append((name, subs, r.i32(), r.i32()))
        for _ in range(r.u32()):
            s["vectors"].append((r.string(), r.i32(), r.i32()))
        out["systems"].append(s)
    for s in out["systems"]:
        s["data"] = [r.doubles(r.u32()) for _ in range(1 + len(s["vectors"]))]
    assert r.p == len(r.b), "trailing bytes"
    return out
