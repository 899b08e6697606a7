"""Restatement of the reference's configuration formats, in Python.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the engine's
format routines (lpmd2_amd/csrc/formats.cpp).  The product package never
imports it.

- the lpmd / zlp format of plugins/lpmd.cc: ReadHeader (:125-169),
  ReadHeaderLPMDOne (:171-210), ReadHeaderZLPOne (:212-226), ReadCell
  (:242-430), WriteHeader (:525-560), WriteCell (:562-640);
- XYZFormat::WriteCell (plugins/xyz.cc:226-249);
- SimulationEngine::Dump / Restore (liblpmd/simulation/simulationbuilder.cc:195-290).

iostream number formats as printf's: std::fixed is %f with the stream's
precision (6 unless set), the default format %g with six significant digits.
Element symbols are ElemSym of liblpmd/support/elements.cc:9-18.
"""
from __future__ import annotations

import math
import struct
import zlib

import numpy as np

ELEM_SYM = (
    "e H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
    "Db Sg Bh Hs Mt Uun Uuu Uub Uut Uuq Uup Uuh Uus Uuo").split()
assert len(ELEM_SYM) == 119


def elem_num(sym: str) -> int:
    """ElemNum (liblpmd/lpmd/elements.h:16-25): first exact match, else 0."""
    try:
        return ELEM_SYM.index(sym)
    except ValueError:
        return 0


def split(line: str):
    """StringSplit(line, ' ') (liblpmd/support/util.cc:32-52): words between
    spaces, tabs, CR and LF."""
    return [w for w in line.replace("\t", " ").replace("\r", " ").replace("\n", " ").split(" ") if w]


def atof(s: str) -> float:
    """C's atof: the longest prefix that parses as a number, else 0."""
    for k in range(len(s), 0, -1):
        try:
            return float(s[:k])
        except ValueError:
            continue
    return 0.0


def _row_len(cell, q):
    r = cell[q]
    return math.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])


def cell_from_lengths(a, b, c, alpha, beta, gamma):
    """NonOrthogonalCell(a, b, c, alpha, beta, gamma) (nonorthogonalcell.h:272-291)."""
    d2r = math.pi / 180.0
    ar, br, gr = alpha * d2r, beta * d2r, gamma * d2r
    t = (math.cos(ar) - math.cos(gr) * math.cos(br)) / math.sin(gr)
    return np.array([[a, 0.0, 0.0], [b * math.cos(gr), b * math.sin(gr), 0.0],
                     [c * math.cos(br), c * t, c * math.sqrt(math.sin(br) * math.sin(br) - t * t)]])


# ---------------------------------------------------------------- reading
def read_lpmd(path):
    """Every frame of an lpmd / zlp file as dict(cell, z, s, v, a, level);
    s are the scaled positions as written."""
    data = open(path, "rb").read()
    pos = 0
    frames = []
    if data[:1] == b"Z":
        # ReadHeaderZLPOne: "Z4L2P", major, minor, revision, two flag bytes
        h = data[:10]
        if h[2:3] != b"L" or h[4:5] != b"P" or h[1:2] != b"4" or h[3:4] != b"2":
            raise ValueError("zlp: Wrong header")
        v0, zlp, hdr, level = h[5], True, None, None
        pos = 10
    elif data[:1] == b"L":
        e = data.index(b"\n")
        words = split(data[:e].decode())
        pos = e + 1
        if not words or words[0] != "LPMD":
            raise ValueError("lpmd: Wrong header (not an LPMD/ZLP file)")
        version = words[1] if len(words) > 1 else ""
        zlp = False
        if len(words) > 2:
            if words[2] == "Z":
                zlp = True
            elif words[2] != "L":
                raise ValueError("lpmd: Wrong compression flag")
        v0 = 0
        if version == "1.0":
            lines = data[pos:].decode().split("\n")
            k = len(split(lines[2])) if len(lines) > 2 else 0
            hdr = ["HDR", "SYM", "X", "Y", "Z"]
            if k == 4:
                level = 0
            elif k == 7:
                hdr += ["VX", "VY", "VZ"]
                level = 1
            elif k == 10:
                hdr += ["VX", "VY", "VZ", "FX", "FY", "FZ"]
                level = 2
            else:
                raise ValueError("lpmd: not an appropriate 1.0 file")
        elif version == "2.0":
            e = data.index(b"\n", pos)
            hdr = split(data[pos:e].decode())
            pos = e + 1
            if not hdr or hdr[0] != "HDR":
                raise ValueError("lpmd: wrong HDR")
            level = 0
        else:
            raise ValueError(f"lpmd: Format version {version} is not supported")
    else:
        raise ValueError("lpmd: Wrong header (not an LPMD/ZLP file)")

    dec = zlib.decompressobj() if zlp else None

    def frame_text():
        # one compressed frame: a width byte, the size (network order, first
        # four bytes of that width), the deflated bytes of the shared stream
        nonlocal pos
        if pos >= len(data):
            return None
        w = data[pos]
        size = struct.unpack(">I", data[pos + 1:pos + 5])[0]
        pos += 1 + w
        chunk = data[pos:pos + size]
        pos += size
        return dec.decompress(chunk).decode()

    if v0 == 1:
        while True:
            t = frame_text()
            if t is None:
                break
            w = split(t)
            lvl, na = int(w[0]), int(w[1])
            cell = np.array([float(q) for q in w[2:11]]).reshape(3, 3)
            p = 11
            z = np.zeros(na, dtype=np.int32)
            s, v, a = (np.zeros((na, 3)) for _ in range(3))
            for i in range(na):
                z[i] = elem_num(w[p])
                p += 1
                s[i] = [float(q) for q in w[p:p + 3]]
                p += 3
                if lvl > 0:
                    v[i] = [float(q) for q in w[p:p + 3]]
                    p += 3
                if lvl > 1:
                    a[i] = [float(q) for q in w[p:p + 3]]
                    p += 3
            frames.append(dict(cell=cell, z=z, s=s, v=v, a=a, level=lvl))
        return frames

    text_lines = None if zlp else data[pos:].decode().split("\n")
    li = 0
    while True:
        if zlp:
            t = frame_text()
            if t is None:
                break
            lines, li = t.split("\n"), 0
        else:
            lines = text_lines
        if li >= len(lines):
            break
        w = split(lines[li])
        li += 1
        if not w:
            break
        na = int(atof(w[0]))
        cw = split(lines[li])
        li += 1
        if len(cw) == 9:
            cell = np.array([atof(q) for q in cw]).reshape(3, 3)
        elif len(cw) == 6:
            cell = cell_from_lengths(*[atof(q) for q in cw])
        else:
            raise ValueError("lpmd: Error ocurred when reading the base vectors")
        z = np.zeros(na, dtype=np.int32)
        s, v, a = (np.zeros((na, 3)) for _ in range(3))
        cols = {"X": (s, 0), "Y": (s, 1), "Z": (s, 2), "VX": (v, 0), "VY": (v, 1), "VZ": (v, 2), "AX": (a, 0),
                "AY": (a, 1), "AZ": (a, 2)}
        for i in range(na):
            aw = split(lines[li]) if li < len(lines) else []
            li += 1
            if not aw:
                raise ValueError("lpmd: the atom file not have elements!")
            for k in range(1, len(hdr)):
                if hdr[k] == "SYM":
                    z[i] = elem_num(aw[k - 1])
                elif hdr[k] in cols:
                    arr, q = cols[hdr[k]]
                    arr[i, q] = atof(aw[k - 1])
        frames.append(dict(cell=cell, z=z, s=s, v=v, a=a, level=level))
    return frames


# ---------------------------------------------------------------- writing
def lpmd_header(zlp: bool, level: int, extra=()):
    cols = ["SYM", "X", "Y", "Z"] + (["VX", "VY", "VZ"] if level >= 1 else []) + \
        (["AX", "AY", "AZ"] if level >= 2 else []) + list(extra)
    return "LPMD 2.0 " + ("Z\n" if zlp else "L\n") + "HDR " + "".join(c + " " for c in cols) + "\n"


def lpmd_frame_text(cell, z, x, v=None, a=None, level=0):
    """WriteCell's text of one frame (orthogonal cell: s = x / |row|)."""
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    out = [f"{len(z)}\n", " ".join("%f" % c for c in cell.reshape(9)) + "\n"]
    L = [_row_len(cell, q) for q in range(3)]
    for i in range(len(z)):
        row = ELEM_SYM[int(z[i])] + " " + " ".join("%f" % (x[i][q] / L[q]) for q in range(3))
        if level >= 1:
            row += " " + " ".join("%f" % (v[i][q] if v is not None else 0.0) for q in range(3))
        if level >= 2:
            row += " " + " ".join("%f" % (a[i][q] if a is not None else 0.0) for q in range(3))
        out.append(row + "\n")
    return "".join(out)


class LpmdWriter:
    """WriteHeader then WriteCell per frame; zlp frames deflated on one stream
    in blocksize pieces, each with a sync flush, with the width byte 8
    (sizeof(unsigned long) on this host) and htonl(size) in 8 bytes."""

    def __init__(self, zlp: bool, level: int = 0, blocksize: int = 1024, compression: int = 6):
        self.zlp, self.level, self.bs = zlp, level, blocksize
        self.out = [lpmd_header(zlp, level).encode()]
        self.comp = zlib.compressobj(compression) if zlp else None

    def frame(self, cell, z, x, v=None, a=None):
        t = lpmd_frame_text(cell, z, x, v, a, self.level).encode()
        if not self.zlp:
            self.out.append(t)
            return
        cb = b""
        for p in range(0, len(t), self.bs):
            cb += self.comp.compress(t[p:p + self.bs]) + self.comp.flush(zlib.Z_SYNC_FLUSH)
        self.out.append(bytes([8]) + struct.pack(">I", len(cb)) + b"\0\0\0\0" + cb)

    def bytes(self):
        return b"".join(self.out)


def xyz_text(cell, z, x, v=None, a=None, level=0, precout=8, inside=False):
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    L = [_row_len(cell, q) for q in range(3)]
    fmt = "%." + str(precout) + "f"
    out = [f"{len(z)}\n", "\n"]
    for i in range(len(z)):
        p = [float(x[i][q]) for q in range(3)]
        if inside:
            # OrthogonalCell::FittedInside (orthogonalcell.h:102-114)
            for q in range(3):
                if p[q] < 0.0 or p[q] > L[q]:
                    p[q] -= L[q] * math.floor(p[q] / L[q])
        row = ELEM_SYM[int(z[i])] + " " + " ".join(fmt % u for u in p)
        if level >= 1:
            row += " " + " ".join(fmt % float(u) for u in v[i])
        if level >= 2:
            row += " " + " ".join(fmt % float(u) for u in a[i])
        out.append(row + "\n")
    return "".join(out)


def dump_text(step, cell, z, x, v, a, tags=None):
    """Dump's text; tags: per atom a list of (key, value)."""
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    out = [f"{step}\n"]
    for r in range(3):
        out.append("".join("%g " % c for c in cell[r]) + "\n")
    out.append(f"{len(z)}\n")
    for i in range(len(z)):
        out.append(f"{int(z[i])} " + "".join("%g " % float(u) for arr in (x, v, a) for u in arr[i]) + "\n")
    out.extend("false \n" for _ in range(len(z)))
    for i in range(len(z)):
        tg = tags[i] if tags is not None else []
        out.append(f"{len(tg)} " + "".join(f"{k} {val} " for k, val in tg) + "\n")
    return "".join(out)
