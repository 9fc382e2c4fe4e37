"""VTK XML PolyData (.vtp) point files without pyvista/VTK.

The reference writes its particle outputs through pyvista
(python_api/lagrbfpy/utils/write_vtp.py:32-50, interface/context.py:459-511:
points, one vertex cell per point, point-data arrays) and reads them back for a
restart (restart.py:12-85).  ``write_vtp`` writes the same content in VTK's
inline "binary" encoding (base64, little-endian, UInt64 byte-count header per
array, no compression); ``read_vtp`` reads the "ascii", "binary" (optionally
zlib-compressed) and "appended" (raw or base64) encodings VTK writers produce.
"""
import base64
import os
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

__all__ = ["write_vtp", "read_vtp"]

_VTK_TYPES = {
    "Float32": np.float32, "Float64": np.float64, "Int8": np.int8, "UInt8": np.uint8, "Int16": np.int16,
    "UInt16": np.uint16, "Int32": np.int32, "UInt32": np.uint32, "Int64": np.int64, "UInt64": np.uint64,
}
_NP_TO_VTK = {np.dtype(v).newbyteorder("<"): k for k, v in _VTK_TYPES.items()}


def _vtk_type(a):
    t = _NP_TO_VTK.get(a.dtype.newbyteorder("<"))
    if t is None:
        raise TypeError(f"array dtype {a.dtype} has no VTK type")
    return t


def _binary_block(a):
    raw = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
    return (base64.b64encode(struct.pack("<Q", len(raw))) + base64.b64encode(raw)).decode("ascii")


def _data_array(a, name, indent):
    a = np.asarray(a)
    ncomp = 1 if a.ndim == 1 else int(np.prod(a.shape[1:]))
    comp = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
    return (f'{indent}<DataArray type="{_vtk_type(a)}" Name="{name}"{comp} format="binary">'
            f"{_binary_block(a.reshape(-1))}</DataArray>\n")


def write_vtp(coords, point_data, filename):
    """Points ``coords`` [n, 2|3] (2-D padded with z = 0) with a vertex cell each
    and the arrays of ``point_data`` (name -> [n] or [n, k]).  ".vtp" is
    appended to ``filename`` when missing; missing directories are created."""
    coords = np.asarray(coords, dtype=np.float64)
    if coords.ndim != 2:
        raise ValueError(f"Expected coords to be a 2D array, got {coords.ndim}D array with shape {coords.shape}")
    if coords.shape[1] == 2:
        coords = np.concatenate((coords, np.zeros((coords.shape[0], 1))), axis=1)
    if coords.shape[1] != 3:
        raise ValueError(f"coords must have 2 or 3 columns, got {coords.shape[1]}")
    n = coords.shape[0]
    if not filename.endswith(".vtp"):
        filename += ".vtp"
    d = os.path.dirname(filename)
    if d and not os.path.exists(d):
        os.makedirs(d)
    parts = ['<?xml version="1.0"?>\n',
             '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
             "  <PolyData>\n",
             f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}" NumberOfLines="0" NumberOfStrips="0" '
             f'NumberOfPolys="0">\n',
             "      <PointData>\n"]
    for name, value in point_data.items():
        v = np.asarray(value)
        if v.shape[0] != n:
            raise ValueError(f"point data '{name}' has {v.shape[0]} rows for {n} points")
        parts.append(_data_array(v, name, "        "))
    parts += ["      </PointData>\n", "      <Points>\n", _data_array(coords, "Points", "        "),
              "      </Points>\n", "      <Verts>\n",
              _data_array(np.arange(n, dtype=np.int64), "connectivity", "        "),
              _data_array(np.arange(1, n + 1, dtype=np.int64), "offsets", "        "),
              "      </Verts>\n", "    </Piece>\n", "  </PolyData>\n", "</VTKFile>\n"]
    with open(filename, "w") as f:
        f.write("".join(parts))
    return filename


def _decode_binary(text, dtype, header, compressed):
    """Inline base64 data: header and data encoded separately (VTK's writers) or
    together; zlib blocks when compressed."""
    text = "".join(text.split()).encode("ascii")
    hs = np.dtype(header).itemsize
    if compressed:
        # [nblocks, block size, last block size, compressed sizes...] then the blocks
        hlen = lambda nw: ((nw * hs + 2) // 3) * 4
        nb = int(np.frombuffer(base64.b64decode(text[:hlen(1)])[:hs], dtype=header)[0])
        head = np.frombuffer(base64.b64decode(text[:hlen(3 + nb)]), dtype=header)
        sizes = head[3:3 + nb].astype(np.int64)
        raw = base64.b64decode(text[hlen(3 + nb):])
        out, off = [], 0
        for s in sizes:
            out.append(zlib.decompress(raw[off:off + s]))
            off += s
        return np.frombuffer(b"".join(out), dtype=dtype)
    hl = ((hs + 2) // 3) * 4
    nbytes = int(np.frombuffer(base64.b64decode(text[:hl])[:hs], dtype=header)[0])
    data = base64.b64decode(text[hl:])
    if len(data) != nbytes:          # header encoded together with the data
        joint = base64.b64decode(text)
        nbytes = int(np.frombuffer(joint[:hs], dtype=header)[0])
        data = joint[hs:hs + nbytes]
    return np.frombuffer(data[:nbytes], dtype=dtype)


def _decode_appended(blob, offset, end, dtype, header, compressed, encoding):
    """Array at ``offset`` of the appended section; with base64 encoding the run
    ends at the next array's offset (``end``)."""
    if encoding == "base64":
        return _decode_binary(blob[offset:end].decode("ascii"), dtype, header, compressed)
    hs = np.dtype(header).itemsize
    if compressed:
        nb = int(np.frombuffer(blob[offset:offset + hs], dtype=header)[0])
        head = np.frombuffer(blob[offset:offset + hs * (3 + nb)], dtype=header)
        sizes = head[3:3 + nb].astype(np.int64)
        off = offset + hs * (3 + nb)
        out = []
        for s in sizes:
            out.append(zlib.decompress(blob[off:off + s]))
            off += s
        return np.frombuffer(b"".join(out), dtype=dtype)
    nbytes = int(np.frombuffer(blob[offset:offset + hs], dtype=header)[0])
    return np.frombuffer(blob[offset + hs:offset + hs + nbytes], dtype=dtype)


def read_vtp(filename):
    """(points [n, 3] float64, point_data dict name -> array [n] or [n, k])."""
    with open(filename, "rb") as f:
        content = f.read()
    appended = b""
    start = content.find(b"<AppendedData")
    if start >= 0:
        us = content.index(b"_", content.index(b">", start)) + 1
        end = content.rfind(b"</AppendedData>")
        appended = content[us:end]
        content = content[:content.index(b">", start) + 1] + b"</AppendedData></VTKFile>"
    root = ET.fromstring(content)
    if root.get("type") != "PolyData":
        raise ValueError(f"{filename}: not a PolyData file ({root.get('type')})")
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError(f"{filename}: big-endian VTK files are not supported")
    header = np.dtype("<u8" if root.get("header_type", "UInt32") == "UInt64" else "<u4")
    compressed = root.get("compressor") is not None
    app = root.find("AppendedData")
    encoding = app.get("encoding", "raw") if app is not None else "raw"
    if encoding == "base64" and appended:
        appended = appended.strip()
    offsets = sorted({int(el.get("offset")) for el in root.iter("DataArray") if el.get("format") == "appended"})

    def array(el):
        dtype = np.dtype(_VTK_TYPES[el.get("type")]).newbyteorder("<")
        fmt = el.get("format", "ascii")
        if fmt == "ascii":
            a = np.array((el.text or "").split(), dtype=dtype)
        elif fmt == "binary":
            a = _decode_binary(el.text or "", dtype, header, compressed)
        elif fmt == "appended":
            off = int(el.get("offset"))
            later = [o for o in offsets if o > off]
            end = later[0] if later else len(appended)
            a = _decode_appended(appended, off, end, dtype, header, compressed, encoding)
        else:
            raise ValueError(f"{filename}: unknown DataArray format {fmt}")
        a = a.astype(dtype.newbyteorder("="))
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    piece = root.find("PolyData/Piece")
    if piece is None:
        raise ValueError(f"{filename}: no PolyData piece")
    pts_el = piece.find("Points/DataArray")
    points = array(pts_el).astype(np.float64).reshape(-1, 3) if pts_el is not None else np.zeros((0, 3))
    data = {}
    pd = piece.find("PointData")
    if pd is not None:
        for el in pd.findall("DataArray"):
            data[el.get("Name")] = array(el)
    return points, data
