"""VTK XML ImageData (.vti) codec for NS3D images and fields.

Covers the subset the reference reads and writes (vtkio.cxx:38-292): one
appended DataArray, ``encoding="base64"``, ``header_type="UInt32"``, zlib
compression (``vtkZLibDataCompressor``) or raw, scalar UInt8 images and
Float64 scalar or 3-component fields.  VTK itself is not required.

Axis convention: the file stores x fastest.  Arrays returned here are
C-ordered with shape ``(nx, ny, nz)`` -- ``a[i, j, k]`` is voxel (x=i, y=j,
z=k), the Blitz++ layout the reference fills at vtkio.cxx:133-143.
"""
from __future__ import annotations

import base64
import re
import struct
import zlib

import numpy as np

_DTYPES = {
    "UInt8": np.uint8, "Int8": np.int8, "UInt16": np.uint16, "Int16": np.int16,
    "UInt32": np.uint32, "Int32": np.int32, "Float32": np.float32, "Float64": np.float64,
}
_NAMES = {np.dtype(v): k for k, v in _DTYPES.items()}


def _attr(tag: str, name: str) -> str | None:
    m = re.search(r'\b' + name + r'="([^"]*)"', tag)
    return m.group(1) if m else None


def read_vti(path: str) -> np.ndarray:
    """Read the first DataArray of a .vti file; returns (nx, ny, nz[, ncomp])."""
    with open(path, "rb") as f:
        raw = f.read()
    head_end = raw.index(b"<AppendedData")
    head = raw[:head_end].decode("latin-1")
    ext = [int(t) for t in _attr(head, "WholeExtent").split()]
    nx, ny, nz = ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1
    da = re.search(r"<DataArray[^>]*>", head).group(0)
    dtype = np.dtype(_DTYPES[_attr(da, "type")])
    ncomp = int(_attr(da, "NumberOfComponents") or 1)
    compressed = "vtkZLibDataCompressor" in head
    htype = _attr(head, "header_type") or "UInt32"
    hsize = 8 if htype == "UInt64" else 4
    hfmt = "<Q" if hsize == 8 else "<I"
    app = raw[head_end:]
    app = app[app.index(b">") + 1:]
    start = app.index(b"_") + 1
    end = app.index(b"</AppendedData>")
    b64 = b"".join(app[start:end].split())
    if compressed:
        first = base64.b64decode(b64[:12])  # 9 bytes: enough for the block count
        nblocks = struct.unpack(hfmt, first[:hsize])[0]
        hbytes = (3 + nblocks) * hsize
        hchars = 4 * ((hbytes + 2) // 3)
        header = base64.b64decode(b64[:hchars])
        vals = struct.unpack("<" + ("Q" if hsize == 8 else "I") * (3 + nblocks), header[:hbytes])
        nb, bsize, last = vals[0], vals[1], vals[2]
        csizes = vals[3:3 + nb]
        data = base64.b64decode(b64[hchars:])
        out = bytearray()
        off = 0
        for cs in csizes:
            out += zlib.decompress(data[off:off + cs])
            off += cs
        payload = bytes(out)
        expect = (nb - 1) * bsize + (last if last else bsize) if nb else 0
        payload = payload[:expect]
    else:
        blob = base64.b64decode(b64)
        n = struct.unpack(hfmt, blob[:hsize])[0]
        payload = blob[hsize:hsize + n]
    arr = np.frombuffer(payload, dtype=dtype)
    shape = (nz, ny, nx) if ncomp == 1 else (nz, ny, nx, ncomp)
    arr = arr.reshape(shape)
    # file is x-fastest: arr[z, y, x] -> return [x, y, z]
    if ncomp == 1:
        return np.ascontiguousarray(arr.transpose(2, 1, 0))
    return np.ascontiguousarray(arr.transpose(2, 1, 0, 3))


def write_vti(path: str, a: np.ndarray, name: str = "ImageScalars", compress: bool = True,
              block: int = 32768) -> None:
    """Write a (nx, ny, nz[, ncomp]) array as appended base64 (zlib) ImageData."""
    a = np.asarray(a)
    ncomp = 1 if a.ndim == 3 else a.shape[3]
    nx, ny, nz = a.shape[:3]
    dt = a.dtype
    if dt not in _NAMES:
        raise ValueError(f"unsupported dtype {dt}")
    if ncomp == 1:
        payload = np.ascontiguousarray(a.transpose(2, 1, 0)).tobytes()
    else:
        payload = np.ascontiguousarray(a.transpose(2, 1, 0, 3)).tobytes()
    if compress:
        blocks = [payload[o:o + block] for o in range(0, len(payload), block)] or [b""]
        comp = [zlib.compress(b) for b in blocks]
        last = len(blocks[-1]) if len(blocks[-1]) != block else 0
        header = struct.pack("<" + "I" * (3 + len(comp)), len(comp), block, last,
                             *[len(c) for c in comp])
        body = base64.b64encode(header) + base64.b64encode(b"".join(comp))
        comp_attr = ' compressor="vtkZLibDataCompressor"'
    else:
        body = base64.b64encode(struct.pack("<I", len(payload)) + payload)
        comp_attr = ""
    rng = (float(a.min()), float(a.max())) if a.size else (0.0, 0.0)
    nc = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
    xml = (
        '<?xml version="1.0"?>\n'
        f'<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" '
        f'header_type="UInt32"{comp_attr}>\n'
        f'  <ImageData WholeExtent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}" Origin="0 0 0" '
        f'Spacing="1 1 1" Direction="1 0 0 0 1 0 0 0 1">\n'
        f'    <Piece Extent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}">\n'
        f'      <PointData Scalars="{name}">\n'
        f'        <DataArray type="{_NAMES[dt]}" Name="{name}"{nc} format="appended" '
        f'RangeMin="{rng[0]:.10g}" RangeMax="{rng[1]:.10g}" offset="0"/>\n'
        '      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n'
        '  </ImageData>\n  <AppendedData encoding="base64">\n   _'
    )
    with open(path, "wb") as f:
        f.write(xml.encode())
        f.write(body)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
