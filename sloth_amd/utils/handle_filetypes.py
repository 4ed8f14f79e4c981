"""VTK XML image files (.vti / .pvti) for electron-density grids, read and written without VTK.

Mirrors ``utils.handle_filetypes`` (src/utils/handle_filetypes.py): ``export_pvti``
(:14-89) writes a 3-D array as the cell data of a vtkImageData file (.vti) plus
the parallel-image file (.pvti) that points at it, and ``pvti_readin`` (:91-120)
reads such a pair back as ``(img, img.shape, spacing)``, the cell array in the
file's Fortran order.  The reference goes through pyvista / vtk, which are not
available here, so both directions implement the VTK XML format directly:

* written: ``ImageData`` with one ``Float32``/``Float64`` cell ``DataArray`` in
  raw appended binary (``header_type="UInt64"``, uncompressed);
* read: ``ascii``, inline ``binary`` (base64) and ``appended`` (``raw`` or
  ``base64``) arrays, ``UInt32``/``UInt64`` headers, uncompressed or
  ``vtkZLibDataCompressor`` blocks, one or several pieces.

``hdf_readin`` (FLASH files through yt) needs yt and h5py and is not provided.
"""
from __future__ import annotations

import base64
import datetime as dt
import os
import re
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_VTK_TYPES = {"Float32": np.float32, "Float64": np.float64, "Int8": np.int8, "UInt8": np.uint8,
              "Int16": np.int16, "UInt16": np.uint16, "Int32": np.int32, "UInt32": np.uint32,
              "Int64": np.int64, "UInt64": np.uint64}
_HEADER_TYPES = {"UInt32": np.uint32, "UInt64": np.uint64}


def _default_fname():
    now = dt.datetime.now()
    return f"./plasma_PVTI_{now.day}_{now.month}_{now.year}_{now.hour}_{now.minute}"


def _write_vti(path, arr, spacing, name="rnec"):
    """One cell array as raw appended binary (UInt64 byte-count header)."""
    spacing = [float(v) for v in spacing]
    a = np.asarray(arr)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    vtype = "Float32" if a.dtype == np.float32 else "Float64"
    nx, ny, nz = a.shape
    data = np.ascontiguousarray(a.flatten(order="F")).astype("<" + a.dtype.str[1:], copy=False)
    raw = data.tobytes()
    rng = (float(np.nanmin(a)), float(np.nanmax(a))) if a.size else (0.0, 0.0)
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        f'  <ImageData WholeExtent="0 {nx} 0 {ny} 0 {nz}" Origin="0 0 0" '
        f'Spacing="{spacing[0]!r} {spacing[1]!r} {spacing[2]!r}" Direction="1 0 0 0 1 0 0 0 1">\n'
        f'    <Piece Extent="0 {nx} 0 {ny} 0 {nz}">\n'
        '      <PointData>\n      </PointData>\n'
        f'      <CellData Scalars="{name}">\n'
        f'        <DataArray type="{vtype}" Name="{name}" format="appended" '
        f'RangeMin="{rng[0]!r}" RangeMax="{rng[1]!r}" offset="0"/>\n'
        '      </CellData>\n    </Piece>\n  </ImageData>\n'
        '  <AppendedData encoding="raw">\n   _')
    with open(path, "wb") as fh:
        fh.write(head.encode())
        fh.write(np.uint64(len(raw)).tobytes())
        fh.write(raw)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")


def export_pvti(arr: np.ndarray, fname: str = None, extent_x=None, extent_y=None, extent_z=None):
    """Write ``arr`` as ``{fname}.vti`` (cell data "rnec") and ``{fname}.pvti``.

    Cell sizes follow the reference: the array spans +-extent along each axis
    (default: half the number of cells), spacing = max(linspace(-e, e, n)) / (n // 2).
    """
    if fname is None:
        fname = _default_fname()
    try:
        np.shape(arr)
        rnec = np.asarray(arr)
    except Exception:
        raise Exception("No electron density currently loaded!")
    shape = rnec.shape
    ext = [e if e is not None else shape[i] // 2
           for i, e in enumerate((extent_x, extent_y, extent_z))]
    cs = [np.linspace(-ext[i], ext[i], shape[i]) for i in range(3)]
    size = [np.max(cs[i]) / (shape[i] // 2) for i in range(3)]
    _write_vti(f"{fname}.vti", rnec, size)
    print(f"wrote {fname}.vti (cell data rnec)")
    _write_pvti(f"{fname}.pvti", shape, size, fname.split("/")[-1] + ".vti",
                "Float32" if rnec.dtype == np.float32 else "Float64")
    print(f"wrote {fname}.pvti (parallel header for {fname}.vti)")


def _write_pvti(path, shape, spacing, source, vtype):
    spacing = [float(v) for v in spacing]
    nx, ny, nz = shape
    content = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="PImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        f'  <PImageData WholeExtent="0 {nx} 0 {ny} 0 {nz}" GhostLevel="0" Origin="0 0 0" '
        f'Spacing="{spacing[0]!r} {spacing[1]!r} {spacing[2]!r}">\n'
        '    <PCellData Scalars="rnec">\n'
        f'      <PDataArray type="{vtype}" Name="rnec"/>\n'
        '    </PCellData>\n'
        f'    <Piece Extent="0 {nx} 0 {ny} 0 {nz}" Source="{source}"/>\n'
        '  </PImageData>\n</VTKFile>\n')
    with open(path, "w") as fh:
        fh.write(content)


# ---------------------------------------------------------------------------
# Reading
# ---------------------------------------------------------------------------

def _b64(s: bytes) -> bytes:
    return base64.b64decode(re.sub(rb"\s+", b"", s))


def _decode_blocks(buf: bytes, htype, compressed: bool, b64: bool):
    """Header + payload of one binary array (inline base64 or appended bytes)."""
    hs = np.dtype(htype).itemsize
    if not compressed:
        if b64:
            # VTK encodes the byte count and the payload separately; a joint
            # encoding decodes the same when the header length is a multiple of 3
            nchar = 4 * ((hs + 2) // 3)
            nbytes = int(np.frombuffer(_b64(buf[:nchar])[:hs], dtype=htype)[0])
            payload = _b64(buf[nchar:])
            if len(payload) < nbytes:
                joint = _b64(buf)
                payload = joint[hs:]
            return payload[:nbytes]
        nbytes = int(np.frombuffer(buf[:hs], dtype=htype)[0])
        return buf[hs:hs + nbytes]
    if b64:
        first = _b64(buf[:4 * ((3 * hs + 2) // 3)])
        nblocks = int(np.frombuffer(first[:hs], dtype=htype)[0])
        hlen = (3 + nblocks) * hs
        nchar = 4 * ((hlen + 2) // 3)
        header = np.frombuffer(_b64(buf[:nchar])[:hlen], dtype=htype)
        body = _b64(buf[nchar:])
    else:
        nblocks = int(np.frombuffer(buf[:hs], dtype=htype)[0])
        hlen = (3 + nblocks) * hs
        header = np.frombuffer(buf[:hlen], dtype=htype)
        body = buf[hlen:]
    sizes = header[3:3 + nblocks].astype(np.int64)
    out, pos = [], 0
    for c in sizes:
        out.append(zlib.decompress(body[pos:pos + c]))
        pos += c
    return b"".join(out)


def _read_vti(path):
    """(cell arrays by name, extent, spacing) of one .vti file."""
    blob = open(path, "rb").read()
    app = blob.find(b"<AppendedData")
    appended = b""
    enc = "raw"
    if app >= 0:
        tag_end = blob.index(b">", app)
        m = re.search(rb'encoding="(\w+)"', blob[app:tag_end])
        enc = m.group(1).decode() if m else "raw"
        start = blob.index(b"_", tag_end) + 1
        end = blob.rfind(b"</AppendedData>")
        appended = blob[start:end]
        xml = blob[:app] + b"</VTKFile>"
    else:
        xml = blob
    root = ET.fromstring(xml)
    htype = _HEADER_TYPES[root.get("header_type", "UInt32")]
    compressed = root.get("compressor") is not None
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise NotImplementedError("big-endian VTK files are not supported")
    img = root.find("ImageData")
    spacing = np.array([float(v) for v in img.get("Spacing", "1 1 1").split()])
    out = []
    for piece in img.findall("Piece"):
        extent = [int(v) for v in piece.get("Extent").split()]
        arrays = {}
        cd = piece.find("CellData")
        for da in (cd.findall("DataArray") if cd is not None else []):
            dtype = _VTK_TYPES[da.get("type")]
            ncomp = int(da.get("NumberOfComponents", "1"))
            fmt = da.get("format", "ascii")
            if fmt == "ascii":
                vals = np.array((da.text or "").split(), dtype=dtype)
            elif fmt == "binary":
                raw = _decode_blocks((da.text or "").strip().encode(), htype, compressed, True)
                vals = np.frombuffer(raw, dtype=np.dtype(dtype).newbyteorder("<"))
            elif fmt == "appended":
                off = int(da.get("offset", "0"))
                raw = _decode_blocks(appended[off:] if enc == "raw" else appended[off:].strip(),
                                     htype, compressed, enc == "base64")
                vals = np.frombuffer(raw, dtype=np.dtype(dtype).newbyteorder("<"))
            else:
                raise ValueError(f"unknown DataArray format {fmt!r}")
            arrays[da.get("Name")] = (vals.astype(dtype), ncomp)
        out.append((extent, arrays))
    return out, spacing


def pvti_readin(filename):
    """Read a .pvti (and the .vti pieces it names): (img, img.shape, spacing).

    img is the first cell array in (nx, ny, nz) order (a trailing component axis
    when it has several), spacing the PImageData spacing.
    """
    root = ET.parse(filename).getroot()
    pimg = root.find("PImageData")
    whole = [int(v) for v in pimg.get("WholeExtent").split()]
    spacing = np.array([float(v) for v in pimg.get("Spacing", "1 1 1").split()])
    dims = [whole[1] - whole[0], whole[3] - whole[2], whole[5] - whole[4]]
    base = os.path.dirname(os.path.abspath(filename))
    img = None
    for piece in pimg.findall("Piece"):
        parts, _ = _read_vti(os.path.join(base, piece.get("Source")))
        for extent, arrays in parts:
            vals, ncomp = next(iter(arrays.values()))
            pd = [extent[1] - extent[0], extent[3] - extent[2], extent[5] - extent[4]]
            shape = pd + ([ncomp] if ncomp > 1 else [])
            block = vals.reshape(shape, order="F")
            if img is None:
                img = np.zeros(dims + ([ncomp] if ncomp > 1 else []), dtype=vals.dtype)
            sl = tuple(slice(extent[2 * a] - whole[2 * a], extent[2 * a + 1] - whole[2 * a])
                       for a in range(3))
            img[sl] = block
    return img, img.shape, spacing
