"""Grid geometry: seeded Voronoi tessellations, random orientations and .vti I/O.

Restates the reference's input generators used to build the benchmark inputs:

* ``damask.seeds.from_random`` (``python/damask/seeds.py:58-61``): seed
  coordinates ``default_rng(seed).random((N, 3)) * size``;
* ``damask.GeomGrid.from_Voronoi_tessellation`` (``python/damask/_geomgrid.py:631-637``):
  periodic nearest seed at the cell centres of ``grid.coordinates0_point``
  (``python/damask/grid.py:217-224``);
* ``damask.Rotation.from_random`` (``python/damask/_rotation.py:1278-1282``):
  normalised Gaussian quaternions with non-negative real part;
* the VTK ImageData (.vti) format read by ``src/grid/VTI.cpp`` and written by
  ``GeomGrid.save`` (``python/damask/_geomgrid.py:786-794``): cell data
  ``material`` flattened x fastest, base64 blocks, optional zlib compression
  with a UInt32 or UInt64 header.
"""
from __future__ import annotations

import base64
import dataclasses
import re
import struct
import zlib
import xml.etree.ElementTree as ET

import numpy as np


@dataclasses.dataclass
class Grid:
    cells: np.ndarray          # (3,) int
    size: np.ndarray           # (3,) float
    origin: np.ndarray         # (3,) float
    material: np.ndarray       # (N1*N2*N3,) int64, 0-based, x fastest

    @property
    def n_cells(self) -> int:
        return int(np.prod(self.cells))


def seeds_from_random(size, N_seeds: int, rng=None) -> np.ndarray:
    rng = np.random.default_rng(rng)
    return rng.random((N_seeds, 3)) * np.asarray(size, float)


def coordinates0_point(cells, size, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    size = np.asarray(size, float)
    cells = np.asarray(cells, np.int64)
    start = np.asarray(origin, float) + size / cells * 0.5
    end = np.asarray(origin, float) + size - size / cells * 0.5
    return np.stack(np.meshgrid(np.linspace(start[0], end[0], cells[0]),
                                np.linspace(start[1], end[1], cells[1]),
                                np.linspace(start[2], end[2], cells[2]), indexing="ij"), axis=-1)


def voronoi_tessellation(cells, size, seeds, periodic: bool = True) -> Grid:
    """Periodic Voronoi map; material IDs are the seed indices (0-based)."""
    cells = np.asarray(cells, np.int64)
    size = np.asarray(size, float)
    coords = coordinates0_point(cells, size).reshape(-1, 3)
    seeds = np.asarray(seeds, float)
    best = np.zeros(len(coords), np.int64)
    chunk = 1 << 15
    for s in range(0, len(coords), chunk):
        c = coords[s:s + chunk, None, :] - seeds[None, :, :]
        if periodic:
            c = c - size * np.round(c / size)
        best[s:s + chunk] = np.argmin(np.einsum("ijk,ijk->ij", c, c), axis=1)
    mat = best.reshape(tuple(cells))          # [x, y, z]
    return Grid(cells, size, np.zeros(3), mat.flatten(order="F"))


def rotation_from_random(n: int, rng=None) -> np.ndarray:
    rng = np.random.default_rng(rng)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    q[q[:, 0] < 0.0] *= -1.0
    return q


# ---------------------------------------------------------------------------- VTI
_DTYPES = {"Int8": np.int8, "UInt8": np.uint8, "Int16": np.int16, "UInt16": np.uint16,
           "Int32": np.int32, "UInt32": np.uint32, "Int64": np.int64, "UInt64": np.uint64,
           "Float32": np.float32, "Float64": np.float64}


def _decode(text: str, dtype, header_t, compressed: bool) -> np.ndarray:
    hsz = np.dtype(header_t).itemsize
    text = re.sub(r"\s+", "", text)
    if not compressed:
        raw = base64.b64decode(text)
        n = int(np.frombuffer(raw[:hsz], header_t)[0])
        return np.frombuffer(raw[hsz:hsz + n], dtype).copy()
    # header: n_blocks, block size, last block size, compressed sizes; its own base64 run
    head0 = base64.b64decode(text[:(((3 * hsz) + 2) // 3) * 4])
    nb = int(np.frombuffer(head0[:hsz], header_t)[0])
    hlen = (3 + nb) * hsz
    hchars = ((hlen + 2) // 3) * 4
    head = np.frombuffer(base64.b64decode(text[:hchars])[:hlen], header_t)
    csz = head[3:3 + nb].astype(np.int64)
    payload = base64.b64decode(text[hchars:])
    out, off = [], 0
    for k in range(nb):
        out.append(zlib.decompress(payload[off:off + csz[k]]))
        off += int(csz[k])
    return np.frombuffer(b"".join(out), dtype).copy()


def read_vti(text_or_path) -> Grid:
    if isinstance(text_or_path, (bytes, bytearray)):
        text = text_or_path.decode()
    elif "<VTKFile" in str(text_or_path):
        text = str(text_or_path)
    else:
        with open(text_or_path) as f:
            text = f.read()
    root = ET.fromstring(text)
    header_t = np.uint64 if root.get("header_type", "UInt32") == "UInt64" else np.uint32
    compressed = root.get("compressor") is not None
    img = root.find("ImageData")
    ext = [int(v) for v in img.get("WholeExtent").split()]
    cells = np.array([ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4]], np.int64)
    spacing = np.array([float(v) for v in img.get("Spacing").split()])
    origin = np.array([float(v) for v in img.get("Origin").split()])
    mat = None
    for arr in img.iter("DataArray"):
        if arr.get("Name") == "material":
            dtype = _DTYPES[arr.get("type")]
            if arr.get("format") == "ascii":
                mat = np.array(arr.text.split(), dtype)
            else:
                mat = _decode(arr.text, dtype, header_t, compressed)
    if mat is None:
        raise ValueError("no 'material' cell data in VTI")
    if mat.size != int(np.prod(cells)):
        raise ValueError("material array does not match the cell count")
    return Grid(cells, cells * spacing, origin, mat.astype(np.int64))


def write_vti(grid: Grid, path=None, compress: bool = True) -> str:
    data = np.asarray(grid.material, np.int64).tobytes()
    if compress:
        blk = 32768
        blocks = [data[i:i + blk] for i in range(0, len(data), blk)] or [b""]
        comp = [zlib.compress(b) for b in blocks]
        head = struct.pack("<" + "I" * (3 + len(comp)), len(comp), blk, len(blocks[-1]),
                           *[len(c) for c in comp])
        payload = base64.b64encode(head).decode() + base64.b64encode(b"".join(comp)).decode()
    else:
        payload = base64.b64encode(struct.pack("<I", len(data)) + data).decode()
    c = grid.cells
    sp = np.asarray(grid.size, float) / c
    comp_attr = ' compressor="vtkZLibDataCompressor"' if compress else ""
    text = (
        '<?xml version="1.0"?>\n'
        f'<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" header_type="UInt32"{comp_attr}>\n'
        f'  <ImageData WholeExtent="0 {c[0]} 0 {c[1]} 0 {c[2]}" Origin="{" ".join(repr(float(v)) for v in grid.origin)}"'
        f' Spacing="{" ".join(repr(float(v)) for v in sp)}" Direction="1 0 0 0 1 0 0 0 1">\n'
        f'  <Piece Extent="0 {c[0]} 0 {c[1]} 0 {c[2]}">\n'
        '    <PointData>\n    </PointData>\n    <CellData>\n'
        f'      <DataArray type="Int64" Name="material" format="binary" RangeMin="{int(grid.material.min())}"'
        f' RangeMax="{int(grid.material.max())}">\n        {payload}\n      </DataArray>\n'
        '    </CellData>\n  </Piece>\n  </ImageData>\n</VTKFile>\n')
    if path is not None:
        with open(path, "w") as f:
            f.write(text)
    return text
