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
    """Periodic Voronoi map; material IDs are the seed indices (0-based): the
    nearest seed of every cell centre (minimum image when periodic), the
    lowest index among equidistant seeds.  A k-d tree answers the queries
    (periodic through its box size); without scipy every distance is formed."""
    cells = np.asarray(cells, np.int64)
    size = np.asarray(size, float)
    coords = coordinates0_point(cells, size).reshape(-1, 3)
    seeds = np.asarray(seeds, float)
    try:
        from scipy.spatial import cKDTree
    except ImportError:
        cKDTree = None
    if cKDTree is not None and len(seeds) > 1:
        tree = cKDTree(np.mod(seeds, size) if periodic else seeds, boxsize=size if periodic else None)
        d, i = tree.query(coords, k=2)
        best = i[:, 0].copy()
        tie = d[:, 0] == d[:, 1]
        best[tie] = np.minimum(i[tie, 0], i[tie, 1])
    else:
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


class VTIError(ValueError):
    """IO_error 844 (VTK) of the reference's VTI reader (src/grid/VTI.cpp:43)."""

    def __init__(self, msg: str):
        super().__init__(f"IO_error 844: {msg}")
        self.code = 844
        self.msg = msg


class VTI:
    """VTK ImageData (.vti) reader with the behaviour of src/grid/VTI.cpp: binary
    DataArrays, base64 with UInt32/UInt64 headers, uncompressed or
    vtkZLibDataCompressor blocks, little endian only."""

    def __init__(self, text_or_path):
        if isinstance(text_or_path, (bytes, bytearray)):
            text = text_or_path.decode()
        elif "<VTKFile" in str(text_or_path):
            text = str(text_or_path)
        else:
            try:
                with open(text_or_path) as f:
                    text = f.read()
            except OSError:
                raise VTIError(f"cannot open file '{text_or_path}'") from None
        try:
            self.root = ET.fromstring(text)
        except ET.ParseError as e:
            raise VTIError(f"invalid XML ({e})") from None
        if self.root.tag != "VTKFile":
            raise VTIError("missing <VTKFile> element")
        self.word = 8 if self.root.get("header_type", "UInt32") == "UInt64" else 4

    @staticmethod
    def read_word(b: bytes, n_bytes_per_word: int) -> int:
        """little-endian unsigned word of 4 or 8 bytes (VTI.cpp:55-59)"""
        return int.from_bytes(bytes(b[:n_bytes_per_word]), "little")

    @staticmethod
    def decode_b64(text: str) -> bytes:
        """whitespace-free base64, possibly several padded runs back to back (VTK writes
        the header of compressed data as a run of its own, VTI.cpp:315-350)"""
        clean = re.sub(r"\s+", "", text).lstrip("=")
        out = bytearray()
        for run in re.findall(r"[^=]+=*", clean):
            try:
                out += base64.b64decode(run, validate=True)
            except (ValueError, TypeError):
                raise VTIError("base64 decode failed (invalid character or malformed input)") from None
        return bytes(out)

    @classmethod
    def decode_uncompressed_vti(cls, text: str, n_bytes_per_word: int) -> bytes:
        """blocks of (byte count, bytes) until a zero count or the end (VTI.cpp:111-131)"""
        d = cls.decode_b64(text)
        out, off = bytearray(), 0
        while off + n_bytes_per_word <= len(d):
            n = cls.read_word(d[off:], n_bytes_per_word)
            off += n_bytes_per_word
            if n == 0:
                break
            if off + n > len(d):
                raise VTIError("VTI uncompressed: data block exceeds payload")
            out += d[off:off + n]
            off += n
        return bytes(out)

    @classmethod
    def decode_compressed_vti(cls, text: str, n_bytes_per_word: int) -> bytes:
        """header (blocks, block size, last block size, compressed sizes) followed by
        the zlib blocks, all in one base64 run (VTI.cpp:61-109)"""
        d = cls.decode_b64(text)
        w = n_bytes_per_word
        if len(d) < 3 * w:
            raise VTIError("header for compressed VTI too short")
        nb, bs, last = (cls.read_word(d[k * w:], w) for k in range(3))
        hsz = 3 * w + nb * w
        if len(d) < hsz:
            raise VTIError("missing header for compressed VTI")
        sizes = [cls.read_word(d[3 * w + k * w:], w) for k in range(nb)]
        payload = d[hsz:]
        out, off = bytearray(), 0
        for k in range(nb):
            if off + sizes[k] > len(payload):
                raise VTIError(f"invalid VTI, defined size overflowing for block {k}")
            want = last if (k + 1 == nb and last > 0) else bs
            try:
                blk = zlib.decompress(payload[off:off + sizes[k]])
            except zlib.error:
                raise VTIError(f"zlib inflate failed on block {k}") from None
            if len(blk) != want:
                raise VTIError(f"zlib inflate failed on block {k}")
            out += blk
            off += sizes[k]
        return bytes(out)

    def _check_format(self):
        """VTI.cpp:149-161"""
        r = self.root
        if r.get("type", "") != "ImageData":
            raise VTIError(f"type is not ImageData (got '{r.get('type', '')}')")
        if r.get("byte_order", "") != "LittleEndian":
            raise VTIError(f"byte_order must be 'LittleEndian' (got '{r.get('byte_order', '')}')")
        comp = r.get("compressor", "")
        if comp and comp != "vtkZLibDataCompressor":
            raise VTIError(f"compressor is not vtkZLibDataCompressor (got '{comp}')")

    def parse_cell_data_array(self, name: str) -> tuple[str, bytes]:
        """(VTK type, raw bytes) of the cell DataArray `name` (VTI.cpp:265-300)"""
        self._check_format()
        img = self.root.find("ImageData")
        if img is None:
            raise VTIError("missing <ImageData> element")
        cd = img.find("Piece/CellData")
        if cd is None:
            raise VTIError("missing <CellData> element")
        arr = next((a for a in cd.findall("DataArray") if a.get("Name") == name), None)
        if arr is None:
            raise VTIError(f"no DataArray with Name='{name}' found")
        if arr.get("format") != "binary":
            raise VTIError(f"DataArray '{name}' is not binary")
        vtk_type = arr.get("type")
        if not vtk_type:
            raise VTIError("DataArray missing 'type' attribute")
        text = arr.text or ""
        raw = (self.decode_compressed_vti(text, self.word) if self.root.get("compressor")
               else self.decode_uncompressed_vti(text, self.word))
        return vtk_type, raw

    def read_geometry(self):
        """cells, size = Spacing * cells, origin and the cell-data labels (VTI.cpp:163-250)"""
        self._check_format()
        img = self.root.find("ImageData")
        if img is None:
            raise VTIError("missing ImageData element")
        d = img.get("Direction", "")
        if d and d.split() != "1 0 0 0 1 0 0 0 1".split():
            raise VTIError(f"unsupported 'Direction' (got '{d}')")
        ext_s = img.get("WholeExtent", "")
        if not ext_s:
            raise VTIError("missing 'WholeExtent'")
        try:
            ext = [int(v) for v in ext_s.split()]
        except ValueError:
            ext = []
        if len(ext) != 6 or ext[0] != 0 or ext[2] != 0 or ext[4] != 0:
            raise VTIError(f"invalid 'WholeExtent' (got '{ext_s}')")

        def three(field, s):
            try:
                v = [float(x) for x in s.split()[:3]]
            except ValueError:
                v = []
            if len(v) != 3:
                raise VTIError(f"bad numeric field for '{field}' (got '{s}')")
            return np.array(v)
        spacing = three("ImageData@Spacing", img.get("Spacing", ""))
        origin = three("ImageData@Origin", img.get("Origin", ""))
        cells = np.array([ext[1], ext[3], ext[5]], np.int64)
        size = spacing * cells
        if np.any(size <= 0):
            raise VTIError("one or more entries <= 0 for 'size'")
        if np.any(cells < 1):
            raise VTIError("one or more entries < 1 for 'cells'")
        labels = []
        cd = img.find("Piece/CellData")
        if cd is not None:
            for a in cd.findall("DataArray"):
                nm = a.get("Name", "")
                if not nm:
                    continue
                if nm in labels:
                    raise VTIError(f"repeated label '{nm}'")
                labels.append(nm)
        return cells, size, origin, labels

    def read_dataset_int(self, name: str, dtype=np.int64) -> np.ndarray:
        vtk_type, raw = self.parse_cell_data_array(name)
        if vtk_type not in _DTYPES or _DTYPES[vtk_type] in (np.float32, np.float64):
            raise VTIError(f"unsupported integer type for dataset '{name}'")
        src = np.dtype(_DTYPES[vtk_type])
        if len(raw) % src.itemsize:
            raise VTIError("size mismatch")
        return np.frombuffer(raw, src).astype(dtype)

    def read_dataset_real(self, name: str) -> np.ndarray:
        vtk_type, raw = self.parse_cell_data_array(name)
        if vtk_type not in _DTYPES:
            raise VTIError(f"unknown VTK type '{vtk_type}'")
        src = np.dtype(_DTYPES[vtk_type])
        if len(raw) % src.itemsize:
            raise VTIError("size mismatch")
        return np.frombuffer(raw, src).astype(np.float64)


def read_vti(text_or_path) -> Grid:
    """Geometry and the 'material' cell data (DAMASK_grid's .vti input)."""
    v = VTI(text_or_path)
    cells, size, origin, labels = v.read_geometry()
    if "material" not in labels:
        raise VTIError("no DataArray with Name='material' found")
    mat = v.read_dataset_int("material")
    if mat.size != int(np.prod(cells)):
        raise VTIError("size mismatch")
    return Grid(cells, size, origin, mat)


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
