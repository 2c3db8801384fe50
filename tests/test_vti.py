"""VTI reader (the .vti input of DAMASK_grid) against the known-answer vectors
and error cases of the reference's unit tests (tests/unit_cpp/test_VTI.cpp:34-290)."""
import numpy as np
import pytest

from damask_amd.geom import VTI, VTIError, read_vti

B64_UC32 = "BAAAAAECAwQAAAAA"                        # [1,2,3,4], UInt32 header, uncompressed
B64_UC64 = "BAAAAAAAAAABAgMEAAAAAAAAAAA="            # [1,2,3,4], UInt64 header
B64_COMP32 = "AgAAAAQAAAADAAAADAAAAAsAAAB4nGNgZGIGAAAOAAd4nOPi5gEAAEMAIg=="    # [0,1,2,3], [10,11,12]
B64_COMP64 = ("AgAAAAAAAAAEAAAAAAAAAAMAAAAAAAAADAAAAAAAAAALAAAAAAAAAHicY2BkYgYA"
              "AA4AB3ic4+LmAQAAQwAi")
UNCOMPRESSED = bytes([1, 2, 3, 4])
COMPRESSED = bytes([0, 1, 2, 3, 10, 11, 12])


def _xml(payload, vtype="Int32", name="mydata", header="", compressor="", byte_order="LittleEndian"):
    head = f' header_type="{header}"' if header else ""
    comp = f' compressor="{compressor}"' if compressor else ""
    return (f'<?xml version="1.0"?><VTKFile type="ImageData" version="1.0" byte_order="{byte_order}"{head}{comp}>'
            '<ImageData WholeExtent="0 1 0 1 0 1" Spacing="1 1 1" Origin="0 0 0"><Piece Extent="0 1 0 1 0 1">'
            f'<CellData><DataArray type="{vtype}" Name="{name}" format="binary">{payload}</DataArray></CellData>'
            '</Piece></ImageData></VTKFile>')


def test_read_word_little_endian():
    assert VTI.read_word(bytes([0x78, 0x56, 0x34, 0x12]), 4) == 0x12345678
    assert VTI.read_word(bytes([0xF0, 0xDE, 0xBC, 0x9A, 0x78, 0x56, 0x34, 0x12]), 8) == 0x123456789ABCDEF0


def test_decode_uncompressed():
    assert VTI.decode_uncompressed_vti(B64_UC32, 4) == UNCOMPRESSED
    assert VTI.decode_uncompressed_vti(B64_UC64, 8) == UNCOMPRESSED


def test_decode_uncompressed_underflow():
    with pytest.raises(VTIError, match="data block exceeds payload"):
        VTI.decode_uncompressed_vti("BQAAAAECAwQ=", 4)           # declares 5 bytes, carries 4


def test_decode_compressed():
    assert VTI.decode_compressed_vti(B64_COMP32, 4) == COMPRESSED
    assert VTI.decode_compressed_vti(B64_COMP64, 8) == COMPRESSED


@pytest.mark.parametrize("payload,compressor,expect", [(B64_UC32, "", UNCOMPRESSED),
                                                       (B64_COMP32, "vtkZLibDataCompressor", COMPRESSED),
                                                       (B64_UC32 + "   ", "", UNCOMPRESSED)])
def test_parse_cell_data_array(payload, compressor, expect):
    v = VTI(_xml(payload, header="UInt32", compressor=compressor))
    vtk_type, raw = v.parse_cell_data_array("mydata")
    assert vtk_type == "Int32" and raw == expect


def test_geometry_extraction():
    xml = ('<?xml version="1.0"?><VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">'
           '<ImageData WholeExtent="0 2 0 3 0 4" Spacing="0.5 1.0 2.0" Origin="10 20 30"></ImageData></VTKFile>')
    cells, size, origin, labels = VTI(xml).read_geometry()
    assert cells.tolist() == [2, 3, 4]
    assert size.tolist() == [1.0, 3.0, 8.0]
    assert origin.tolist() == [10.0, 20.0, 30.0]
    assert labels == []


def test_missing_array():
    with pytest.raises(VTIError, match="no DataArray with Name='testdata' found"):
        VTI(_xml(B64_UC32, name="other")).parse_cell_data_array("testdata")


def test_rejects_big_endian():
    with pytest.raises(VTIError, match="byte_order must be 'LittleEndian'"):
        VTI(_xml(B64_UC32, byte_order="BigEndian")).parse_cell_data_array("mydata")


def test_rejects_other_compressor():
    with pytest.raises(VTIError, match="compressor is not vtkZLibDataCompressor"):
        VTI(_xml(B64_UC32, compressor="vtkLZ4DataCompressor")).parse_cell_data_array("mydata")


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_int64_dataset(dtype):
    v = VTI(_xml("CAAAACkAAAAAAAAA", vtype="Int64"))          # 41 as Int64
    out = v.read_dataset_int("mydata", dtype)
    assert out.dtype == dtype and out.tolist() == [41]


def test_golden_geometry_files():
    """the reference's own result inputs: 6x7x8 cells of 0.75 x 0.875 x 1, 6 grains"""
    import os
    from conftest import GOLDEN
    g = read_vti(os.path.join(GOLDEN, "fcc_6grains6x7x8.vti"))
    assert g.cells.tolist() == [6, 7, 8] and np.allclose(g.size, [0.75, 0.875, 1.0])
    assert sorted(set(g.material.tolist())) == list(range(6))
