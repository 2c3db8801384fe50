"""DADF5 writer: layout, cell_to maps and labels against the reference's result
file of the golden case (read back with h5dump; skipped without HDF5)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, fcc_golden_inputs

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP) or not os.path.exists("/opt/conda/lib/libhdf5.so"),
                                reason="HDF5 tools not available")


def _read(path, name, dtype):
    out = path + ".bin"
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, stdout=subprocess.DEVNULL)
    a = np.fromfile(out, dtype)
    os.unlink(out)
    return a


def _attr(path, name):
    out = subprocess.run([H5DUMP, "-a", name, path], capture_output=True, text=True).stdout
    return out


def test_writer_layout_and_maps(tmp_path):
    from damask_amd import _lib, dadf5
    grid, mat, _ = fcc_golden_inputs()
    maps = mat.cell_maps(grid.material)
    ref = np.load(os.path.join(GOLDEN, "fcc_6grains6x7x8.npz"))
    # the index maps are bit-exact against the reference's /cell_to
    assert np.array_equal(maps["phase_entry"], ref["cell_to_phase_entry"])
    assert np.array_equal(maps["homogenization_entry"], ref["cell_to_homogenization_entry"])
    path = str(tmp_path / "out.hdf5")
    lib = _lib.load()
    systems = [dadf5.slip_system_labels(lib, p) for p in mat.phases]
    w = dadf5.DADF5Writer(path, grid.cells, grid.size, grid.origin, maps["phase"], [p.name for p in mat.phases],
                          [p.lattice for p in mat.phases], maps["homogenization"],
                          [h[0] for h in mat.homogenizations])
    N = grid.n_cells
    rng = np.random.default_rng(0)
    P = rng.standard_normal((N, 3, 3))
    O = rng.standard_normal((N, 4))
    xi = rng.standard_normal((N, 12))
    nn = int(np.prod(np.asarray(grid.cells) + 1))
    w.add_increment(40, 20.0, {"P": P, "O": O, "xi_sl": xi}, np.ones((N, 3)), np.zeros((nn, 3)), systems)
    w.close()
    listing = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    for grp in ("cell_to", "geometry", "increment_40", "homogenization", mat.phases[0].name, "mechanical"):
        assert f'GROUP "{grp}"' in listing
    for ds in ("phase", "homogenization", "cells", "size", "origin", "u_p", "u_n", "P", "O", "xi_sl"):
        assert f'DATASET "{ds}"' in listing
    assert "SIMPLE { ( 336, 3, 3 ) / ( 336, 3, 3 ) }" in listing
    assert "SIMPLE { ( 336, 1 ) / ( 336, 1 ) }" in listing        # /cell_to/phase
    name = mat.phases[0].name
    got = _read(path, f"/increment_40/phase/{name}/mechanical/P", "<f8").reshape(N, 3, 3)
    assert np.array_equal(got, P)
    assert np.array_equal(_read(path, "/geometry/cells", "<i4"), np.asarray(grid.cells))
    # cell_to/phase entries (compound {label, entry}: the int64 follows the label bytes)
    text = subprocess.run([H5DUMP, "-d", "/cell_to/phase", path], capture_output=True, text=True).stdout
    entries = np.array([int(x) for x in re.findall(r'"\s*,\s*(-?\d+)\s*\n', text)])
    assert np.array_equal(entries, ref["cell_to_phase_entry"])
    # attributes as the reference writes them
    assert '"first Piola-Kirchhoff stress"' in _attr(path, f"/increment_40/phase/{name}/mechanical/P/description")
    assert '"Pa"' in _attr(path, f"/increment_40/phase/{name}/mechanical/P/unit")
    assert '"cF"' in _attr(path, f"/increment_40/phase/{name}/lattice")
    sysattr = _attr(path, f"/increment_40/phase/{name}/mechanical/xi_sl/systems")
    assert '"[ 0  1 -1]( 1  1  1)"' in sysattr and '"[-1 -1  0](-1  1 -1)"' in sysattr
    assert "20" in _attr(path, "/increment_40/t\\/s")


def test_slip_labels_match_reference_result():
    """the 12 FCC labels of the golden result's xi_sl 'systems' attribute"""
    from damask_amd import _lib, dadf5
    _, mat, _ = fcc_golden_inputs()
    labels = dadf5.slip_system_labels(_lib.load(), mat.phases[0])
    expect = ["[ 0  1 -1]( 1  1  1)", "[-1  0  1]( 1  1  1)", "[ 1 -1  0]( 1  1  1)", "[ 0 -1 -1](-1 -1  1)",
              "[ 1  0  1](-1 -1  1)", "[-1  1  0](-1 -1  1)", "[ 0 -1  1]( 1 -1 -1)", "[-1  0 -1]( 1 -1 -1)",
              "[ 1  1  0]( 1 -1 -1)", "[ 0  1  1](-1  1 -1)", "[ 1  0 -1](-1  1 -1)", "[-1 -1  0](-1  1 -1)"]
    assert labels == expect


@pytest.mark.gpu
def test_gpu_run_writes_reference_layout(gpu_lib, tmp_path):
    """the golden case through the Python driver with result output every
    f_out = 4 increments; increment 40 against the reference's result file"""
    from damask_amd import dadf5
    from damask_amd.grid import DamaskGrid, GridMechSpectralBasic
    grid, mat, lcs = fcc_golden_inputs()
    ref = np.load(os.path.join(GOLDEN, "fcc_6grains6x7x8.npz"))
    path = str(tmp_path / "golden.hdf5")
    solver = GridMechSpectralBasic(grid, mat)
    drv = DamaskGrid(solver)
    setup = [(os.path.join(GOLDEN, "fcc_singlePhase.yaml"), "material configuration"),
             (os.path.join(GOLDEN, "fcc_tensionY.yaml"), "load case definition (grid solver)"),
             (os.path.join(GOLDEN, "fcc_6grains6x7x8.vti"), "geometry definition (grid solver)")]
    out = dadf5.ResultOutput(path, solver, drv, setup_files=setup)
    drv.run(lcs, on_output=out)
    out.close()
    setup_listing = subprocess.run([H5DUMP, "-H", "-g", "/setup", path], capture_output=True, text=True,
                                   check=True).stdout
    assert all(f'DATASET "{os.path.basename(f)}"' in setup_listing for f, _ in setup)
    N = grid.n_cells
    name = mat.phases[0].name
    listing = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    assert all(f'GROUP "increment_{i}"' in listing for i in range(0, 41, 4))
    F = _read(path, f"/increment_40/phase/{name}/mechanical/F", "<f8").reshape(N, 3, 3)
    assert np.abs(F - ref["F_40"].reshape(N, 3, 3)).max() < 2e-5
    xi = _read(path, f"/increment_40/phase/{name}/mechanical/xi_sl", "<f8").reshape(N, 12)
    assert np.abs(xi - ref["xi_sl_40"]).max() < 2e-5 * ref["xi_sl_40"].max()
    up = _read(path, "/increment_40/geometry/u_p", "<f8").reshape(N, 3)
    assert np.abs(up - ref["u_p_40"]).max() < 1e-4 * np.abs(ref["u_n_40"]).max()


def test_setup_group_matches_reference_result(tmp_path):
    """/setup (src/result.f90:147-148,287-307): the golden case's input files
    (byte copies of the /setup datasets of the reference's result file),
    stored under the names the reference used, give datasets of the reference's
    type (NUL-terminated fixed-length string of the text's length + 1), shape
    (1,), bytes and descriptions; a repeated name gets the '.1' suffix"""
    from damask_amd import dadf5
    grid, mat, _ = fcc_golden_inputs()
    maps = mat.cell_maps(grid.material)
    path = str(tmp_path / "setup.hdf5")
    w = dadf5.DADF5Writer(path, grid.cells, grid.size, grid.origin, maps["phase"], [p.name for p in mat.phases],
                          [p.lattice for p in mat.phases], maps["homogenization"],
                          [h[0] for h in mat.homogenizations])
    files = (("fcc_singlePhase.yaml", "singlePhase.yaml", "material configuration"),
             ("fcc_tensionY.yaml", "tensionY.yaml", "load case definition (grid solver)"),
             ("fcc_6grains6x7x8.vti", "6grains6x7x8.vti", "geometry definition (grid solver)"))
    for src, name, desc in files:
        with open(os.path.join(GOLDEN, src)) as fh:
            assert w.add_setup_file(fh.read(), "some/dir/" + name, desc) == name
    assert w.add_setup_file("x", "tensionY.yaml", "load case definition (grid solver)") == "tensionY.yaml.1"
    w.close()
    ours = subprocess.run([H5DUMP, "-H", "-g", "/setup", path], capture_output=True, text=True, check=True).stdout
    assert '"input data used to run the simulation"' in _attr(path, "/setup/description")
    for _, name, desc in files:
        block = ours[ours.index(f'DATASET "{name}"'):]
        n = len(open(os.path.join(GOLDEN, dict((b, a) for a, b, _ in files)[name])).read().encode())
        assert f"STRSIZE {n + 1};" in block.split("ATTRIBUTE")[0]
        assert "DATASPACE  SIMPLE { ( 1 ) / ( 1 ) }" in block.split("ATTRIBUTE")[0]
        assert f'"{desc}"' in _attr(path, f"/setup/{name}/description")
        out = path + ".bin"
        subprocess.run([H5DUMP, "-d", f"/setup/{name}", "-b", "MEMORY", "-o", out, path], check=True,
                       stdout=subprocess.DEVNULL)
        with open(out, "rb") as fh, open(os.path.join(GOLDEN, dict((b, a) for a, b, _ in files)[name]), "rb") as src:
            assert fh.read() == src.read(), name
