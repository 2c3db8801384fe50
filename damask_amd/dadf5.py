"""DADF5 result files with the reference's layout (src/result.f90:85-218,
479-842; materialpoint_result, src/materialpoint.f90:144-157; the phase and
geometry writers it calls), written through the HDF5 C library with ctypes.

Layout, as DAMASK_grid writes it:

    /                      DADF5_version_major/minor, creator, created, solver, call
    /geometry              attributes cells, size, origin; datasets of the same name
    /cell_to/phase         (N, 1) compound {label: fixed string, entry: int64}
    /cell_to/homogenization (N,) compound {label, entry}
    /setup                 attribute description; one string dataset per input file
                           (material.yaml, numerics.yaml, load case, geometry)
    /increment_<i>         attribute t/s
        geometry/u_p, u_n  displacements of the cell centres / nodes
        homogenization/<name>/mechanical
        phase/<name>       attribute lattice
            mechanical/F, P, F_e, F_p, L_p (N_ph, 3, 3), O (N_ph, 4), xi_sl, gamma_sl (N_ph, N_sl)

Tensors are stored T_ij at [e, i, j] (the reference writes its column-major
arrays transposed, src/result.f90:490-498).  Every dataset carries the
created / creator / description / unit attributes.

The HDF5 library is located through DAMASK_HDF5_LIB, then /opt/conda/lib and
the loader's search path; it is needed only for result output.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import datetime
import getpass
import os
import socket

import numpy as np

DADF5_VERSION = (1, 2)
CREATOR = "damask_amd (MI355X spectral basic)"

# (description, unit) of the mechanical outputs
# (src/phase_mechanical.f90:339-373, phenopowerlaw.f90:404-437, src/grid/spectral_utilities.f90 u_p/u_n)
FIELD_INFO = {
    "F": ("deformation gradient", "1"),
    "P": ("first Piola-Kirchhoff stress", "Pa"),
    "F_e": ("elastic deformation gradient", "1"),
    "F_p": ("plastic deformation gradient", "1"),
    "L_p": ("plastic velocity gradient", "1/s"),
    "O": ("crystal orientation as quaternion q_0 (q_1 q_2 q_3)", "1"),
    "xi_sl": ("resistance against plastic slip", "Pa"),
    "gamma_sl": ("plastic shear", "1"),
}
U_INFO = {"u_p": ("displacements of the materialpoints (cell centers)", "m"),
          "u_n": ("displacements of the nodes", "m")}

_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_ulonglong
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_COMPOUND = 6
H5T_STR_SPACEPAD = 2
H5T_VARIABLE = C.c_size_t(-1).value


def _find_lib() -> str:
    cands = [os.environ.get("DAMASK_HDF5_LIB", ""), "/opt/conda/lib/libhdf5.so",
             ctypes.util.find_library("hdf5") or ""]
    for c in cands:
        if c and (os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise OSError("no HDF5 C library found (set DAMASK_HDF5_LIB)")


class _H5:
    """The few HDF5 1.10 C functions the writer needs."""

    def __init__(self):
        L = C.CDLL(_find_lib())
        L.H5open.restype = _herr
        if L.H5open() < 0:
            raise OSError("H5open failed")
        sig = {
            "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
            "H5Fclose": (_herr, [_hid]),
            "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
            "H5Gclose": (_herr, [_hid]),
            "H5Screate": (_hid, [C.c_int]),
            "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sclose": (_herr, [_hid]),
            "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Dclose": (_herr, [_hid]),
            "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
            "H5Awrite": (_herr, [_hid, _hid, C.c_void_p]),
            "H5Aclose": (_herr, [_hid]),
            "H5Tcopy": (_hid, [_hid]),
            "H5Tset_size": (_herr, [_hid, C.c_size_t]),
            "H5Tset_strpad": (_herr, [_hid, C.c_int]),
            "H5Tcreate": (_hid, [C.c_int, C.c_size_t]),
            "H5Tinsert": (_herr, [_hid, C.c_char_p, C.c_size_t, _hid]),
            "H5Tclose": (_herr, [_hid]),
            "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
            "H5Gopen2": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Aread": (_herr, [_hid, _hid, C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L = L
        g = lambda n: _hid.in_dll(L, n).value   # noqa: E731  (global type ids, valid after H5open)
        self.F64 = g("H5T_IEEE_F64LE_g")
        self.I32 = g("H5T_STD_I32LE_g")
        self.I64 = g("H5T_STD_I64LE_g")
        self.NATIVE_DOUBLE = g("H5T_NATIVE_DOUBLE_g")
        self.NATIVE_INT32 = g("H5T_NATIVE_INT_g")
        self.NATIVE_INT64 = g("H5T_NATIVE_LLONG_g")
        self.C_S1 = g("H5T_C_S1_g")

    def ok(self, r, what):
        if r < 0:
            raise OSError(f"HDF5: {what} failed")
        return r


def _now() -> str:
    return datetime.datetime.now().astimezone().strftime("%Y-%m-%d %H:%M:%S%z")


class DADF5Writer:
    """One DADF5 file for a grid run.  ``phase_of``/``ho_of`` are the 0-based
    phase / homogenization index per cell, ``phase_names``/``ho_names`` their
    labels in material.yaml order, ``lattices`` the phase lattices."""

    def __init__(self, path: str, cells, size, origin, phase_of, phase_names, lattices, ho_of=None,
                 ho_names=("SX",), call: str = ""):
        self.h5 = _H5()
        self.cells = np.asarray(cells, np.int32)
        self.N = int(np.prod(self.cells))
        self.phase_of = np.asarray(phase_of, np.int64)
        self.ho_of = np.zeros(self.N, np.int64) if ho_of is None else np.asarray(ho_of, np.int64)
        self.phase_names = list(phase_names)
        self.ho_names = list(ho_names)
        self.lattices = list(lattices)
        H = self.h5
        self.f = H.ok(H.L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create " + path)
        self._attr_int(self.f, "DADF5_version_major", [DADF5_VERSION[0]], scalar=True)
        self._attr_int(self.f, "DADF5_version_minor", [DADF5_VERSION[1]], scalar=True)
        self._attr_str(self.f, "creator", CREATOR)
        self._attr_str(self.f, "created", _now())
        self._attr_str(self.f, "solver", "grid")
        self._attr_str(self.f, "call", call)
        try:
            self._attr_str(self.f, "user", getpass.getuser())
        except Exception:
            pass
        self._attr_str(self.f, "host", socket.gethostname())
        # geometry (src/grid/discretization_grid.f90:106-254 -> result_addSetupFile / geometry)
        g = self._group(self.f, "geometry")
        self._attr_int(g, "cells", self.cells)
        self._attr_f64(g, "size", np.asarray(size, float))
        self._attr_f64(g, "origin", np.asarray(origin, float))
        for name, val, unit, t in (("cells", self.cells, "1", "i"), ("size", np.asarray(size, float), "m", "f"),
                                   ("origin", np.asarray(origin, float), "m", "f")):
            d = self._dataset(g, name, val, int32=(t == "i"))
            self._attr_str(d, "unit", unit)
            H.L.H5Dclose(d)
        H.L.H5Gclose(g)
        # cell_to maps (src/result.f90:534-692): entry = rank of the cell within its phase
        g = self._group(self.f, "cell_to")
        self._attr_str(g, "description", "mappings to place data in space")
        self.entry_phase = np.zeros(self.N, np.int64)
        for p in range(len(self.phase_names)):
            m = self.phase_of == p
            self.entry_phase[m] = np.arange(int(m.sum()))
        self.entry_ho = np.zeros(self.N, np.int64)
        for h in range(len(self.ho_names)):
            m = self.ho_of == h
            self.entry_ho[m] = np.arange(int(m.sum()))
        self._mapping(g, "phase", [self.phase_names[p] for p in self.phase_of], self.entry_phase, (self.N, 1),
                      "cell ID and constituent ID to phase results")
        self._mapping(g, "homogenization", [self.ho_names[h] for h in self.ho_of], self.entry_ho, (self.N,),
                      "cell ID to homogenization results")
        H.L.H5Gclose(g)
        # input files (src/result.f90:147-148,287-307)
        self.setup = self._group(self.f, "setup")
        self._attr_str(self.setup, "description", "input data used to run the simulation")
        self.setup_names = set()

    def add_setup_file(self, content: str, fname: str, description: str) -> str:
        """result_addSetupFile (src/result.f90:287-307): the file's text as a
        fixed-length, NUL-terminated string dataset of shape (1,) named after the
        file (directories stripped, '.1', '.2', ... appended to repeated names)."""
        H = self.h5
        name = os.path.basename(fname)
        suffix, i = "", 0
        while name + suffix in self.setup_names:
            i += 1
            suffix = f".{i}"
        name += suffix
        self.setup_names.add(name)
        data = content.rstrip(" ").encode()
        t = H.L.H5Tcopy(H.C_S1)
        H.L.H5Tset_size(t, len(data) + 1)
        s = self._space((1,))
        buf = C.create_string_buffer(data, len(data) + 1)
        d = H.ok(H.L.H5Dcreate2(self.setup, name.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                 "setup " + name)
        H.ok(H.L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(buf, C.c_void_p)), "write setup " + name)
        self._stamp(d, description)
        H.L.H5Dclose(d)
        H.L.H5Sclose(s)
        H.L.H5Tclose(t)
        return name

    # ------------------------------------------------------------ primitives
    def _group(self, loc, name):
        H = self.h5
        return H.ok(H.L.H5Gcreate2(loc, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group " + name)

    def _space(self, shape):
        H = self.h5
        if shape is None:
            return H.ok(H.L.H5Screate(H5S_SCALAR), "scalar space")
        dims = (_hsize * len(shape))(*shape)
        return H.ok(H.L.H5Screate_simple(len(shape), dims, None), "space")

    def _attr_raw(self, loc, name, ftype, mtype, shape, buf):
        H = self.h5
        s = self._space(shape)
        a = H.ok(H.L.H5Acreate2(loc, name.encode(), ftype, s, H5P_DEFAULT, H5P_DEFAULT), "attribute " + name)
        H.ok(H.L.H5Awrite(a, mtype, buf), "write attribute " + name)
        H.L.H5Aclose(a)
        H.L.H5Sclose(s)

    def _attr_int(self, loc, name, values, scalar=False):
        v = np.ascontiguousarray(values, np.int32)
        self._attr_raw(loc, name, self.h5.I32, self.h5.NATIVE_INT32, None if scalar else v.shape,
                       v.ctypes.data_as(C.c_void_p))

    def _attr_f64(self, loc, name, values, scalar=False):
        v = np.ascontiguousarray(values, np.float64)
        self._attr_raw(loc, name, self.h5.F64, self.h5.NATIVE_DOUBLE, None if scalar else v.shape,
                       v.ctypes.data_as(C.c_void_p))

    def _vlstr(self):
        H = self.h5
        t = H.L.H5Tcopy(H.C_S1)
        H.L.H5Tset_size(t, H5T_VARIABLE)
        return t

    def _attr_str(self, loc, name, text):
        t = self._vlstr()
        p = (C.c_char_p * 1)(str(text).encode())
        self._attr_raw(loc, name, t, t, None, C.cast(p, C.c_void_p))
        self.h5.L.H5Tclose(t)

    def _attr_strs(self, loc, name, texts):
        t = self._vlstr()
        p = (C.c_char_p * len(texts))(*[s.encode() for s in texts])
        self._attr_raw(loc, name, t, t, (len(texts),), C.cast(p, C.c_void_p))
        self.h5.L.H5Tclose(t)

    def _dataset(self, loc, name, data, int32=False):
        H = self.h5
        a = np.ascontiguousarray(data, np.int32 if int32 else np.float64)
        s = self._space(a.shape)
        ftype, mtype = (H.I32, H.NATIVE_INT32) if int32 else (H.F64, H.NATIVE_DOUBLE)
        d = H.ok(H.L.H5Dcreate2(loc, name.encode(), ftype, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "dataset " + name)
        H.ok(H.L.H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), "write " + name)
        H.L.H5Sclose(s)
        return d

    def _stamp(self, d, description, unit=None):
        self._attr_str(d, "created", _now())
        self._attr_str(d, "creator", CREATOR)
        self._attr_str(d, "description", description)
        if unit is not None:
            self._attr_str(d, "unit", unit)

    def _mapping(self, loc, name, labels, entries, shape, description):
        H = self.h5
        width = max(1, max(len(s) for s in labels))
        st = H.L.H5Tcopy(H.C_S1)
        H.L.H5Tset_size(st, width)
        H.L.H5Tset_strpad(st, H5T_STR_SPACEPAD)
        rec = np.dtype([("label", f"S{width}"), ("entry", "<i8")])
        ct = H.L.H5Tcreate(H5T_COMPOUND, rec.itemsize)
        H.L.H5Tinsert(ct, b"label", rec.fields["label"][1], st)
        H.L.H5Tinsert(ct, b"entry", rec.fields["entry"][1], H.I64)
        buf = np.zeros(len(labels), rec)
        buf["label"] = [s.ljust(width).encode() for s in labels]
        buf["entry"] = entries
        s = self._space(shape)
        d = H.ok(H.L.H5Dcreate2(loc, name.encode(), ct, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "mapping")
        H.ok(H.L.H5Dwrite(d, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data_as(C.c_void_p)), "write mapping")
        self._stamp(d, description)
        H.L.H5Dclose(d)
        H.L.H5Sclose(s)
        H.L.H5Tclose(ct)
        H.L.H5Tclose(st)

    # ------------------------------------------------------------ increments
    def add_increment(self, inc: int, t: float, phase_fields: dict, u_p=None, u_n=None, systems=None):
        """``phase_fields``: output name -> per-cell array over all N cells
        ((N,3,3) tensors T_ij at [e,i,j], (N,4) O, (N,N_sl) slip fields); the
        rows of each phase go to its group.  ``systems``: per phase the slip
        system labels of its xi_sl / gamma_sl datasets."""
        H = self.h5
        gi = self._group(self.f, f"increment_{inc}")
        self._attr_f64(gi, "t/s", np.array([t]), scalar=True)
        gg = self._group(gi, "geometry")
        for name, val in (("u_p", u_p), ("u_n", u_n)):
            if val is None:
                continue
            d = self._dataset(gg, name, np.asarray(val, float).reshape(-1, 3))
            self._stamp(d, *U_INFO[name])
            H.L.H5Dclose(d)
        H.L.H5Gclose(gg)
        gh = self._group(gi, "homogenization")
        for name in self.ho_names:
            g1 = self._group(gh, name)
            H.L.H5Gclose(self._group(g1, "mechanical"))
            H.L.H5Gclose(g1)
        H.L.H5Gclose(gh)
        gp = self._group(gi, "phase")
        for p, name in enumerate(self.phase_names):
            g1 = self._group(gp, name)
            self._attr_str(g1, "lattice", self.lattices[p])
            gm = self._group(g1, "mechanical")
            rows = np.nonzero(self.phase_of == p)[0]
            for out, val in phase_fields.items():
                val = np.asarray(val)
                desc, unit = FIELD_INFO[out]
                data = val[rows]
                if out in ("xi_sl", "gamma_sl"):
                    nsl = len(systems[p]) if systems else data.shape[1]
                    data = data[:, :nsl]
                d = self._dataset(gm, out, data)
                self._stamp(d, desc, unit)
                if out == "O":
                    self._attr_str(d, "lattice", self.lattices[p])
                if out in ("xi_sl", "gamma_sl") and systems:
                    self._attr_strs(d, "systems", systems[p])
                H.L.H5Dclose(d)
            H.L.H5Gclose(gm)
            H.L.H5Gclose(g1)
        H.L.H5Gclose(gp)
        H.L.H5Gclose(gi)

    def close(self):
        if self.f is not None:
            if self.setup is not None:
                self.h5.L.H5Gclose(self.setup)
                self.setup = None
            self.h5.L.H5Fclose(self.f)
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def slip_system_labels(lib, phase) -> list[str]:
    """'[d1 d2 d3](n1 n2 n3)' labels of the active slip systems, each index in
    two characters (crystal.f90:2133-2181)."""
    import ctypes as C
    lat = 0 if phase.lattice == "cF" else 1
    N_sl = (C.c_int32 * len(phase.N_sl))(*phase.N_sl)
    out = (C.c_double * 6)()
    labels = []
    for a in range(phase.nsl):
        if lib.dmk_slip_system(lat, len(phase.N_sl), N_sl, a, out):
            raise ValueError(f"slip system {a} of phase {phase.name}")
        v = [int(round(x)) for x in out]
        labels.append("[" + " ".join(f"{i:2d}" for i in v[:3]) + "](" + " ".join(f"{i:2d}" for i in v[3:]) + ")")
    return labels


class ResultOutput:
    """Result writer for a GridMechSpectralBasic run: increment 0 at creation,
    then one increment per call (use as DamaskGrid.run's ``on_output``).
    Outputs are those listed in material.yaml (phase mechanical and plastic
    ``output`` keys)."""

    def __init__(self, path: str, solver, driver=None, call: str = "", setup_files=()):
        self.s = solver
        self.driver = driver
        g = solver.grid
        mat = solver.material
        self.w = DADF5Writer(path, g.cells, g.size, g.origin, solver.maps["phase"], [p.name for p in mat.phases],
                             [p.lattice for p in mat.phases], solver.maps["homogenization"],
                             [h[0] for h in mat.homogenizations], call=call)
        # the input files in the order the reference reads them: material and
        # numerics configuration (src/config.f90:39-52), load case
        # (src/grid/DAMASK_grid.f90:159), geometry (src/grid/discretization_grid.f90:162)
        for fname, description in setup_files:
            with open(fname) as fh:
                self.w.add_setup_file(fh.read(), fname, description)
        self.systems = [slip_system_labels(solver.lib, p) for p in mat.phases]
        outs = []
        for p in mat.phases:
            for o in list(p.output) + list(p.output_plastic):
                if o in FIELD_INFO and o not in outs:
                    outs.append(o)
        self.outputs = outs
        cells = np.asarray(g.cells)
        step = np.asarray(g.size, float) / cells
        k, j, i = np.meshgrid(*[np.arange(n) for n in cells[::-1]], indexing="ij")
        self.X0p = (np.stack([i, j, k], -1).reshape(-1, 3) + 0.5) * step + np.asarray(g.origin, float)
        k, j, i = np.meshgrid(*[np.arange(n + 1) for n in cells[::-1]], indexing="ij")
        self.X0n = np.stack([i, j, k], -1).reshape(-1, 3) * step + np.asarray(g.origin, float)
        self(0, 0.0, zero_displacement=True)

    def __call__(self, inc: int, t: float, zero_displacement: bool = False):
        fields = {o: self.s.field(o) for o in self.outputs}
        if zero_displacement:
            u_p, u_n = np.zeros_like(self.X0p), np.zeros_like(self.X0n)
        else:
            x_p, x_n = self.driver.coords if self.driver is not None and hasattr(self.driver, "coords") \
                else self.s.updateCoords()
            u_p, u_n = x_p - self.X0p, x_n - self.X0n
        self.w.add_increment(inc, t, fields, u_p, u_n, self.systems)

    def close(self):
        self.w.close()
