"""<job>_restart.hdf5 in the reference's layout, written and read through the
HDF5 C library (the same ctypes binding as the result writer).

Layout (src/grid/DAMASK_grid.f90:429-451 restart trigger,
src/grid/grid_mech_spectral_basic.f90:378-416 restartWrite,
src/phase_mechanical.f90:1248-1261 mechanical_restartWrite,
src/homogenization.f90:492-517 homogenization_restartWrite):

    /                              attribute increment
    /solver                        F, F_lastinc (N, 3, 3); P_aim, F_aim, F_aim_lastinc,
                                   F_aimDot (3, 3); C_volAvg, C_volAvgLastInc,
                                   C_minMaxAvg (3, 3, 3, 3)
    /phase/<name>                  omega_plastic (N_ph, 2 N_sl: xi_sl then gamma_sl),
                                   F, F_p, F_i, L_p, L_i (N_ph, 3, 3)
    /homogenization/<name>         omega_mechanical (N_ho, 0): isostrain has no state

The reference writes its Fortran arrays as they lie in memory, so an HDF5
reader sees a 3x3 tensor at [e, j, i] = T_ij and a 3x3x3x3 at [l, k, j, i].
The read side follows the reference's restart read: the phase data become the
increment-start state (phase_mechanical.f90:1267-1280), S is not stored, and
the solver then evaluates the response at the restored F with dt = 0 before
reading the stiffness averages (grid_mech_spectral_basic.f90:198-232).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .dadf5 import H5F_ACC_RDONLY, H5F_ACC_TRUNC, H5P_DEFAULT, H5S_ALL, _H5, _hsize

SOLVER_33 = ("P_aim", "F_aim", "F_aim_lastinc", "F_aimDot")
SOLVER_3333 = ("C_volAvg", "C_volAvgLastInc", "C_minMaxAvg")


def _T(a: np.ndarray) -> np.ndarray:
    """(N, 3, 3) [e, i, j] <-> the reference's stored [e, j, i] (an involution)"""
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


class _File:
    def __init__(self, path: str, write: bool):
        self.h5 = _H5()
        L = self.h5.L
        if write:
            self.f = self.h5.ok(L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create " + path)
        else:
            self.f = self.h5.ok(L.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "open " + path)
        self.open = []

    def group(self, loc, name, create):
        L = self.h5.L
        g = L.H5Gcreate2(loc, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) if create \
            else L.H5Gopen2(loc, name.encode(), H5P_DEFAULT)
        self.open.append(self.h5.ok(g, "group " + name))
        return g

    def write(self, loc, name, a):
        H, L = self.h5, self.h5.L
        a = np.ascontiguousarray(a, np.float64)
        dims = (_hsize * a.ndim)(*a.shape)
        s = H.ok(L.H5Screate_simple(a.ndim, dims, None), "space")
        d = H.ok(L.H5Dcreate2(loc, name.encode(), H.F64, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "dataset " + name)
        H.ok(L.H5Dwrite(d, H.NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), name)
        L.H5Dclose(d)
        L.H5Sclose(s)

    def read(self, loc, name):
        H, L = self.h5, self.h5.L
        d = H.ok(L.H5Dopen2(loc, name.encode(), H5P_DEFAULT), "dataset " + name)
        s = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(s)
        dims = (_hsize * max(nd, 1))()
        L.H5Sget_simple_extent_dims(s, dims, None)
        a = np.zeros(tuple(dims[i] for i in range(nd)), np.float64)
        if a.size:
            H.ok(L.H5Dread(d, H.NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), name)
        L.H5Sclose(s)
        L.H5Dclose(d)
        return a

    def attr_int(self, name, value=None):
        H, L = self.h5, self.h5.L
        v = np.zeros(1, np.int32)
        if value is None:
            a = H.ok(L.H5Aopen(self.f, name.encode(), H5P_DEFAULT), "attribute " + name)
            H.ok(L.H5Aread(a, H.NATIVE_INT32, v.ctypes.data_as(C.c_void_p)), name)
        else:
            v[0] = value
            s = H.ok(L.H5Screate(0), "scalar")
            a = H.ok(L.H5Acreate2(self.f, name.encode(), H.I32, s, H5P_DEFAULT, H5P_DEFAULT), "attribute " + name)
            H.ok(L.H5Awrite(a, H.NATIVE_INT32, v.ctypes.data_as(C.c_void_p)), name)
            L.H5Sclose(s)
        L.H5Aclose(a)
        return int(v[0])

    def close(self):
        for g in reversed(self.open):
            self.h5.L.H5Gclose(g)
        self.h5.L.H5Fclose(self.f)


def write_restart(path: str, solver, increment: int) -> None:
    """restartWrite of the solver and the material points after a converged increment"""
    d = solver.restart_write()
    f = _File(path, True)
    try:
        f.attr_int("increment", increment)
        g = f.group(f.f, "solver", True)
        f.write(g, "F", _T(d["F"]))
        f.write(g, "F_lastinc", _T(d["F_lastinc"]))
        for k in SOLVER_33:
            f.write(g, k, d[k].T)
        for k in SOLVER_3333:
            f.write(g, k, np.asarray(d[k]).transpose(3, 2, 1, 0))
        phase_of = solver.maps["phase"]
        gp = f.group(f.f, "phase", True)
        for p, ph in enumerate(solver.material.phases):
            m = phase_of == p
            nsl = ph.nsl
            g = f.group(gp, ph.name, True)
            f.write(g, "omega_plastic", np.concatenate([d["xi_sl"][m][:, :nsl], d["gamma_sl"][m][:, :nsl]], axis=1))
            f.write(g, "F", _T(d["F"][m]))
            f.write(g, "F_p", _T(d["F_p"][m]))
            f.write(g, "F_i", _T(np.broadcast_to(np.eye(3), (int(m.sum()), 3, 3))))
            f.write(g, "L_p", _T(d["L_p"][m]))
            f.write(g, "L_i", np.zeros((int(m.sum()), 3, 3)))
        gh = f.group(f.f, "homogenization", True)
        ho_of = solver.maps["homogenization"]
        for h, (name, _, _) in enumerate(solver.material.homogenizations):
            g = f.group(gh, name, True)
            f.write(g, "omega_mechanical", np.zeros((int((ho_of == h).sum()), 0)))
    finally:
        f.close()


def read_restart(path: str, solver) -> int:
    """restart read of a fresh solver (grid_mech_spectral_basic.f90:198-232,
    phase_mechanical.f90:1267-1280); returns the stored increment, after which
    the driver's next load case skips increments 1..increment"""
    f = _File(path, False)
    try:
        inc = f.attr_int("increment")
        g = f.group(f.f, "solver", False)
        N = solver.N
        d = {"F": _T(f.read(g, "F")), "F_lastinc": _T(f.read(g, "F_lastinc"))}
        for k in SOLVER_33:
            d[k] = f.read(g, k).T
        for k in SOLVER_3333:
            d[k] = f.read(g, k).transpose(3, 2, 1, 0)
        nsl = solver.nsl
        d["F_p"] = np.zeros((N, 3, 3))
        d["L_p"] = np.zeros((N, 3, 3))
        d["xi_sl"] = np.zeros((N, nsl))
        d["gamma_sl"] = np.zeros((N, nsl))
        d["S"] = np.zeros((N, 3, 3))            # not stored: S_0 = 0 as after the reference's restart read
        phase_of = solver.maps["phase"]
        gp = f.group(f.f, "phase", False)
        for p, ph in enumerate(solver.material.phases):
            m = phase_of == p
            gph = f.group(gp, ph.name, False)
            om = f.read(gph, "omega_plastic")
            d["xi_sl"][np.ix_(m, np.arange(ph.nsl))] = om[:, :ph.nsl]
            d["gamma_sl"][np.ix_(m, np.arange(ph.nsl))] = om[:, ph.nsl:2 * ph.nsl]
            d["F_p"][m] = _T(f.read(gph, "F_p"))
            d["L_p"][m] = _T(f.read(gph, "L_p"))
    finally:
        f.close()
    solver.restart_read(d, restart_inc=inc)
    return inc
