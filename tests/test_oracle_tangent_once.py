"""The oracle's tangent-once mode (dPdF, orientation and stiffness averages
formed after each solve from the last response, as the GPU path does, instead
of after every response; the CPU baseline runs in this mode) gives the run of
the default mode bit for bit: spectral_basic reads C_volAvg and C_minMaxAvg
only between solves (grid_mech_spectral_basic.f90:261-262,306-311).  One
thread: the oracle's OpenMP dot-product reductions round differently run to
run (1e-12) when several threads combine them."""
import os

import numpy as np


def test_tangent_once_reproduces_the_default_run(oracle):
    from damask_amd import config, workloads
    grid, mat = workloads.polycrystal([8, 7, 6], 6, 2)
    mat = config.parse_material(mat)
    _, lcs = config.parse_load(workloads.tension_x(6.0, 6))
    for num in (config.Numerics(), config.Numerics(update_gamma=True, N_iter_max=8)):
        runs = []
        for once in (False, True):
            pr = oracle.Problem(grid, mat, num, tangent_once=once, threads=1)
            r = pr.run_loadcase(lcs[0])
            runs.append((r, pr.solver_get(4), pr.solver_get(5), pr.get(7, 4), pr.solver_get(0)))
            pr.close()
        (a, Ca, Ma, Oa, Fa), (b, Cb, Mb, Ob, Fb) = runs
        assert a["rc"] == b["rc"]
        assert np.array_equal(a["stats"], b["stats"])
        assert np.array_equal(a["P_av"], b["P_av"])
        assert np.array_equal(Ca, Cb) and np.array_equal(Ma, Mb) and np.array_equal(Fa, Fb)
        assert np.array_equal(Oa, Ob)
    oracle.lib().orc_set_threads(min(16, os.cpu_count() or 1))
