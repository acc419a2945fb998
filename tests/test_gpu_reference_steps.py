"""The HIP path against whole time steps composed from the reference's own
Fortran routines (tests/golden/ref_steps.npz, written by
tests/golden/make_golden.py through oracle/ref_harness.f90 kernel 9):
first-order steps 1..5 on the 5x4x3 jittered golden box with its wall hole
and blanked cells, from a restart with total = 0, both flux methods.  The
fused hp_step and the per-routine chain must equal the reference's state,
ghost history, step counter and time bit for bit after every step."""
import os

import numpy as np
import pytest

from hydraphase import mesh as M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = np.load(os.path.join(HERE, "golden", "ref_steps.npz"))
KERN = np.load(os.path.join(HERE, "golden", "ref_kernels.npz"))


def golden_mesh():
    f = KERN
    return M.Mesh(x=f["mesh/x"], y=f["mesh/y"], z=f["mesh/z"], nod=f["mesh/nod"], nc=f["mesh/nc"],
                  iblank=f["mesh/iblank"], gparent=f["mesh/gparent"], gtype=f["mesh/gtype"],
                  gside=f["mesh/gside"], shape=tuple(f["mesh/shape"]))


@pytest.mark.parametrize("mode", ["fused", "per_routine"])
@pytest.mark.parametrize("flux", [2, 1])
def test_gpu_equals_reference_composed_steps(flux, mode):
    from hydraphase import Solver
    m = golden_mesh()
    p = M.water_air_params(cfl=0.5, viscous_gas=True, myflux=flux)
    p["myorder"] = 1
    s = Solver(m, p)
    s.upload(STEPS[f"f{flux}/in/cu"], STEPS[f"f{flux}/in/gp"], 0.0, 0)
    for k in range(1, 6):
        pre = f"f{flux}/step{k}/"
        if mode == "fused":
            s.step(1)
        else:
            s.primitive()
            _, _, t, tot = s.download()
            dt = s.dt_min() * p["cfl"]
            s.set_dt(dt, t + dt, tot + 1)
            s.smooth()
            s.rk()
            d = s.end_step(True)
            np.testing.assert_array_equal(d, STEPS[pre + "diffs"], err_msg=f"step {k} monitor")
        cu, gp, t, tot = s.download()
        np.testing.assert_array_equal(cu, STEPS[pre + "cu"], err_msg=f"step {k} cu")
        np.testing.assert_array_equal(gp, STEPS[pre + "gp"], err_msg=f"step {k} ghost history")
        assert tot == k and t == float(STEPS[pre + "time"])
    s.close()
