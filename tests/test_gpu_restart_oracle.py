"""Restart files as the channel between this library and the reference's
algorithm: the GPU writes its state after two RK3 steps in the reference's
instantaneous-field format (domain1_<field>_<iter>.bin, FP64, Fortran order,
no header; src/io_restart.f90:55-130), the oracle reads those files with the
format's own reader and both continue for two more steps from them: the
trajectories agree to the parity bar.  And the other way round: files the
oracle's state was written to start the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LX, LZ = 4 * np.pi, 4 * np.pi / 3
CFG = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_restart_files_carry_the_state_both_ways(po, tmp_path):
    from chapsim2_amd import restart
    from chapsim2_amd.domain import PROJ_EXACT, Domain
    from chapsim2_amd.initial import channel_fields

    nc = (32, 24, 16)
    fields = channel_fields(nc, LX, -1.0, 1.0, LZ, 2, 0.075, noise=0.1, seed=31)
    dm = Domain(nc, LX, -1.0, 1.0, LZ, projection=PROJ_EXACT, **CFG)
    of = po.Flow(list(nc), LX, -1.0, 1.0, LZ, wall_y=True, iacc=4, poisson=2, projection=1, **CFG)
    try:
        dm.set_state(*fields)
        dm.update_re(0)
        dm.substep(0)
        dm.run(0, 2)
        # GPU -> files -> oracle
        restart.write_flow(dm, 2, directory=str(tmp_path))
        shapes = [dm.shapes[k] for k in ("qx", "qy", "qz", "pr")]
        st = [restart.read_array(str(tmp_path / restart.file_name(k, 2)), sh)
              for k, sh in zip(("qx", "qy", "qz", "pr"), shapes)]
        for a, b in zip(st, dm.get_state()):
            assert np.array_equal(a, b)
        of.set_state(*st)
        for it in (3, 4):
            of.step(it)
        dm.run(2, 2)
        g, r = dm.get_state(), of.get_state()
        errs = [_rel(x, y) for x, y in zip(g[:3], r[:3])]
        print("GPU files -> oracle, 2 more steps: velocity", errs, flush=True)
        assert max(errs) < 1e-12
        # oracle -> files -> GPU
        for k, a in zip(("qx", "qy", "qz", "pr"), of.get_state()):
            restart.write_array(str(tmp_path / restart.file_name(k, 4)), a)
        restart.read_flow(dm, 4, directory=str(tmp_path))
        dm.run(4, 1)
        of.step(5)
        g, r = dm.get_state(), of.get_state()
        errs = [_rel(x, y) for x, y in zip(g[:3], r[:3])]
        print("oracle files -> GPU, 1 more step: velocity", errs, flush=True)
        assert max(errs) < 1e-12
    finally:
        dm.close()
