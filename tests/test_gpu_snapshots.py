"""Snapshot stream on the device: states written by SnapshotWriter during a run
equal, bit for bit, the states of plain runs of the same number of steps."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def _vortex():
    import atum_hip as A
    K, N = 8, 4
    v = P.julia_range(-5, 5, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    dg = A.DGSEM(law=A.EulerLaw(2), grid=grid,
                 volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RusanovFlux(), precision="fast")
    q0 = P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0))
    return dg, q0, 0.1 * (10 / K) / N


@pytest.mark.parametrize("ring", [1, 3])
def test_snapshots_equal_plain_runs(tmp_path, ring):
    import atum_hip as A
    dg, q0, dt = _vortex()
    q = dg.fieldarray(q0)
    solver = A.LSRK54(dg, q, dt)
    w = A.SnapshotWriter(q, str(tmp_path / "v_{step:04d}.atm"), every=5, ring=ring,
                         meta={"law": "euler"}, solver=solver)
    for step in range(1, 21):  # solve!'s loop without the clipped final step
        t = solver.time
        solver.dostep(q)
        w(step, t, q)
    files = w.close()
    assert len(files) == 4
    for f in sorted(files):
        got, head = A.read_snapshot(f)
        n = head["step"]
        ref = dg.fieldarray(q0)
        A.LSRK54(dg, ref, dt).run(ref, n)
        dg.ctx.sync()
        np.testing.assert_array_equal(got, ref.numpy())
        assert head["law"] == "euler" and head["time"] == pytest.approx(n * dt, rel=1e-12)
    np.testing.assert_array_equal(A.read_snapshot(sorted(files)[-1])[0], q.numpy())


def test_snapshot_without_staging_and_query():
    """Direct copy (no staging buffer): the compute stream waits for the copy."""
    import atum_hip as A
    from atum_hip import _lib
    from atum_hip.snapshots import PinnedBuffer
    dg, q0, dt = _vortex()
    q = dg.fieldarray(q0)
    host = PinnedBuffer(dg.ctx, q.shape)
    dg.ctx.check(_lib.lib.atm_snapshot(dg.ctx.h, 5, host.ptr, None, q.ptr, q.nbytes), "snapshot")
    A.LSRK54(dg, q, dt).run(q, 3)  # may start only after the copy has read q
    dg.ctx.check(_lib.lib.atm_snapshot_wait(dg.ctx.h, 5), "wait")
    assert _lib.lib.atm_snapshot_query(dg.ctx.h, 5) == 1
    np.testing.assert_array_equal(host.array, q0)


def test_checkpoint_restart_lsrk_and_ark(tmp_path):
    """Checkpoint/resume: a run stopped at a snapshot and continued from it (new
    field, new solver with t0 = the snapshot's time) ends bit for bit where the
    uninterrupted run ends, for LSRK54 and for IMEX ARK23 with banded solves."""
    import atum_hip as A
    from serve_requests import _imex
    dg, q0, dt = _vortex()
    ex, lin, qi0, dti = _imex(2, 3, 3, 7)({})
    cases = [("lsrk", dg, q0, dt, lambda d, q, t0: A.LSRK54(d, q, dt, t0=t0)),
             ("ark", lin, qi0, dti, lambda d, q, t0: A.ARK23(ex, lin, q, dti, t0=t0, split_rhs=False))]
    for name, d, qa, h, make in cases:
        full = d.fieldarray(qa)
        make(d, full, 0.0).run(full, 12)
        q = d.fieldarray(qa)
        solver = make(d, q, 0.0)
        w = A.SnapshotWriter(q, str(tmp_path / (name + "_{step:04d}.atm")), every=7, solver=solver)
        for step in range(1, 8):
            t = solver.time
            solver.dostep(q)
            w(step, t, q)
        (path,) = w.close()
        q2, step, t0 = A.load_checkpoint(path, d)
        assert step == 7 and t0 == solver.time
        make(d, q2, t0).run(q2, 5)
        d.ctx.sync()
        np.testing.assert_array_equal(q2.numpy(), full.numpy(), err_msg=name)


def test_vtk_snapshots(tmp_path):
    """fmt="vtk": the snapshot stream writes .vtu files of q - q_ref (interpolated to
    equally spaced points, as the reference's experiments write) and a .pvd of the run."""
    import xml.etree.ElementTree as ET
    import atum_hip as A
    from atum_hip import vtk
    dg, q0, dt = _vortex()
    q = dg.fieldarray(q0)
    solver = A.LSRK54(dg, q, dt)
    ref = np.ones_like(q0)
    w = A.SnapshotWriter(q, str(tmp_path / "v_{step:03d}.vtu"), every=3, solver=solver, fmt="vtk", grid=dg.grid,
                         names=["rho", "rhou", "rhov", "rhoe"], reference=ref, pvd=str(tmp_path / "v.pvd"))
    for step in range(1, 7):
        t = solver.time
        solver.dostep(q)
        w(step, t, q)
    files = w.close()
    assert len(files) == 2
    last = ET.parse(sorted(files)[-1]).getroot()
    got = {da.get("Name"): np.array(da.text.split(), dtype=float) for da in last.iter("DataArray") if da.get("Name")}
    want = vtk.to_equispaced((q.numpy()[0] - 1.0).reshape(dg.grid.Np, dg.grid.K, order="F"), dg.grid.N, 2)
    np.testing.assert_array_equal(got["rho"], want.reshape(-1, order="F"))
    pvd = ET.parse(str(tmp_path / "v.pvd")).getroot().findall("Collection/DataSet")
    assert [float(d.get("timestep")) for d in pvd] == pytest.approx([3 * dt, 6 * dt], rel=1e-12)
