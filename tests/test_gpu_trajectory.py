"""Trajectory reporting from integrate() and optimize() on the GPU, after
the reference's tests/test_trajectory.py:863-1175: step 0 is written, runs
append and resume after the files' last step, uneven files refuse to resume,
truncation across files; the fused engine's reported frames are the states
of an engine run stopped at those steps (bit for bit, FP64 files), and the
per-step path's frames are its states."""

import numpy as np
import pytest
import torch

import torch_sim_amd as tsa
from torch_sim_amd.trajectory import TorchSimTrajectory, TrajectoryReporter

pytestmark = pytest.mark.gpu
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)


def _setup(dev, seed=81):
    st = tsa.workloads.cu_batch(2, device=dev, seed=seed, reps=[(3, 3, 3), (3, 3, 4)])
    return st, tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)


def _steps(f):
    with TorchSimTrajectory(f) as t:
        return t.get_steps("positions").tolist()


def test_integrate_writes_step_zero_appends_and_resumes(dev, tmp_path):
    st, model = _setup(dev)
    files = [tmp_path / f"i{k}.traj" for k in range(2)]
    kw = dict(model=model, timestep=0.001, temperature=300.0,
              integrator=tsa.Integrator.nvt_langevin)
    out = tsa.integrate(st, n_steps=5, trajectory_reporter=TrajectoryReporter(
        files, state_frequency=1), **kw)
    assert all(_steps(f) == list(range(6)) for f in files)
    rep2 = TrajectoryReporter(files, state_frequency=1, trajectory_kwargs=dict(mode="a"))
    tsa.integrate(out, n_steps=7, trajectory_reporter=rep2, **kw)
    assert all(_steps(f) == list(range(13)) for f in files)
    rep2.truncate_to_step(min(rep2.last_steps))
    assert rep2.last_steps == [12, 12]
    with TorchSimTrajectory(files[1], mode="a") as t:  # make the files uneven
        t.truncate_to_step(3)
    with pytest.raises(ValueError, match="Cannot resume integration from inconsistent states"):
        tsa.integrate(st, n_steps=4, trajectory_reporter=TrajectoryReporter(
            files, state_frequency=1, trajectory_kwargs=dict(mode="a")), **kw)


def test_integrate_default_properties_from_a_dict(dev, tmp_path):
    """A dict reporter gets the default MD properties at prop_frequency 10
    and states every 100 steps (runners.py:34-70, 338-343)."""
    st, model = _setup(dev, seed=82)
    files = [str(tmp_path / f"d{k}.traj") for k in range(2)]
    tsa.integrate(st, model=model, timestep=0.002, temperature=500.0, n_steps=30,
                  integrator=tsa.Integrator.nvt_langevin,
                  trajectory_reporter={"filenames": files})
    with TorchSimTrajectory(files[0]) as t:
        assert t.get_steps("positions").tolist() == [0]
        for name in ("kinetic_energy", "potential_energy", "temperature"):
            assert t.get_steps(name).tolist() == [0, 10, 20, 30], name
        temp = t.get_array("temperature")
        assert (temp > 50).all() and (temp < 5000).all()


@pytest.mark.parametrize("engine", [True, False])
def test_reported_frames_are_the_run_states(dev, tmp_path, engine):
    """The frames integrate() writes are the states of the same run stopped
    at those steps: the fused engine (cut at the report steps, each closed
    there) against a BatchedMD run in chunks, the per-step path against its
    own step loop."""
    st, model = _setup(dev, seed=83)
    files = [tmp_path / f"e{k}.traj" for k in range(2)]
    rep = TrajectoryReporter(files, state_frequency=10, prop_calculators={5: {
        "potential_energy": lambda s: s.energy}},
        trajectory_kwargs=dict(coerce_to_float32=False))
    st.rng = 9
    kT = 800.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    tsa.integrate(st, model=model, timestep=0.002, temperature=800.0, n_steps=30,
                  integrator=tsa.Integrator.nvt_langevin, trajectory_reporter=rep,
                  engine=engine, seed=4, skin=0.5)
    s2, _ = _setup(dev, seed=83)
    s2.rng = 9
    md = tsa.nvt_langevin_init(s2, model, kT=kT)
    frames, energies = {}, {}
    if engine:
        eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.5, seed=4,
                            tpa=model.tpa)
        for k in range(1, 31):
            eng.run(1)
            if k % 5 == 0:
                s = eng.state()
                frames[k], energies[k] = s.positions, s.energy
    else:
        for k in range(1, 31):
            md = tsa.nvt_langevin_step(md, model, dt=dt, kT=kT)
            if k % 5 == 0:
                frames[k], energies[k] = md.positions.clone(), md.energy.clone()
    for s, f in enumerate(files):
        with TorchSimTrajectory(f) as t:
            assert t.get_steps("positions").tolist() == [0, 10, 20, 30]
            assert t.get_steps("potential_energy").tolist() == [0, 5, 10, 15, 20, 25, 30]
            pos = t.get_array("positions")
            sel = (s2.system_idx == s).cpu()
            for j, k in enumerate((10, 20, 30)):
                np.testing.assert_array_equal(pos[j + 1], frames[k].cpu()[sel].numpy())
            pe = t.get_array("potential_energy")
            for j, k in enumerate((5, 10, 15, 20, 25, 30)):
                assert pe[j + 1, 0] == energies[k][s].item()


def test_optimize_writes_step_zero_appends_and_follows_the_refill(dev, tmp_path):
    """optimize() (runners.py:669-743): step 0, then every step of every
    system in its own file while systems leave and enter the batch; a second
    run appends after the last steps."""
    st, model = _setup(dev, seed=84)
    files = [tmp_path / f"o{k}.traj" for k in range(2)]
    rep = TrajectoryReporter(files, state_frequency=1)
    out = tsa.optimize(st, model, max_steps=5, optimizer=tsa.Optimizer.fire,
                       trajectory_reporter=rep, steps_between_swaps=100)
    assert all(_steps(f) == list(range(6)) for f in files)
    rep2 = TrajectoryReporter(files, state_frequency=1, trajectory_kwargs=dict(mode="a"))
    tsa.optimize(out, model, max_steps=7, optimizer=tsa.Optimizer.fire,
                 trajectory_reporter=rep2, steps_between_swaps=100)
    assert all(_steps(f) == list(range(8)) for f in files)
    # three systems through an 80-atom budget (two in flight, the third refilled
    # when they finish): each file follows its system
    st3 = tsa.workloads.cu_batch(3, device=dev, seed=85, reps=[(2, 2, 2), (2, 2, 3), (2, 3, 3)])
    files3 = [tmp_path / f"f{k}.traj" for k in range(3)]
    tsa.optimize(st3, model, max_steps=4, optimizer=tsa.Optimizer.fire,
                 trajectory_reporter=TrajectoryReporter(files3, state_frequency=2),
                 steps_between_swaps=2, autobatcher=tsa.InFlightAutoBatcher(max_atoms=80))
    assert all(_steps(f) == [0, 2, 4] for f in files3)
    for k, f in enumerate(files3):
        with TorchSimTrajectory(f) as t:
            assert t.get_array("positions").shape[1] == int((st3.system_idx == k).sum())
