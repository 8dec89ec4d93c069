"""Trajectory files and the reporter on the CPU (torch_sim/trajectory.py),
following the reference's tests/test_trajectory.py: array writes and reads,
state frames, dtype coercion, invalid writes, file modes, truncation,
resumption, a record cut short, and the reporter's frequencies, per-system
files and property values."""

import os

import numpy as np
import pytest
import torch

import torch_sim_amd as tsa
from torch_sim_amd.integrators import MDState
from torch_sim_amd.trajectory import TorchSimTrajectory, TrajectoryReporter

F64 = torch.float64


def _md(n_sys=2, seed=1):
    st = tsa.workloads.cu_batch(n_sys, device="cpu", seed=seed, reps=[(2, 2, 2), (2, 2, 3)][:n_sys])
    g = torch.Generator().manual_seed(seed)
    return MDState.from_state(st, momenta=torch.randn(st.n_atoms, 3, generator=g, dtype=F64),
                              energy=torch.randn(st.n_systems, generator=g, dtype=F64),
                              forces=torch.randn(st.n_atoms, 3, generator=g, dtype=F64))


def test_write_and_read_arrays(tmp_path):
    """test_trajectory.py:57-84: one frame per int step, several with a list."""
    with TorchSimTrajectory(tmp_path / "a.traj", mode="w", metadata={"who": "me"}) as t:
        x = np.arange(12, dtype=np.float32).reshape(4, 3)
        t.write_arrays({"x": x}, 0)
        t.write_arrays({"x": np.stack([x + 1, x + 2])}, [3, 5])
        assert t.array_registry["x"] == ((4, 3), np.dtype(np.float32))
        assert t.get_array("x").shape == (3, 4, 3)
        np.testing.assert_array_equal(t.get_array("x", start=1, stop=3)[1], x + 2)
        np.testing.assert_array_equal(t.get_steps("x"), [0, 3, 5])
        assert t.metadata == {"who": "me"}
        assert t.last_step is None  # no positions array
    with TorchSimTrajectory(tmp_path / "a.traj") as t:  # read back from disk
        np.testing.assert_array_equal(t.get_array("x", step=2)[1], x + 2)


def test_coercion_and_invalid_writes(tmp_path):
    """test_trajectory.py:147-185, 270-368: float64 stored as float32 by
    default (int64 as int32 when asked); shape / dtype changes and
    non-increasing steps are errors."""
    with TorchSimTrajectory(tmp_path / "c.traj", mode="w", coerce_to_int32=True) as t:
        t.write_arrays({"f": np.ones((2, 3)), "i": np.arange(4, dtype=np.int64),
                        "b": np.array([True, False])}, 1)
        assert t.get_array("f").dtype == np.float32 and t.get_array("i").dtype == np.int32
        assert t.get_array("b").dtype == np.bool_
        with pytest.raises(ValueError, match="shape mismatch"):
            t.write_arrays({"f": np.ones((3, 3))}, 2)
        with pytest.raises(ValueError, match="must be greater than the last recorded"):
            t.write_arrays({"f": np.ones((2, 3))}, 1)
        with pytest.raises(ValueError, match="dtype mismatch"):
            t.write_arrays({"i": np.arange(4, dtype=np.float32)}, 3)
        with pytest.raises(ValueError, match="Unsupported"):
            t.write_arrays({"u": np.arange(3, dtype=np.uint8)}, 4)
        with pytest.raises(ValueError, match="Number of steps"):
            t.write_arrays({"f": np.ones((3, 2, 3))}, [5, 6])
    with TorchSimTrajectory(tmp_path / "n.traj", mode="w", coerce_to_float32=False,
                            compress_data=False) as t:
        t.write_arrays({"f": torch.ones(2, dtype=F64)}, 0)
        assert t.get_array("f").dtype == np.float64
    assert os.path.getsize(tmp_path / "n.traj") > 16


def test_file_modes_and_append(tmp_path):
    """test_trajectory.py:248-268: 'w' starts over, 'a' keeps and appends,
    'r' cannot write."""
    p = tmp_path / "m.traj"
    with TorchSimTrajectory(p, mode="w") as t:
        t.write_arrays({"positions": np.zeros((2, 3))}, 1)
    with TorchSimTrajectory(p, mode="a") as t:
        assert t.last_step == 1
        t.write_arrays({"positions": np.ones((2, 3))}, 2)
    with TorchSimTrajectory(p, mode="r") as t:
        assert len(t) == 2 and t.last_step == 2
        with pytest.raises(ValueError, match="read-only"):
            t.write_arrays({"positions": np.ones((2, 3))}, 3)
    with TorchSimTrajectory(p, mode="w") as t:
        assert t.last_step is None


def test_a_record_cut_short_is_ignored(tmp_path):
    """A writer that died mid-record leaves a tail the next open drops."""
    p = tmp_path / "cut.traj"
    with TorchSimTrajectory(p, mode="w", compress_data=False) as t:
        t.write_arrays({"positions": np.zeros((5, 3))}, 1)
        t.write_arrays({"positions": np.ones((5, 3))}, 2)
    size = os.path.getsize(p)
    with open(p, "r+b") as fh:
        fh.truncate(size - 7)
    with TorchSimTrajectory(p, mode="a") as t:
        assert t.last_step == 1
        t.write_arrays({"positions": np.full((5, 3), 3.0)}, 2)
    with TorchSimTrajectory(p) as t:
        np.testing.assert_array_equal(t.get_steps("positions"), [1, 2])
        assert t.get_array("positions")[1, 0, 0] == 3.0


def test_write_state_and_get_state(tmp_path):
    """test_trajectory.py:85-146, 464-503: positions per frame, optional
    velocities / forces, the cell per frame by default, masses and atomic
    numbers once, pbc as a global array; get_state rebuilds the frame."""
    md = _md(1)
    with TorchSimTrajectory(tmp_path / "s.traj", mode="w", coerce_to_float32=False) as t:
        t.write_state(md, 0, save_velocities=True, save_forces=True)
        md2 = md.clone()
        md2.positions = md2.positions + 0.5
        t.write_state([md2, md2], [1, 2])
        assert len(t) == 3
        np.testing.assert_array_equal(t.get_steps("masses"), [0])
        np.testing.assert_array_equal(t.get_steps("cell"), [0, 1, 2])
        assert t.get_array("velocities").shape == (1, md.n_atoms, 3)
        st = t.get_state(1, device=torch.device("cpu"))
        assert torch.equal(st.positions, md2.positions)
        assert torch.equal(st.cell, md.cell) and torch.equal(st.masses, md.masses)
        assert st.pbc.tolist() == [True, True, True]
        with pytest.raises(ValueError, match="velocities requested"):
            t.write_state(tsa.workloads.cu_batch(1, device="cpu", reps=[(2, 2, 2)]), 3,
                          save_velocities=True)
        with pytest.raises(ValueError, match="System index must be specified"):
            t.write_state(_md(2), 4)
        with pytest.raises(ImportError):
            t.get_atoms()


def test_truncate_to_step(tmp_path):
    """test_trajectory.py:960-1017: frames after the step go, arrays written
    once at step 0 stay (pbc keeps its three entries)."""
    p = tmp_path / "t.traj"
    md = _md(1)
    with TorchSimTrajectory(p, mode="w") as t:
        for k in range(6):
            t.write_state(md, k)
            t.write_arrays({"v": md.momenta}, k)
    with TorchSimTrajectory(p, mode="a") as t:
        t.truncate_to_step(3)
        for name in t.array_registry:
            n, steps = 4, [0, 1, 2, 3]
            if name in ("atomic_numbers", "masses"):
                n, steps = 1, [0]
            if name == "pbc":
                n, steps = 3, [0]
            assert len(t.get_array(name)) == n, name
            np.testing.assert_array_equal(t.get_steps(name), steps)
        with pytest.raises(ValueError, match=r"self\.last_step=3 < step=10"):
            t.truncate_to_step(10)
        with pytest.raises(ValueError, match="Step must be larger than 0"):
            t.truncate_to_step(0)
        t.write_state(md, 4)
    with TorchSimTrajectory(p) as t:
        np.testing.assert_array_equal(t.get_steps("positions"), [0, 1, 2, 3, 4])


def test_reporter_frequencies_files_and_properties(tmp_path):
    """test_trajectory.py:553-758: one file per system; states every
    state_frequency steps, each calculator at its own frequency; calculators
    of one argument get the model argument; the per-system values come back;
    without filenames only the values."""
    md = _md(2)
    files = [tmp_path / f"r{k}.traj" for k in range(2)]
    calls = []

    def ke(state):
        calls.append(state.n_systems)
        return 0.5 * (state.momenta ** 2 / state.masses[:, None]).sum()

    rep = TrajectoryReporter(files, state_frequency=2, prop_calculators={
        1: {"ke": ke}, 3: {"e": lambda state, model: state.energy}})
    for step in range(0, 7):
        out = rep.report(md, step)
        assert len(out) == 2 and "ke" in out[0] and (("e" in out[1]) == (step % 3 == 0))
    rep.close()
    assert set(calls) == {1}  # a user calculator sees one system at a time
    for k, f in enumerate(files):
        with TorchSimTrajectory(f) as t:
            np.testing.assert_array_equal(t.get_steps("positions"), [0, 2, 4, 6])
            np.testing.assert_array_equal(t.get_steps("ke"), range(7))
            np.testing.assert_array_equal(t.get_steps("e"), [0, 3, 6])
            sel = md.system_idx == k
            want = 0.5 * (md.momenta[sel] ** 2 / md.masses[sel][:, None]).sum()
            assert abs(float(t.get_array("ke")[2, 0]) - float(want)) < 1e-4 * float(want)
    # without files every calculator sees the whole batch (trajectory.py:342-394):
    # a batch total goes to each system
    props = TrajectoryReporter(None, prop_calculators={1: {"ke": ke}}).report(md, 5)
    total = 0.5 * (md.momenta ** 2 / md.masses[:, None]).sum()
    assert len(props) == 2 and props[1]["ke"].shape == (1,)
    assert torch.allclose(props[0]["ke"], total) and torch.allclose(props[1]["ke"], total)
    with pytest.raises(ValueError, match="doesn't match number of trajectory files"):
        TrajectoryReporter(files[:1]).report(md, 0)
    with pytest.raises(ValueError, match="unique"):
        TrajectoryReporter([files[0], files[0]])


def test_reporter_truncate_reopen_and_last_steps(tmp_path):
    """test_trajectory.py:1019-1068: truncation across files, bounds; the
    files can be re-pointed (the in-flight optimizer does so)."""
    md = _md(2)
    files = [tmp_path / f"q{k}.traj" for k in range(2)]
    rep = TrajectoryReporter(files, state_frequency=1)
    for step in range(6):
        rep.report(md, step)
    assert rep.last_steps == [5, 5]
    with pytest.raises(ValueError, match=r"Step 7 is greater than the minimum last step"):
        rep.truncate_to_step(7)
    with pytest.raises(ValueError, match="Step must be greater than 0"):
        rep.truncate_to_step(-2)
    rep.truncate_to_step(3)
    assert rep.last_steps == [3, 3]
    rep.reopen_trajectories([files[1]])
    rep.report(select_first(md), 4)
    assert rep.last_steps == [4]
    rep.close()
    with TorchSimTrajectory(files[0]) as t0, TorchSimTrajectory(files[1]) as t1:
        assert t0.last_step == 3 and t1.last_step == 4


def select_first(md):
    return tsa.select_systems(md, [1])
