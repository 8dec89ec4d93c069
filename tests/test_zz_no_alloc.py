"""tsx_nl_build_noalloc (the list build whose workspace is the library's
static device array) against tsx_nl_build with a caller-allocated
workspace: the same rows, row order, displacement references, bin-relative
frame and status word, bit for bit, on two calls in a row where the second
input is smaller than the first (so it runs on what the first call left in
the array), plus the host-side refusal of sizes beyond the array."""

import ctypes

import pytest
import torch

import torch_sim_amd as tsa
from torch_sim_amd import _lib
from torch_sim_amd.neighbors import NeighborList, SystemIndex, plan_capacity, jbits_for

pytestmark = pytest.mark.gpu


def _outputs(nl, n, n_sys, dev, stencil):
    o = {
        "nbr": torch.full_like(nl.nbr, -1),
        "count": torch.full_like(nl.count, -1),
        "order": torch.full_like(nl.order, -1),
        "xref": torch.full((max(1, n), 4), float("nan"), dtype=torch.float64, device=dev),
        "status": torch.zeros(1, dtype=torch.int32, device=dev),
    }
    if stencil:
        o["anchor"] = torch.full((n, 4), float("nan"), dtype=torch.float64, device=dev)
        o["rel"] = torch.full((n, 4), float("nan"), dtype=torch.float32, device=dev)
        o["ustep"] = torch.full((n_sys, 12), float("nan"), dtype=torch.float64, device=dev)
        o["frame"] = _lib.NLFrame(o["anchor"].data_ptr(), o["rel"].data_ptr(),
                                  o["ustep"].data_ptr())
    return o


def _build(entry, nl, st, si, pos, out):
    frame = out.get("frame")
    args = [
        _lib.ptr(pos), 4, _lib.ptr(st.cell), _lib.ptr(st.pbc.to(torch.uint8).contiguous()), 0,
        _lib.ptr(si.sys_start), _lib.ptr(si.atom_sys), si.n_atoms, si.n_systems, nl.r_list,
        nl.bins_per_rlist, 0, 0, nl.jbits, nl.tpa, nl.kmax, nl.ld, nl.bin_cap,
    ]
    if entry == "tsx_nl_build":
        args += [1, _lib.ptr(nl.ws)]
    args += [
        _lib.ptr(out["nbr"]), _lib.ptr(out["count"]), _lib.ptr(out["order"]), _lib.ptr(out["xref"]),
        _lib.ptr(out["status"]), None, None, None, nl.fill_mode, nl.word_mode,
        None if frame is None else ctypes.addressof(frame), si.max_atoms, _lib.stream(),
    ]
    _lib.call(entry, *args)


def _valid_entries(nl, count):
    """Flat indices of the stored entries of every row and their mask."""
    lt = nl.tpa.bit_length() - 1
    dev = count.device
    t = torch.arange(nl.ld, device=dev).view(-1, 1)
    k = torch.arange(nl.kmax, device=dev).view(1, -1)
    flat = (k >> lt) * (nl.ld << lt) + (t << lt) + (k & (nl.tpa - 1))
    return flat, k < count[: nl.ld].view(-1, 1)


def _case(dev, st, stencil, fill_mode):
    n, n_sys = st.n_atoms, st.n_systems
    si = SystemIndex(st.system_idx, n_sys)
    pos = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    pos[:, :3] = st.positions
    nl = NeighborList(6.6, tpa=2, bins_per_rlist=2, fill_mode=fill_mode,
                      word_mode=1 if stencil else 0)
    bins, kmax = plan_capacity(st.cell.cpu().numpy(), st.pbc.cpu().numpy(), si.counts, 6.6, 2)
    nl.ensure(n, n_sys, kmax, bins + 64, jbits_for(si.max_atoms), dev)
    ref = _outputs(nl, n, n_sys, dev, stencil)
    got = _outputs(nl, n, n_sys, dev, stencil)
    _build("tsx_nl_build", nl, st, si, pos, ref)
    _build("tsx_nl_build_noalloc", nl, st, si, pos, got)
    torch.cuda.synchronize()
    print(f"case: {n_sys} systems, {n} atoms, stencil words {stencil}, fill mode {fill_mode}, "
          f"status {int(ref['status'].item())}", flush=True)
    assert int(ref["status"].item()) == 0
    for name in ("count", "order", "status"):
        assert torch.equal(got[name], ref[name]), name
    assert torch.equal(got["xref"][:n], ref["xref"][:n])
    flat, mask = _valid_entries(nl, ref["count"])
    assert torch.equal(got["nbr"][flat][mask], ref["nbr"][flat][mask])
    assert int(mask.sum()) > 50 * n
    if stencil:
        for name in ("anchor", "rel", "ustep"):
            assert torch.equal(got[name], ref[name]), name


def test_noalloc_build_equals_workspace_build_on_two_calls(dev):
    # first call: two systems above 2048 atoms (grid-wide build), per-atom fill
    big = tsa.workloads.cu_batch(3, device=dev, seed=51, reps=[[9, 9, 9], [3, 3, 3], [10, 9, 8]])
    _case(dev, big, stencil=True, fill_mode=2)
    # second call, smaller: one workgroup per system, per-bin fill, shift words
    small = tsa.workloads.cu_batch(5, device=dev, seed=52)
    assert small.n_atoms < big.n_atoms
    _case(dev, small, stencil=False, fill_mode=1)


def test_noalloc_build_refuses_sizes_beyond_its_array(dev):
    too_many = 524288 + 1
    with pytest.raises(_lib.TsxError, match="n_atoms"):
        _lib.call("tsx_nl_build_noalloc", None, 3, None, None, 0, None, None, too_many, 1, 6.0,
                  2, 0, 0, 20, 2, 2, too_many, 1, None, None, None, None, None, None, None, None,
                  0, 0, None, 0, _lib.stream())
    with pytest.raises(_lib.TsxError, match="bin_cap"):
        _lib.call("tsx_nl_build_noalloc", None, 3, None, None, 0, None, None, 10, 1, 6.0,
                  2, 0, 0, 20, 2, 2, 10, 524288 + 1, None, None, None, None, None, None, None,
                  None, 0, 0, None, 0, _lib.stream())
