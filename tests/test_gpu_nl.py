"""GPU parity of the HIP cell-list neighbor list against the FP64 oracle.

Pair sets must be identical after sorting (bit-exact integer triples)."""

import json
import os

import numpy as np
import pytest
import torch

import torch_sim_amd as tsa
from torch_sim_amd.neighbors import NeighborList, SystemIndex
from oracle import nl as onl

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
F64 = torch.float64


def _gpu_nl(pos, cell_row, pbc, cutoff, idx, si=False, dev="cuda"):
    out = tsa.torchsim_nl(pos.to(dev), cell_row.to(dev), pbc.to(dev), cutoff, idx.to(dev), si)
    return onl.sort_pairs(*(t.cpu() for t in out))


def _assert_same(a, b):
    assert a[0].shape == b[0].shape, (a[0].shape, b[0].shape)
    assert torch.equal(a[0], b[0])
    assert torch.equal(a[1], b[1])
    assert torch.equal(a[2].to(F64), b[2].to(F64))


def _cacrp2o7():
    g = GOLD["cacrp2o7"]
    pos = torch.tensor(g["positions"], dtype=F64)
    return pos, torch.tensor(g["cell_row"], dtype=F64).view(1, 3, 3), torch.tensor([[True] * 3]), torch.zeros(len(pos), dtype=torch.long)


@pytest.mark.parametrize("cutoff", [2.0, 5.0, 7.0, 11.0])
@pytest.mark.parametrize("self_interaction", [False, True])
def test_triclinic_pair_set_exact(dev, cutoff, self_interaction):
    pos, cell, pbc, idx = _cacrp2o7()
    ref = onl.brute_force_nl(pos, cell, pbc, cutoff, idx, self_interaction)
    _assert_same(_gpu_nl(pos, cell, pbc, cutoff, idx, self_interaction), ref)


@pytest.mark.parametrize("shift_mode", ["uniform", "per_atom"])
def test_unwrapped_input_pair_set_exact(dev, shift_mode):
    pos, cell, pbc, idx = _cacrp2o7()
    n = pos.shape[0]
    if shift_mode == "uniform":
        ints = torch.tensor(GOLD["lattice_shift_triplets"]["uniform"][1], dtype=F64).expand(n, 3)
    else:
        ar = torch.arange(n)
        ints = torch.stack([(ar % 3) - 1, (ar % 5) - 2, (ar % 4) - 2], 1).to(F64)
    shifted = pos + ints @ cell[0]
    ref = onl.brute_force_nl(shifted, cell, pbc, 5.0, idx)
    _assert_same(_gpu_nl(shifted, cell, pbc, 5.0, idx), ref)


def test_batched_uneven_triclinic_exact(dev):
    st = tsa.workloads.soft_sphere_batch(6, n_min=16, n_max=96, device="cpu", seed=4)
    cell_row = st.cell.mT.contiguous()
    for cutoff in (1.4, 2.3):
        ref = onl.brute_force_nl(st.positions, cell_row, st.pbc, cutoff, st.system_idx)
        _assert_same(_gpu_nl(st.positions, cell_row, st.pbc, cutoff, st.system_idx), ref)


def test_small_cells_many_images_exact(dev):
    # Cu 2x2x2 (7.22 A) and 1x1x1 (3.61 A) cells with rc = 6 A: each pair appears
    # under several image shifts
    st = tsa.workloads.cu_batch(3, device="cpu", reps=[[2, 2, 2], [1, 1, 1], [1, 2, 3]], seed=2)
    ref = onl.brute_force_nl(st.positions, st.cell.mT, st.pbc, 6.0, st.system_idx)
    _assert_same(_gpu_nl(st.positions, st.cell.mT.contiguous(), st.pbc, 6.0, st.system_idx), ref)


def test_per_system_pbc_exact(dev):
    st = tsa.workloads.soft_sphere_batch(3, n_min=20, n_max=40, device="cpu", seed=7)
    pbc = torch.tensor([[True, True, True], [True, False, True], [False, False, False]])
    ref = onl.brute_force_nl(st.positions, st.cell.mT, pbc, 1.8, st.system_idx)
    _assert_same(_gpu_nl(st.positions, st.cell.mT.contiguous(), pbc, 1.8, st.system_idx), ref)


def test_edge_cases_and_empty(dev):
    g = GOLD["nl_edge_case"]
    for p in (True, False):
        for si in (True, False):
            pos = torch.tensor(g["positions"], dtype=F64)
            cell = torch.eye(3, dtype=F64) * g["cell_diag"]
            out = tsa.torchsim_nl(pos.cuda(), cell.cuda(), torch.tensor([p] * 3).cuda(), g["cutoff"],
                                  torch.zeros(2, dtype=torch.long).cuda(), si)
            assert out[0].shape[1] > 0 and (out[1] == 0).all()
    g = GOLD["nl_empty"]
    out = tsa.torchsim_nl(torch.tensor(g["positions"], dtype=F64).cuda(), (torch.eye(3, dtype=F64) * g["cell_diag"]).cuda(),
                          torch.tensor(g["pbc"]).cuda(), g["cutoff"], torch.zeros(2, dtype=torch.long).cuda())
    assert out[0].shape == (2, 0) and out[1].numel() == 0 and out[2].shape == (0, 3)


def test_half_list_matches_full_to_half(dev):
    pos, cell, pbc, idx = _cacrp2o7()
    full = onl.brute_force_nl(pos, cell, pbc, 7.0, idx)
    ref = onl.sort_pairs(*onl.full_to_half_list(*full))
    si = SystemIndex(idx.cuda(), 1)
    nl = NeighborList(7.0, half=True, tpa=2)
    nl.build(pos.cuda(), cell.transpose(1, 2).contiguous().cuda(), pbc.cuda(), si)
    got = onl.sort_pairs(*(t.cpu() for t in nl.pairs(si, F64)))
    _assert_same(got, ref)


@pytest.mark.parametrize("tpa,q", [(1, 1), (4, 1), (8, 2)])
def test_layouts_and_bin_sizes_give_same_set(dev, tpa, q):
    st = tsa.workloads.cu_batch(4, device="cpu", seed=5)
    ref = onl.brute_force_nl(st.positions, st.cell.mT, st.pbc, 6.0, st.system_idx)
    si = SystemIndex(st.system_idx.cuda(), st.n_systems)
    nl = NeighborList(6.0, tpa=tpa, bins_per_rlist=q)
    nl.build(st.positions.cuda(), st.cell.cuda().contiguous(), st.pbc.cuda(), si)
    _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(si, F64))), ref)


def _raw_rows(nl, n):
    """Every row of the ELL list as (count, entries) in thread order."""
    cnt = nl.count[:n].cpu()
    nbr = nl.nbr.cpu()
    rows = []
    for t in range(n):
        c = int(cnt[t])
        idx = [(k // nl.tpa) * (nl.ld * nl.tpa) + t * nl.tpa + k % nl.tpa for k in range(c)]
        rows.append(nbr[idx].tolist())
    return nl.order[:n].cpu().tolist(), rows


@pytest.mark.parametrize("q", [1, 2, 3])
def test_fill_kernels_write_identical_rows(dev, q):
    """The per-bin and per-atom fill kernels write every row with the same
    neighbour words in the same order (stencil image, then atom index), on
    small triclinic cells (many periodic images) and larger ones, full and
    half lists.  Three bins per list radius give 343 stencil images, more
    than one batch of the per-bin fill."""
    st = tsa.workloads.cu_batch(6, device="cpu", seed=8)
    small = tsa.workloads.cu_batch(3, device="cpu", reps=[[2, 2, 2], [1, 1, 1], [1, 2, 3]], seed=2)
    tri = tsa.workloads.soft_sphere_batch(4, n_min=40, n_max=120, device="cpu", seed=3)
    for s_ in (st, small, tri):
        si = SystemIndex(s_.system_idx.cuda(), s_.n_systems)
        for half in (False, True):
            outs = []
            for mode in (1, 2):
                print("fill mode", mode, "q", q, "half", half, "atoms", s_.n_atoms, flush=True)
                nl = NeighborList(6.0, half=half, tpa=4, bins_per_rlist=q, fill_mode=mode)
                nl.build(s_.positions.cuda(), s_.cell.cuda().contiguous(), s_.pbc.cuda(), si)
                outs.append(_raw_rows(nl, s_.n_atoms))
            assert outs[0] == outs[1]
        nl = NeighborList(6.0, tpa=4, bins_per_rlist=q, fill_mode=1)
        nl.build(s_.positions.cuda(), s_.cell.cuda().contiguous(), s_.pbc.cuda(), si)
        ref = onl.brute_force_nl(s_.positions, s_.cell.mT, s_.pbc, 6.0, s_.system_idx)
        _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(si, F64))), ref)


@pytest.mark.parametrize("q", [1, 2, 3])
def test_wave_fill_writes_the_same_stencil_rows(dev, q):
    """Stencil-word lists (the MD engine's): the one-wave-per-bin fill (mode
    3, and modes 4 / 5 with 4 / 8 centre atoms per pass) writes every row
    entry for entry as the per-bin and per-atom fills do, on dense Cu
    systems (bins of up to 10 atoms: one and two passes of either width),
    small cells with many periodic images and triclinic soft-sphere cells."""
    st = tsa.workloads.cu_batch(6, device="cpu", seed=8)
    small = tsa.workloads.cu_batch(3, device="cpu", reps=[[2, 2, 2], [1, 1, 1], [1, 2, 3]], seed=2)
    tri = tsa.workloads.soft_sphere_batch(4, n_min=40, n_max=120, device="cpu", seed=3)
    for s_ in (st, small, tri):
        si = SystemIndex(s_.system_idx.cuda(), s_.n_systems)
        outs = []
        for mode in (1, 2, 3, 4, 5):
            print("stencil words, fill mode", mode, "q", q, "atoms", s_.n_atoms, flush=True)
            nl = NeighborList(6.0, tpa=2, bins_per_rlist=q, fill_mode=mode, word_mode=1)
            nl.build(s_.positions.cuda(), s_.cell.cuda().contiguous(), s_.pbc.cuda(), si)
            outs.append(_raw_rows(nl, s_.n_atoms))
        assert all(o == outs[0] for o in outs[1:])


@pytest.mark.parametrize("mode", [1, 2])
def test_fill_exact_on_edge_cases(dev, mode):
    """Both fill kernels against brute force: unwrapped input ~140 cells from
    the cell (wrap counts too large for the unchecked word path, so rows take
    the checked path while every S' still fits), self interaction, per-system
    pbc, and a dense list (rc = 11 A in Cu, ~470 neighbours)."""
    pos, cell, pbc, idx = _cacrp2o7()
    n = pos.shape[0]
    ar = torch.arange(n)
    far = torch.stack([(ar % 3) + 139, (ar % 2) - 136, (ar % 3) + 149], 1).to(F64)
    shifted = pos + far @ cell[0]
    for p_, si_ in ((shifted, False), (pos, True)):
        print("fill mode", mode, "cacrp2o7 self_interaction", si_, flush=True)
        ref = onl.brute_force_nl(p_, cell, pbc, 5.0, idx, si_)
        sidx = SystemIndex(idx.cuda(), 1)
        nl = NeighborList(5.0, self_interaction=si_, tpa=2, bins_per_rlist=2, fill_mode=mode)
        nl.build(p_.cuda(), cell.transpose(1, 2).contiguous().cuda(), pbc.cuda(), sidx)
        _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(sidx, F64))), ref)
    st = tsa.workloads.soft_sphere_batch(3, n_min=20, n_max=40, device="cpu", seed=7)
    pb = torch.tensor([[True, True, True], [True, False, True], [False, False, False]])
    ref = onl.brute_force_nl(st.positions, st.cell.mT, pb, 1.8, st.system_idx)
    sidx = SystemIndex(st.system_idx.cuda(), st.n_systems)
    print("fill mode", mode, "per-system pbc", flush=True)
    nl = NeighborList(1.8, tpa=2, bins_per_rlist=2, fill_mode=mode)
    nl.build(st.positions.cuda(), st.cell.cuda().contiguous(), pb.cuda(), sidx)
    _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(sidx, F64))), ref)
    big = tsa.workloads.cu_batch(1, device="cpu", reps=[[7, 7, 7]], seed=3)
    ref = onl.sort_pairs(*onl.linked_cell_nl(big.positions, big.cell.mT, big.pbc, 11.0,
                                             big.system_idx))
    sidx = SystemIndex(big.system_idx.cuda(), 1)
    print("fill mode", mode, "dense Cu rc 11", flush=True)
    nl = NeighborList(11.0, tpa=4, bins_per_rlist=2, fill_mode=mode)
    nl.build(big.positions.cuda(), big.cell.cuda().contiguous(), big.pbc.cuda(), sidx)
    _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(sidx, F64))), ref)


@pytest.mark.parametrize("case", ["cu", "dilute", "many"])
def test_one_block_build_matches_grid_build(dev, case):
    """Systems of <= 2048 atoms are binned and sorted by one workgroup each;
    the rows equal those of the grid-wide kernels, including a dilute box with
    more bins than the workgroup counts in LDS (its counters are global).
    Batches of at most 16 systems take the chunked bin scan on the grid path,
    larger ones ("many", 20 systems) the one-workgroup-per-system scan."""
    if case in ("cu", "many"):
        st = tsa.workloads.cu_batch(5 if case == "cu" else 20, device="cpu", seed=9)
        rc = 6.0
        pos, cell_col, pbc, idx, ns = st.positions, st.cell, st.pbc, st.system_idx, st.n_systems
    else:
        g = torch.Generator().manual_seed(11)
        n0, n1 = 150, 90
        box = torch.tensor([60.0, 55.0, 70.0], dtype=F64)
        pos = torch.rand(n0 + n1, 3, generator=g, dtype=F64) * box
        cell_col = torch.diag(box).expand(2, 3, 3).contiguous()
        pbc = torch.tensor([True, True, True])
        idx = torch.cat([torch.zeros(n0, dtype=torch.long), torch.ones(n1, dtype=torch.long)])
        ns, rc = 2, 6.5
    n = pos.shape[0]
    si = SystemIndex(idx.cuda(), ns)
    for q in (1, 2):
        rows = []
        for one_block in (True, False):
            print("build case", case, "q", q, "one_block", one_block, flush=True)
            nl = NeighborList(rc, tpa=2, bins_per_rlist=q, one_block_build=one_block)
            nl.build(pos.cuda(), cell_col.cuda().contiguous(), pbc.cuda(), si)
            rows.append(_raw_rows(nl, n))
        assert rows[0] == rows[1]
    ref = onl.brute_force_nl(pos, cell_col.mT, pbc, rc, idx)
    _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(si, F64))), ref)


def test_wrap_positions_matches_oracle(dev):
    g = GOLD["wrap_batched"]
    cell = torch.tensor(g["cell_row"], dtype=F64)
    pos = torch.tensor(g["positions"], dtype=F64)
    idx = torch.tensor(g["system_idx"])
    pbc = torch.tensor(g["pbc"])
    w, s = tsa.neighbors.wrap_positions(pos.cuda(), cell.cuda(), idx.cuda(), pbc.cuda())
    rw, rs = onl.wrap_positions(pos, cell, idx, pbc)
    torch.testing.assert_close(w.cpu(), rw, atol=1e-12, rtol=0)
    assert torch.equal(s.cpu(), rs)
    g = GOLD["wrap_non_periodic_axis"]
    w, s = tsa.neighbors.wrap_positions(torch.tensor(g["positions"], dtype=F64).cuda(),
                                        torch.tensor(g["cell_row"], dtype=F64).cuda(),
                                        torch.zeros(1, dtype=torch.long).cuda(), torch.tensor(g["pbc"]).cuda())
    assert w.cpu().tolist() == g["expected_wrapped"] and s.cpu().tolist() == g["expected_shifts"]


def test_overflow_retry_resizes(dev):
    st = tsa.workloads.cu_batch(2, device="cpu", seed=8)
    si = SystemIndex(st.system_idx.cuda(), st.n_systems)
    nl = NeighborList(6.0, tpa=1)
    nl.ensure(st.n_atoms, st.n_systems, 8, 2, 11, torch.device("cuda"))
    nl.build(st.positions.cuda(), st.cell.cuda().contiguous(), st.pbc.cuda(), si)
    assert nl.kmax >= 78 and int(nl.count.max()) >= 78
    ref = onl.brute_force_nl(st.positions, st.cell.mT, st.pbc, 6.0, st.system_idx)
    _assert_same(onl.sort_pairs(*(t.cpu() for t in nl.pairs(si, F64))), ref)
