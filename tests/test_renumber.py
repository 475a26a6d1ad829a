"""Renumbering at world size 1 (CPU): the compiled edge order against the
oracle's stable sort, row-wise summation order preserved under RCM, and the
oracle's sort itself against numpy's stable argsort."""
import numpy as np
import pytest
import torch

from easier_amd.core.compiler import compile_program
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import swe as wswe
from oracle import renumber
from oracle import swe as oswe


def test_edge_order_is_a_stable_sort():
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 17, size=400)
    o = renumber.edge_order(idx, 17)
    assert np.array_equal(o, np.argsort(idx, kind='stable'))
    # a planted mistake (an unstable order) is told apart
    bad = np.argsort(-np.arange(idx.size), kind='stable')[np.argsort(idx[::-1], kind='stable')]
    assert not np.array_equal(bad, o)


def _swe(n, reorder):
    mesh = wmesh.tri2(n)
    geo = oswe.geometry(mesh, centre=(0.5, 0.5), width=100.0)
    eqn = wswe.ShallowWaterEquation(mesh, geo, dt=1e-3, g=9.81)
    prog = compile_program([eqn], reorder, torch.device('cpu'))
    return mesh, eqn, prog


@pytest.mark.parametrize('reorder', ['none', 'rcm'])
def test_compiled_edge_order_keeps_reference_row_order(reorder):
    mesh, eqn, prog = _swe(9, reorder)
    nc = mesh['cells'].shape[0]
    cell_order = eqn.h.easier_order
    cell_order = np.arange(nc) if cell_order is None else cell_order.numpy()
    assert np.array_equal(np.sort(cell_order), np.arange(nc))
    inv = np.empty(nc, np.int64)
    inv[cell_order] = np.arange(nc)
    for red, user_idx in ((eqn.scatter, mesh['dst']), (eqn.scatter_b, mesh['bcells'])):
        spaces = [e for e, r in prog.bound_reducer.items() if r is red]
        assert len(spaces) == 1
        o = prog.space_order.get(spaces[0])
        o = np.arange(user_idx.shape[0]) if o is None else np.asarray(o)
        # compiled row of every compiled edge, sorted
        rows = red.easier_host_idx
        assert np.array_equal(rows, inv[np.asarray(user_idx)[o]])
        assert np.all(np.diff(rows) >= 0)
        # within every row the user's edges keep their relative order
        for lst in renumber.row_edge_lists(user_idx, o, nc):
            assert lst == sorted(lst)
        if reorder == 'none':
            assert np.array_equal(o, renumber.edge_order(user_idx, nc))
