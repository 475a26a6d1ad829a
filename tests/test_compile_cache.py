"""Compile cache (esr.dump / compile(load_dir=...), reference dump.py:341-503,
824-918): a dump of a compiled program gives the same renumbering without
running it again; a changed program, a missing or a corrupted dump compiles
from scratch."""
import json
import os

import numpy as np
import pytest
import torch

from easier_amd.core import cache, compiler
from easier_amd.core.compiler import compile_program
from easier_amd.workloads import mesh as wmesh
from easier_amd.workloads import swe as wswe
from oracle import swe as oswe


def _program(n=12, amp=0.1):
    mesh = wmesh.tri2(n)
    geo = oswe.geometry(mesh, amp=amp)
    return wswe.ShallowWaterEquation(mesh, geo, dt=0.001, g=9.81)


def _compile(m, cache_=None):
    prog = compile_program([m], 'rcm', torch.device('cpu'), cache_)
    for mod in prog.modules:
        mod.easier_program = prog
    return prog


def _orders(prog):
    return {prog.space_name[k]: v for k, v in prog.space_order.items()}


def _indices(prog):
    out = {}
    for m in prog.modules:
        for name, obj in m.named_modules():
            if hasattr(obj, 'easier_host_idx'):
                out[name] = np.asarray(obj.easier_host_idx)
    return out


def test_dump_and_reload_gives_the_same_program(tmp_path, monkeypatch):
    m1 = _program()
    p1 = _compile(m1)
    assert any(k == 'rcm' for k in p1.space_kind.values())
    assert any(k == 'edge' for k in p1.space_kind.values())
    cache.dump([m1], str(tmp_path))
    meta = json.load(open(tmp_path / 'jit' / cache.JSON_NAME))
    assert sorted(meta['spaces']) == sorted(_orders(p1))

    # the renumbering is not run again
    def no_rcm(*a, **k):
        raise AssertionError('RCM ran although the cache matches')
    monkeypatch.setattr(compiler, 'rcm_order', no_rcm)
    m2 = _program()
    p2 = _compile(m2, cache.load(str(tmp_path)))
    o1, o2 = _orders(p1), _orders(p2)
    assert o1.keys() == o2.keys()
    for k in o1:
        np.testing.assert_array_equal(o1[k], o2[k])
    kinds = lambda p: {p.space_name[k]: v for k, v in p.space_kind.items()}
    assert kinds(p1) == kinds(p2)
    i1, i2 = _indices(p1), _indices(p2)
    assert i1.keys() == i2.keys()
    for k in i1:
        np.testing.assert_array_equal(i1[k], i2[k])
    for t1, t2 in ((m1.h, m2.h), (m1.uh, m2.uh), (m1.area, m2.area)):
        assert torch.equal(t1.data, t2.data)
    # the same edge spaces are bound to their Reducers
    assert sorted(p1.space_name[k] for k in p1.bound_reducer) == \
        sorted(p2.space_name[k] for k in p2.bound_reducer)


def test_changed_program_or_bad_dump_compiles_from_scratch(tmp_path, monkeypatch):
    m1 = _program()
    _compile(m1)
    cache.dump([m1], str(tmp_path))
    calls = []
    real = compiler.rcm_order

    def counting(*a, **k):
        calls.append(1)
        return real(*a, **k)
    monkeypatch.setattr(compiler, 'rcm_order', counting)
    # a different mesh: other indices, so the fingerprint differs
    p = _compile(_program(n=14), cache.load(str(tmp_path)))
    assert calls and p.fingerprint != json.load(
        open(tmp_path / 'jit' / cache.JSON_NAME))['fingerprint']
    # no dump at all
    calls.clear()
    assert cache.load(str(tmp_path / 'nothing')) is None
    # a corrupted array file
    with open(tmp_path / 'jit' / cache.NPZ_NAME, 'wb') as f:
        f.write(b'not an npz')
    assert cache.load(str(tmp_path)) is None
    _compile(_program(), None)
    assert calls


def test_dump_needs_compiled_modules(tmp_path):
    from easier_amd.core.utils import EasierJitException
    with pytest.raises(EasierJitException):
        cache.dump([_program()], str(tmp_path))
