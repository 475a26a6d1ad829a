"""Recognition of the fused solvers' operator and preconditioner in the
traced forwards of the solver modules.

The reference's CG and GMRES take any operator and any callable `M`
(easier/numeric/solver/cg.py:74-104, gmres.py:139-194); tracing inlines a
plain nn.Module, a function or a lambda into the solver modules' graphs
(easier/core/jit.py:108-118). The fused device iterations apply when the
traced graph computes

    A(x) = reducer(selector(x) * Ae) + x * Av        (numeric/linsys.py:23-24)

with esr.Tensors Ae (edges) and Av (rows) -- operand order of the products
and of the sum free -- and

    M(r) = r               or        M(r) = r / d

with d a partition tensor, whatever object produced those operations
(numeric.Linsys, preconditioner.Jacobi, a user module or a lambda).
"""
import operator
from typing import Optional

import torch
from torch import fx

from ...core import module as esr
from ...core.compiler import resolve_attr

_MUL = {operator.mul, torch.mul}
_ADD = {operator.add, torch.add}
_DIV = {operator.truediv, torch.div, torch.true_divide}


def _attrs(module, graph: fx.Graph):
    env = {}
    for node in graph.nodes:
        if node.op == 'get_attr':
            try:
                env[node] = resolve_attr(module, node.target)
            except AttributeError:
                pass
    return env


def _base(node):
    """Look through full-slice views (x[:], x[...])."""
    while isinstance(node, fx.Node) and node.op == 'call_function' and \
            node.target is operator.getitem and len(node.args) == 2 and \
            node.args[1] in (slice(None), Ellipsis):
        node = node.args[0]
    return node


def _is(node, op, targets):
    if not isinstance(node, fx.Node) or node.op != op:
        return False
    if op == 'call_method':
        return node.target in targets
    return node.target in targets


def _binary(node, fns, methods):
    """(a, b) of a binary call node (function or method form), else None."""
    if _is(node, 'call_function', fns) and len(node.args) == 2 and not node.kwargs:
        return node.args
    if _is(node, 'call_method', methods) and len(node.args) == 2 and not node.kwargs:
        return node.args
    return None


def _tensor(env, node):
    node = _base(node)
    t = env.get(node) if isinstance(node, fx.Node) else None
    return t if isinstance(t, esr.Tensor) and t.is_partition else None


def match_linsys(module, graph: fx.Graph, x: esr.Tensor) -> Optional[dict]:
    """{'selector', 'reducer', 'Ae', 'Av'} of the first
    reducer(selector(x) * Ae) + x * Av in `graph`, or None."""
    env = _attrs(module, graph)

    def is_x(n):
        return _tensor(env, n) is x

    for node in graph.nodes:
        ab = _binary(node, _ADD, {'add', '__add__'})
        if ab is None:
            continue
        for red_n, diag_n in (ab, ab[::-1]):
            if not (isinstance(red_n, fx.Node) and red_n.op == 'call_module'):
                continue
            red = module.get_submodule(red_n.target)
            if not isinstance(red, esr.Reducer) or red_n.kwargs or len(red_n.args) != 1:
                continue
            d = _binary(diag_n, _MUL, {'mul', '__mul__'})
            if d is None:
                continue
            Av = _tensor(env, d[1]) if is_x(d[0]) else \
                (_tensor(env, d[0]) if is_x(d[1]) else None)
            if Av is None or Av is x:
                continue
            m = _binary(red_n.args[0], _MUL, {'mul', '__mul__'})
            if m is None:
                continue
            for sel_n, ae_n in (m, m[::-1]):
                if not (isinstance(sel_n, fx.Node) and sel_n.op == 'call_module'):
                    continue
                sel = module.get_submodule(sel_n.target)
                if not isinstance(sel, esr.Selector) or len(sel_n.args) != 1 or \
                        not is_x(sel_n.args[0]):
                    continue
                Ae = _tensor(env, ae_n)
                if Ae is None:
                    continue
                return dict(selector=sel, reducer=red, Ae=Ae, Av=Av)
    return None


def match_preconditioner(module, graph: fx.Graph, value: fx.Node, arg_ok):
    """What M computed at `value` (the traced M(arg) node): ('identity',
    None) when value is the argument itself (arg_ok(value)), ('jacobi', d)
    for arg / d with d a partition tensor, else None."""
    env = _attrs(module, graph)
    if arg_ok(value):
        return ('identity', None)
    ab = _binary(value, _DIV, {'div', 'truediv', '__truediv__'})
    if ab is not None and arg_ok(ab[0]):
        d = _tensor(env, ab[1])
        if d is not None:
            return ('jacobi', d)
    return None


def cg_preconditioner(update_module, graph: fx.Graph, r: esr.Tensor, z: esr.Tensor):
    """M of CG's Update (cg.py:66-71: z1 = M(r); ...; self.z[:] = z1)."""
    env = _attrs(update_module, graph)
    for node in graph.nodes:
        if node.op == 'call_function' and node.target is operator.setitem and \
                _tensor(env, node.args[0]) is z:
            return match_preconditioner(update_module, graph, node.args[2],
                                        lambda n: _tensor(env, n) is r)
    return None


def gmres_preconditioner(updatex_module, graph: fx.Graph, x: esr.Tensor):
    """M of GMRES's UpdateX (gmres.py:133-136: x.add_(M(dx)), dx the
    squeezed product of the basis and y)."""
    env = _attrs(updatex_module, graph)
    for node in graph.nodes:
        if node.op == 'call_method' and node.target == 'add_' and \
                _tensor(env, node.args[0]) is x and len(node.args) == 2:
            return match_preconditioner(
                updatex_module, graph, node.args[1],
                lambda n: isinstance(n, fx.Node) and n.op == 'call_method' and
                n.target == 'squeeze')
    return None


def graph_of(prog, module):
    for m, g in zip(prog.modules, prog.graphs):
        if m is module:
            return g
    return None
