"""CPU oracle: a restatement of the reference (jack-pan-ai/EASIER) algorithms
on the path the hip backend replaces.

TEST INFRASTRUCTURE ONLY. Nothing in easier_amd imports, links or runs this
package; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use it, always as the checker (or the timed CPU baseline), never as the thing
measured or shipped.

The reference's CPU backend (esr.compile(..., backend='cpu')) evaluates each
program node with torch CPU operators; the oracle calls the same operators
(t[idx], zeros().scatter_reduce_(..., include_self=False), torch.sum/norm)
in the same order, without the FX interpreter.

Modules and the reference lines each follows:
  ops.py      Selector/Reducer/aggregators  easier/core/module.py:364-371,
              441-461, 825-884; dataflow_distribution.py:197-289
  mesh.py     triangular mesh generators    csrc/triangular_mesh.cpp:10-142
  solvers.py  Linsys, CG, GMRES             numeric/linsys.py:23-24,
              numeric/solver/cg.py:12-137, numeric/solver/gmres.py:12-266
  swe.py      shallow-water geometry + RK4  tutorial/shallow_water_equation/
              assemble_shallow_water.py:154-253, swe_main.py:86-141
  fem.py      P1 finite elements (no reference counterpart: the configs need
              P1-FEM Poisson; pinned by closed forms, see its docstring)
  renumber.py index renumbering at world size 1
              passes/sparse_encoding/sparse_encoding.py:244-338,
              passes/utils.py:301-336

Parity pinning of every function is listed in tests/test_oracle.py.
"""
