"""The Poisson solves at the benchmark's size: the 2048 x 2048 2-triangle mesh
with a variable coefficient (4,198,401 vertices, 25,174,016 couplings), f64,
relative tolerance 1e-8. The oracle's CG would need minutes on the host, so
the checks are size-independent: the true residual ||b - A x|| recomputed on
the host with scipy, and two independent methods (fused Jacobi-PCG, the
three-kernel form at this size, and multigrid flexible CG with damped-Jacobi
or the trained network as smoother, on the geometric hierarchy and on the
smoothed-aggregation one) agreeing on x."""
import os

import pytest

from easier_amd.numeric.preconditioner import learned as lp
from easier_amd.workloads import learned as wl

WEIGHTS = os.path.join(os.path.dirname(wl.__file__), 'gnn_smoother_l3_h32.npz')

pytestmark = pytest.mark.gpu


def test_full_size_jacobi_pcg_and_multigrid_fcg_agree(gpu):
    print('full-size Poisson 2048x2048: Jacobi-PCG and multigrid FCG', flush=True)
    res = wl.evaluate_compiled(lp.load(WEIGHTS), 2048, 1, rtol=1e-8, maxiter=20000, reps=1,
                               methods=('jacobi', 'mg', 'mg-learned', 'amg'))
    jac, mg, ml, amg = res['jacobi'], res['mg'], res['mg-learned'], res['amg']
    assert ml['true_rel_residual'] <= 1e-8 * (1 + 1e-6), ml
    assert ml['rel_diff_x'] <= 1e-6 and ml['iters'] <= 20 and ml['fallback_at'] is None, ml
    assert res['vertices'] == 2049 * 2049
    assert jac['true_rel_residual'] <= 1e-8 * (1 + 1e-6), jac
    assert mg['true_rel_residual'] <= 1e-8 * (1 + 1e-6), mg
    # both stop at 1e-8 relative residual; on this problem (condition
    # number ~1e7) that leaves x within ~1e-6 of each other
    assert mg['rel_diff_x'] <= 1e-6, mg
    assert 5000 <= jac['iters'] <= 12000 and mg['iters'] <= 20 and mg['fallback_at'] is None
    assert amg['true_rel_residual'] <= 1e-8 * (1 + 1e-6), amg
    assert amg['rel_diff_x'] <= 1e-6 and amg['iters'] <= 30 and amg['fallback_at'] is None, amg
