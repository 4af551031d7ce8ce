"""The reference's box patch test on the oracle: the one-step tension with prescribed
displacement, solved by the oracle's Newton, gives the reference's value of the volume-averaged
PK2 S_xx (4.6, verification_tests/output/stress_tension_disp.out) for both the linear- and the
quadratic-tet box, within 1e-12 (the reference checks its own run at 1e-14)."""
import numpy as np
import pytest

import oracle as O
import patch_deck
from problems import oracle_newton_step


@pytest.mark.parametrize("lt", ["LT", "QT"])
def test_patch_tension_average_stress(lt):
    prob = patch_deck.load(lt)
    assert prob.ne == 24 and np.array_equal(prob.deck.defl, [1.0])
    st = O.OracleState(prob.nqp, False)
    defl = prob.deck.defl.copy()
    r, it, _ = oracle_newton_step(prob, st, np.zeros(prob.ndof), defl, dict(NORM=0.0, first=True), tol=1e-14,
                                  max_iter=40, computer_zero=1e-13)
    S, E, _, _ = O.output_variables(prob.mesh, prob.mats, st.F_np1)
    sxx = patch_deck.average_S_xx(prob, S)
    assert abs(sxx - patch_deck.expected()) < 1e-12, (sxx, it)
    # the solution is homogeneous: F = diag(2, 1, 1) at every QP
    F = st.F_np1.reshape(3, 3, -1)
    assert np.abs(F - np.diag([2.0, 1.0, 1.0])[:, :, None]).max() < 1e-12
