"""Pins of the global Newton line search (LINE_S3_MP, src/LINE.cc:58-250) on the CPU.

The oracle's step-length rule is checked against closed forms: for a quadratic merit the
first cut is the exact minimiser, and for a cubic merit the second cut is the exact minimiser
of the cubic, in both root branches; the caps at half and a tenth of the current step hold.
The driver's own rule (pgfem-3d_amd/newton.py) must give the same steps.  Finally, on a
four-grain crystal-plasticity box, the oracle's full-step Newton cycles while the line
search converges."""
import numpy as np
import pytest

import oracle as O
import pgfem3d_amd as P
from newton import backtrack_gamma
from problems import CPBoxProblem, oracle_newton_step


def _cubic_min(a, b):
    """Positive minimiser of phi(x) = 1 - 2x + b x^2 + a x^3 (phi'(x) = 0, phi'' > 0)."""
    roots = np.roots([3.0 * a, 2.0 * b, -2.0])
    roots = roots[np.isreal(roots)].real
    return float(min(x for x in roots if x > 0 and 2 * b + 6 * a * x > 0))


@pytest.mark.parametrize("c", [2.5, 4.0, 9.0])
def test_first_cut_minimises_the_quadratic(c):
    phi = lambda x: 1.0 - 2.0 * x + c * x * x  # noqa: E731
    assert O.line_search_gamma(1.0, phi(1.0)) == pytest.approx(1.0 / c, rel=1e-14)


@pytest.mark.parametrize("a,b", [(5.0, 5.541666666666667), (34.07407407407407, -1.0), (12.0, 2.0)])
def test_second_cut_minimises_the_cubic(a, b):
    """The cubic through phi(0), phi'(0) and trials at 1 and 0.5 is phi itself, so the cut is
    its minimiser (chosen inside the caps [0.05, 0.25]; b > 0 and b <= 0 take the two root
    forms)."""
    phi = lambda x: 1.0 - 2.0 * x + b * x * x + a * x ** 3  # noqa: E731
    x = _cubic_min(a, b)
    assert 0.05 < x < 0.25
    assert O.line_search_gamma(0.5, phi(0.5), (1.0, phi(1.0))) == pytest.approx(x, rel=1e-12)


def test_caps():
    # a merit that rises steeply: the quadratic proposal is below a tenth of the step
    assert O.line_search_gamma(1.0, 1e6) == pytest.approx(0.1)
    # a cubic minimiser past half the current step is capped at half
    phi = lambda x: 1.0 - 2.0 * x + 0.9 * x * x + 0.2 * x ** 3  # noqa: E731
    assert _cubic_min(0.2, 0.9) > 0.25
    assert O.line_search_gamma(0.5, phi(0.5), (1.0, phi(1.0))) == pytest.approx(0.25, rel=1e-14)


def test_driver_rule_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        g = rng.uniform(0.01, 1.0)
        phi = rng.uniform(0.5, 50.0)
        prev = None if rng.random() < 0.3 else (g * rng.uniform(1.5, 10.0), rng.uniform(0.5, 50.0))
        if prev is None:
            g = 1.0
        assert backtrack_gamma(g, phi, prev) == pytest.approx(O.line_search_gamma(g, phi, prev), rel=1e-12)


def test_crystal_plasticity_needs_the_line_search():
    prob = CPBoxProblem(3, P.PGF_TET4, P.fcc_crystal(), n_grains=4, seed=2)
    guess, defl = prob.displacement(1e-3, 0.0, seed=0)
    tables = (prob.gP, prob.gL)
    st = O.CPState(prob.nqp, prob.mats[0].g0)
    with pytest.raises(AssertionError, match="did not converge"):
        oracle_newton_step(prob, st, guess, defl, dict(NORM=0.0, first=True), tol=1e-12, max_iter=15,
                           cp_tables=tables, line_search=False)
    st = O.CPState(prob.nqp, prob.mats[0].g0)
    _, it, gams = oracle_newton_step(prob, st, guess, defl, dict(NORM=0.0, first=True), tol=1e-12,
                                     cp_tables=tables)
    assert it <= 12 and min(gams) < 1.0, (it, gams)
