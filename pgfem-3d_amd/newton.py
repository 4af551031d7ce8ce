"""Global Newton driver over the device element path (the role of Newton_Raphson,
src/Newton_Raphson.cc:886-1430, for the total Lagrangian CM analysis on one GPU).

Per load step, as the reference takes it when given the step's prescribed increment
(load_step(r, defl_n, defl_d=increment)): the first iteration is the reference's predictor --
the tangent and the load vector of the prescribed increment f_defl (pgf_prescribed_load) at
(u_n, defl_n), K du = -f_defl (Newton_Raphson.cc:1700-1741, stiffmat_fd.cc:722-745) -- after
which every iteration runs the fused integration + residual + tangent pass
(pgf_residual_tangent) at defl_n + increment, forms R = RR - f_int (RR = 0: displacement
driven), checks ||R|| / NORM against the tolerance and solves K du = R.  NORM is the largest
norm of a step's first residual (the one after the predictor) seen so far, re-taken while it
is below the tolerance (fv->NORM, :1207-1216).  Called with the full prescribed values of step
n+1 instead (defl_d None), the step starts from them at u_n with no predictor.  Linear solves
run on the device with pgf_solve (Jacobi-preconditioned BiCGStab), or on the host with a sparse
direct solve.  A step that fails (positive INFO from the element pass, a diverging or
non-finite residual, the iteration cap) returns its status so the caller can subdivide and
pgf_reset_step, as the reference does.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import pgfem3d_amd as P


@dataclass
class StepResult:
    converged: bool
    iterations: int
    norms: list = field(default_factory=list)  # normalised residual norm per iteration
    info: int = 0                              # element-pass INFO of the failing iteration
    solver_iters: list = field(default_factory=list)
    cuts: int = 0                              # subdivisions used (increment())
    gammas: list = field(default_factory=list)  # step length taken per iteration
    alpha: float = 0.0                          # physics-based subdivision measure at convergence


ALPHA_RESTART = 1.25  # Newton_Raphson.cc:91


def substeps_after_failure(alpha: float) -> int:
    """Pieces a failed step is cut into at its first subdivision
    (SubdivisionScheme::do_subdivision, src/subdivision.cc:139-170): the step shrinks by
    1/alpha when the physics asks for more than 4/3 of it, else by 0.75; the count is the
    rounded ratio (round1, src/matice.cc:43-47), and at least one more than before."""
    gama = 1.0 / alpha if alpha > 4.0 / 3.0 else 0.75
    i = int(1.0 / gama + 0.5)
    return i if i > 1 else 2


def backtrack_gamma(gam, phi, prev=None):
    """Step length for the next trial of the line search (LINE.cc:99-124).  phi is
    ||R(u + gam du)||^2 / ||R(u)||^2 at the rejected trial gam (so phi(0) = 1 and the Newton
    direction's slope is -2); prev = (gam, phi) of the trial before, or None after the full
    step.  The first cut is the minimiser of the quadratic through phi(0), phi'(0) and phi(1);
    later cuts the minimiser of the cubic through phi(0), phi'(0) and the last two trials,
    capped at half the current step.  No cut goes below a tenth of the current step."""
    if prev is None:
        lam = 1.0 / (phi + 1.0)  # -phi'(0) / (2 (phi(1) - phi(0) - phi'(0)))
    else:
        g1, p1 = prev
        e0 = (phi - 1.0 + 2.0 * gam) / (gam * gam)  # excess over the linear model, / gam^2
        e1 = (p1 - 1.0 + 2.0 * g1) / (g1 * g1)
        a = (e0 - e1) / (gam - g1)  # phi(x) = 1 - 2 x + b x^2 + a x^3
        b = (gam * e1 - g1 * e0) / (gam - g1)
        if a == 0.0:
            lam = 1.0 / b
        else:
            disc = b * b + 6.0 * a
            if disc < 0.0:
                lam = 0.5 * gam
            elif b <= 0.0:
                lam = (np.sqrt(disc) - b) / (3.0 * a)
            else:
                lam = 2.0 / (b + np.sqrt(disc))  # same root, without cancellation
        lam = min(lam, 0.5 * gam)
    return max(lam, 0.1 * gam)


class NewtonDriver:
    def __init__(self, h: P.Handle, Ap, Ai, tol=1e-10, max_iter=25, solver="device", lin_tol=1e-13,
                 lin_maxit=20000, computer_zero=1e-15, line_search=True, iter_max=None,
                 alpha_restart=ALPHA_RESTART):
        """tol is the reference's nor_min; iter_max its iteration cap (None: max_iter only):
        past iter_max iterations a residual above 20 tol, or any past iter_max + 3, ends the
        step -- converged if below 5 tol, else failed (Newton_Raphson.cc:1333-1352).
        alpha_restart: a converged step whose physics-based subdivision measure
        (pgf_subdivision_param) exceeds it fails for subdivision (Newton_Raphson.cc:1364-1398);
        None skips the check."""
        import torch

        self.h, self.tol, self.max_iter = h, tol, max_iter
        self.solver, self.lin_tol, self.lin_maxit = solver, lin_tol, lin_maxit
        self.computer_zero, self.line_search, self.iter_max = computer_zero, line_search, iter_max
        self.alpha_restart = alpha_restart
        self.Ap, self.Ai = Ap, Ai
        dev = torch.device("cuda:0")
        self.f = torch.empty(h.ndof, dtype=torch.float64, device=dev)
        self.vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
        self.du = torch.zeros(h.ndof, dtype=torch.float64, device=dev)
        self.rt = torch.empty(h.ndof, dtype=torch.float64, device=dev)  # line-search trial point
        self.NORM = 0.0
        self.first = True
        self.last_relres = 0.0

    def _solve(self, R):
        """du = K^-1 R; ok is False when the solve failed the reference's check
        (Newton_Raphson.cc:996-1011): a non-finite residual norm, or one above 500 times the
        solver tolerance (BiCGStab breakdown or its iteration limit).  The step then fails
        with INFO = 1 and is subdivided."""
        import torch

        if self.solver == "device":
            self.du.zero_()
            it, rr = self.h.solve(self.vals, R, self.du, tol=self.lin_tol, maxit=self.lin_maxit)
            self.last_relres = rr
            return self.du, it, bool(np.isfinite(rr) and rr <= 500.0 * self.lin_tol)
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla

        K = sp.csr_matrix((self.vals.cpu().numpy(), self.Ai, self.Ap), shape=(self.h.ndof, self.h.ndof))
        Rh = R.cpu().numpy()
        x = spla.spsolve(K.tocsc(), Rh)
        nb = np.linalg.norm(Rh)
        rr = float(np.linalg.norm(K @ x - Rh) / nb) if nb > 0 else 0.0
        self.last_relres = rr
        ok = bool(np.all(np.isfinite(x)) and np.isfinite(rr) and rr <= 500.0 * self.lin_tol)
        return torch.from_numpy(np.ascontiguousarray(x)).to(R.device), 0, ok

    def _accept(self, res, dt):
        """A converged step: commit it unless its physics-based subdivision measure asks for
        a smaller step (INFO = 1, Newton_Raphson.cc:1364-1398)."""
        if self.alpha_restart is not None:
            res.alpha = self.h.subdivision_param(dt)
            if not res.alpha <= self.alpha_restart:
                res.info = 1
                return
        res.converged = True
        self.h.commit_step()

    def increment(self, r, defl_from, defl_to, dt=1.0, max_cuts=4, predictor=True) -> StepResult:
        """Take the prescribed values from defl_from (converged) to defl_to, subdividing on
        failure as the reference does (a failing step is reset with pgf_reset_step and the
        remaining increment is applied in halves, up to max_cuts levels; the time step is
        split with it).  A step rejected by the physics-based measure alpha > 4/3 is cut into
        round(alpha) pieces instead of halves (substeps_after_failure).  predictor: each piece
        starts with the reference's predictor."""
        out = StepResult(False, 0)
        pieces = [(0.0, 1.0, 0)]
        r_conv = r.clone()
        while pieces:
            a, b, lvl = pieces.pop(0)
            if predictor:
                res = self.load_step(r, defl_from + a * (defl_to - defl_from), dt=dt * (b - a),
                                     defl_d=(b - a) * (defl_to - defl_from))
            else:
                res = self.load_step(r, defl_from + b * (defl_to - defl_from), dt=dt * (b - a))
            out.iterations += res.iterations
            out.norms += res.norms
            if res.converged:
                r_conv.copy_(r)
                continue
            self.h.reset_step()
            r.copy_(r_conv)
            if lvl >= max_cuts:
                out.info = res.info
                return out
            out.cuts += 1
            k = substeps_after_failure(res.alpha)
            pieces = [(a + (b - a) * j / k, a + (b - a) * (j + 1) / k, lvl + 1) for j in range(k)] + pieces
        out.converged = True
        return out

    def load_step(self, r, defl, d_r=None, defl_d=None, dt=1.0) -> StepResult:
        """Solve one load step in place: r (free DOFs, device float64) enters as the converged
        u(n) and leaves as u(n+1).  With defl_d (device, the step's prescribed increment) defl
        holds the prescribed values of step n and the step opens with the predictor; without
        it defl holds those of step n+1."""
        import torch

        res = StepResult(False, 0)
        it0 = 0
        tf = getattr(self.h, "three_field", False)
        if defl_d is not None:
            # predictor: K and f_defl at (u_n, defl_n), K du = -f_defl, no line search
            info = self.h.tangent(self.vals, r, d_r=d_r, defl=defl, dt=dt)
            if not info:
                info = self.h.prescribed_load(self.f, r, defl_d, d_r=d_r, defl=defl, dt=dt)
            if info:
                res.info = info
                return res
            du, lit, ok = self._solve(-self.f)
            res.solver_iters.append((lit, self.last_relres))
            if not ok:
                res.info = 1
                return res
            r.add_(du)
            res.gammas.append(1.0)
            it0 = 1

        def run(x):
            return self.h.residual_tangent(self.f, self.vals, x, d_r=d_r, defl=defl, defl_d=defl_d, dt=dt)

        info = run(r)
        for it in range(it0, self.max_iter):
            if info:
                res.info = info
                return res
            R = -self.f
            nor2 = float(R.norm())
            # NORM from a step's first residual (Newton_Raphson.cc:1207-1216)
            if it == it0 and (self.first or self.NORM < self.tol):
                self.NORM = max(self.NORM, nor2, self.computer_zero)
                self.first = False
            nor = nor2 / self.NORM
            res.norms.append(nor)
            res.iterations = it
            if not np.isfinite(nor):
                return res
            # iteration cap: this residual follows the reference's iteration it - 1
            if (self.iter_max is not None and it > self.iter_max and nor > self.tol
                    and (nor > 20.0 * self.tol or it > self.iter_max + 3)):
                if nor < 5.0 * self.tol:
                    self._accept(res, dt)
                return res
            if nor <= self.tol or nor2 <= self.computer_zero:
                self._accept(res, dt)
                return res
            du, lit, ok = self._solve(R)
            res.solver_iters.append((lit, self.last_relres))
            if not ok:
                res.info = 1
                return res
            if tf:  # the elements' volume and pressure increments that go with du
                self.h.tf_update(du)
            gam, prev = 1.0, None
            while True:
                torch.add(r, du, alpha=gam, out=self.rt)
                if tf:
                    self.h.tf_accept(gam)
                info = run(self.rt)
                if info or it == 0 or not self.line_search:
                    break
                phi = float(self.f.norm()) ** 2 / (nor2 * nor2)
                if not np.isfinite(phi):
                    return res
                if phi <= 1.0 - 2e-4 * gam:  # sufficient decrease (LINE.cc:96)
                    break
                if tf:
                    self.h.tf_accept(0.0)  # back to the start of the trials
                gam, prev = backtrack_gamma(gam, phi, prev), (gam, phi)
                if gam < 0.1 * self.tol:  # the line search failed (LINE.cc:126-127)
                    if nor < 5.0 * self.tol:  # soft convergence: keep u (Newton_Raphson.cc:1246)
                        if run(r) == 0:
                            self._accept(res, dt)
                    return res
            res.gammas.append(gam)
            r.copy_(self.rt)
        return res
