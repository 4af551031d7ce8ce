"""Seeded synthetic problems shared by the tests (box meshes of the unit cube)."""
import numpy as np

import oracle as O
import pgfem3d_amd as P


class BoxProblem:
    def __init__(self, n, etype, material, jitter=0.1, seed=1):
        self.mesh = P.mesh_box(n, etype, jitter, seed)
        self.supports = P.box_supports(self.mesh)
        self.ids, self.ndof = P.generate_dof_ids(self.mesh, self.supports)
        self.Ap, self.Ai = P.psparse_apai(self.mesh, self.ids, self.ndof)
        self.mats = [material]
        self.nqp = self.mesh.n_elem * O.nint(etype)

    def displacement(self, stretch, noise, seed):
        """Free-DOF vector of a uniaxial stretch u_z = stretch*z plus lateral contraction and
        seeded noise, and the prescribed top value."""
        rng = np.random.default_rng(seed)
        X = self.mesh.X
        u = np.zeros_like(X)
        u[:, 2] = stretch * X[:, 2]
        u[:, 0] = -0.3 * stretch * (X[:, 0] - 0.5) * X[:, 2]
        u[:, 1] = -0.3 * stretch * (X[:, 1] - 0.5) * X[:, 2]
        u += noise * rng.standard_normal(u.shape) * X[:, 2:3] * (1 - X[:, 2:3])
        r = np.zeros(self.ndof)
        mask = self.ids > 0
        r[self.ids[mask] - 1] = u[mask]
        return r, np.array([stretch])
