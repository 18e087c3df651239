"""Viscosity, mirroring the reference's ``Viscosity`` (explicit, viscosity.py:111-164) and
``Weiler18`` (implicit, weiler18.py:343-401).

The explicit kernel of the reference updates v in place while neighbours read it, so its result
depends on execution order; here every particle reads the pre-step velocities (Jacobi order).
"""


class Viscosity:
    def __init__(self, simulation_data, neighbour_search, config):
        self.ps = simulation_data
        self.ns = neighbour_search
        self.cfg = config

    def reset(self):
        pass

    def solve(self, coeff: float, dt: float):
        if coeff == 0.0:
            return  # the force is proportional to the coefficient
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        ps.ctx.viscosity_explicit(coeff, dt)
        ps.sync_from_library(v=True)


class Weiler18(Viscosity):
    """Block-Jacobi PCG on (I - dt/rho L) v_new = v_old, tol |r|^2 < 1e-5, cap 1000."""

    def __init__(self, simulation_data, neighbour_search, config):
        super().__init__(simulation_data, neighbour_search, config)
        self.max_cg_iterations = 1000
        self.tol = 1e-5
        self.use_preconditioner = True
        self.warm_start = False
        self.last_cg_iters = 0

    def solve(self, coeff: float, dt: float):
        if self.warm_start or not self.use_preconditioner:
            raise NotImplementedError("Weiler18: only the default cold-start, block-Jacobi solve is implemented")
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        self.last_cg_iters = ps.ctx.viscosity_weiler18(coeff, dt)
        ps.sync_from_library(v=True)
