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
    """Block-Jacobi PCG on (I - dt/rho L) v_new = v_old, tol |r|^2 < 1e-5, cap 1000.

    ``use_preconditioner`` and ``warm_start`` are the reference's mutable attributes
    (weiler18.py:210-211; its GUI toggles them live): plain CG without the preconditioner, and
    a PCG that starts from the previous solve's correction, kept per particle by the library
    and zeroed by Framework.reset (weiler18.py:213-220)."""

    def __init__(self, simulation_data, neighbour_search, config):
        super().__init__(simulation_data, neighbour_search, config)
        self.max_cg_iterations = 1000
        self.tol = 1e-5
        self.use_preconditioner = True
        self.warm_start = False
        self.last_cg_iters = 0

    def apply_options(self, ctx):
        """Hand the mutable attributes to the library (read at every Weiler18 solve)."""
        ctx.w18_warm_start = self.warm_start
        ctx.w18_preconditioner = self.use_preconditioner

    def reset(self):
        """The kept correction is zeroed by the library's ivp_reset, which Framework.reset calls
        (the reference's Framework.reset calls this method for the same purpose)."""

    def solve(self, coeff: float, dt: float):
        ps = self.ps
        ps.sync_to_library(positions=False, velocities=True)
        self.apply_options(ps.ctx)
        self.last_cg_iters = ps.ctx.viscosity_weiler18(coeff, dt)
        ps.sync_from_library(v=True)
