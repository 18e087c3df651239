"""Neighbour search, mirroring the reference's ``NeighbourSearch``
(scripts/sim/neighbour_search.py:81-179): ``build(positions, n)`` then
``build_cache(positions, n, cache_size, count, ids, gradW)``.

The HashGrid is replaced by the library's uniform grid and counting sort (ivp_neighbors_build),
which also evaluates density and Aii in the same pass; ``PressureBase.compute_density`` and
``compute_Aii`` read those results back.  One behaviour differs on purpose: a particle with more
than ``cache_size`` neighbours raises ``RuntimeError`` instead of being truncated silently
(neighbour_search.py:45), because truncated rows make the neighbour relation asymmetric.
"""

import math

from . import _native as N
from .simulation_data import owner_of


class NeighbourSearch:
    def __init__(self, domain_size: tuple, support_radius: float):
        self.support_radius = support_radius
        # grid dimensions as the reference reports them (neighbour_search.py:105-107); the library
        # grid uses cells of h (1 + 2^-10) with a two-cell margin, see include/ivp.h
        self.grid_dims = tuple(max(1, int(math.ceil(d / support_radius))) for d in domain_size)
        self.last_status = 0

    def build(self, positions, num_particles: int):
        ps = owner_of(positions)
        if int(num_particles) != ps.particle_num:
            raise ValueError("num_particles must equal SimulationData.particle_num")
        ps.sync_to_library(positions=True)
        status = ps.ctx.neighbors_build()
        self.last_status = status
        ps._cache_stale = True
        if status == N.IVP_NEIGHBOR_OVERFLOW:
            raise RuntimeError(f"[NeighbourSearch] a particle has more than neighborCacheSize={ps.cache_size} "
                               f"neighbours; raise neighborCacheSize")

    def build_cache(self, positions, num_particles: int, cache_size: int, neighbor_count=None, neighbor_ids=None,
                    neighbor_gradW=None):
        """Rows are built by ``build``.  Caller-owned arrays are filled in the reference layout;
        SimulationData's own cache attributes are exported when read."""
        ps = owner_of(positions)
        if int(cache_size) != ps.cache_size:
            raise ValueError("cache_size must equal SimulationData.cache_size")
        own = ps._cache
        if neighbor_count is None or (own is not None and neighbor_count.data_ptr() == own[0].data_ptr()):
            return
        ps.ctx.neighbors_export(neighbor_count, neighbor_ids, neighbor_gradW.reshape(-1), cache_size=cache_size)
