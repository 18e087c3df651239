"""Scene configuration, mirroring the reference's ``SimConfig``
(scripts/io/config_builder.py:10-115).

Typed keys are parsed from the scene JSON ``Configuration`` block exactly as the reference
does (``from_json``, config_builder.py:54-100); every other key is read lazily through
``get_cfg(name, default)`` (config_builder.py:102-115), which is how the IP solver reads
``pressureMethod``, ``densityErrorTol``, ``maxOuterIters``, ``fixedOuterIters``, ``omega``,
``tol_pcg`` and ``neighborCacheSize``.
"""

import json
from dataclasses import dataclass, field
from typing import Optional


@dataclass
class SimConfig:
    domain_size: tuple = (1.0, 1.0, 1.0)
    domain_start: tuple = (0.0, 0.0, 0.0)
    particle_radius: float = 0.01
    support_radius: float = 0.04
    particle_diameter: float = 0.02
    density0: float = 1000.0
    dt: float = 0.0016
    gravity: tuple = (0.0, -9.8, 0.0)
    viscosity_method: int = 0
    surface_tension_method: int = 0
    viscosity_coeff: float = 0.01
    surface_tension_coeff: float = 0.01
    scene_data: Optional[dict] = field(default=None)

    def __post_init__(self):
        # config_builder.py:46-52: d = 2r, h = 4r
        self.particle_diameter = 2.0 * self.particle_radius
        self.support_radius = 4.0 * self.particle_radius

    @classmethod
    def from_dict(cls, scene_data: dict) -> "SimConfig":
        """Same parsing as ``SimConfig.from_json`` (config_builder.py:54-100) on a loaded dict."""
        cfg = scene_data.get("Configuration", {})
        domain_start = tuple(cfg.get("domainStart", [0.0, 0.0, 0.0]))
        domain_end = tuple(cfg.get("domainEnd", [1.0, 1.0, 1.0]))
        domain_size = tuple(e - s for s, e in zip(domain_start, domain_end))
        return cls(
            domain_size=domain_size,
            domain_start=domain_start,
            particle_radius=float(cfg.get("particleRadius", 0.01)),
            dt=float(cfg.get("timeStepSize", 0.0016)),
            density0=float(cfg.get("density0", 1000.0)),
            gravity=tuple(cfg.get("gravitation", [0.0, -9.8, 0.0])),
            viscosity_method=int(cfg.get("viscosityMethod", 0)),
            surface_tension_method=int(cfg.get("surfaceTensionMethod", 0)),
            viscosity_coeff=float(cfg.get("viscosityCoeff", 0.01)),
            surface_tension_coeff=float(cfg.get("surfaceTensionCoeff", 0.01)),
            scene_data=scene_data,
        )

    @classmethod
    def from_json(cls, scene_file_path: str) -> "SimConfig":
        with open(scene_file_path, "r") as f:
            return cls.from_dict(json.load(f))

    def get_cfg(self, name: str, default=None):
        if self.scene_data is None:
            return default
        return self.scene_data.get("Configuration", {}).get(name, default)

    def get_fluid_blocks(self):
        return [] if self.scene_data is None else self.scene_data.get("FluidBlocks", [])

    def get_generated_bodies(self):
        return [] if self.scene_data is None else self.scene_data.get("GeneratedBodies", [])

    def get_rigid_bodies(self):
        return [] if self.scene_data is None else self.scene_data.get("RigidBodies", [])

    def get_fluid_bodies(self):
        return [] if self.scene_data is None else self.scene_data.get("FluidBodies", [])

    def get_solid_bodies(self):
        return [] if self.scene_data is None else self.scene_data.get("SolidBodies", [])

    def get_scheduled_spawns(self):
        return [] if self.scene_data is None else self.scene_data.get("ScheduledSpawns", [])
