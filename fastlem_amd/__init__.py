"""fastlem_amd -- MI355X (gfx950) uplift-erosion solve with fastlem's API.

The public surface mirrors the reference crate's paths used by
``generate()``: ``TopographicalParameters`` (src/core/parameters.rs),
``TerrainModel2DBulider`` / ``TerrainModel2D`` / ``Terrain2D`` / ``Site2D``
(src/models/surface/), ``TerrainGenerator`` and ``GenerationError``
(src/lem/generator.rs).  The iteration loop runs as HIP kernels through the
C ABI declared in include/fastlem_hip.h.
"""
from .generator import (GenerationError, InvalidNumberOfParameters, ModelNotSet, ParameterArrays,
                        ParametersNotSet, TerrainGenerator, jitter_elevations)
from .parameters import TopographicalParameters
from .surface import Site2D, Terrain2D, TerrainModel2D, TerrainModel2DBulider

__all__ = [
    "GenerationError", "InvalidNumberOfParameters", "ModelNotSet", "ParametersNotSet", "ParameterArrays",
    "TerrainGenerator", "jitter_elevations", "TopographicalParameters", "Site2D", "Terrain2D", "TerrainModel2D",
    "TerrainModel2DBulider",
]
