"""Fast spectral inelastic collision operator (drop-in for reference src/collision)."""
from .spectral_mesh import SpectralMesh  # noqa: F401
from .inelastic_collisions import FSInelasticVHSCollision  # noqa: F401
