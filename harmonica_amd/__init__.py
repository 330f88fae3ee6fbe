"""
harmonica_amd -- MI355X (gfx950) implementation of harmonica's pairwise
gravity forward-modelling path.

Drop-in functions with the reference's Python API (fatiando/harmonica):

* ``prism_gravity`` (all ten fields, singular-point handling),
* ``prism_layer`` and its ``.prism_layer`` accessor (``.gravity``),
* ``point_gravity`` (Cartesian and spherical),
* ``EquivalentSources`` (jacobian / predict / fit / grid / profile),
  ``EquivalentSourcesGB`` (gradient-boosted windowed fit) and
  ``EquivalentSourcesSph`` (spherical coordinates),
* ``prism_magnetic`` (b_e, b_n, b_u or the vector ``b``),
* ``tesseroid_gravity`` (potential and g_z, adaptive discretisation) and
  ``tesseroid_layer`` with its ``.tesseroid_layer`` accessor,

computed by hand-written HIP kernels in libhmx.so (C-ABI in include/hmx.h),
called through ctypes, with PyTorch-ROCm used only for device buffers and
streams.  There is no CPU fallback: without the library or a GPU the
functions raise.
"""

from ._eqs import EquivalentSources
from ._eqs_gb import EquivalentSourcesGB
from ._eqs_sph import EquivalentSourcesSph
from ._forward import point_gravity, prism_gravity
from ._layer import DatasetAccessorPrismLayer, prism_layer
from ._magnetic import prism_magnetic
from ._tesseroid import tesseroid_gravity
from ._tesseroid_layer import DatasetAccessorTesseroidLayer, tesseroid_layer
from .constants import GRAVITATIONAL_CONST

__all__ = [
    "GRAVITATIONAL_CONST",
    "DatasetAccessorPrismLayer",
    "EquivalentSources",
    "EquivalentSourcesGB",
    "EquivalentSourcesSph",
    "point_gravity",
    "prism_gravity",
    "prism_layer",
    "prism_magnetic",
    "tesseroid_gravity",
    "tesseroid_layer",
    "DatasetAccessorTesseroidLayer",
]

__version__ = "0.1.0"
