"""Python entry points of the integral path, one module per integral as in the reference
(reference molecular_integral/{overlap,nuclear_attraction,two_center,three_center}.py)."""

from .overlap import overlap
from .nuclear_attraction import nuclear_attraction
from .two_center import two_center
from .three_center import three_center
