"""MI355X (gfx950) implementation of BookLeaf's Lagrangian step and ALE remap
behind the reference's driver interface (src/packages/hydro, src/packages/ale)
and dataAPI field layout (src/dataAPI)."""
from . import abi, problem  # noqa: F401

__all__ = ["abi", "problem", "device"]
