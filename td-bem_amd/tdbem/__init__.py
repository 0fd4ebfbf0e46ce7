"""MI355X-native TD-BEM hot paths: Python face of the C ABI (include/tdbem.h) plus the workload
generators.  The compute runs in libtdbem.so (HIP, gfx950); nothing here computes G/H or marches."""
from . import api, configs, deck, meshgen  # noqa: F401

__all__ = ["api", "configs", "deck", "meshgen"]
