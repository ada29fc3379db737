"""Python access to the MI355X ELP move engine.

Thin ctypes bindings over the C ABI in ``include/elp_gpu.h`` (``libelp.so``)
and over the handle API of the drop-in host libraries
(``libgpupol3_ring.so``, ``libgpupol_lin.so``).  The libraries are built
in-tree by ``make -C mc-polymer-simulation_amd``; loading fails loudly when
they are missing -- there is no CPU fallback in this package.
"""
from .core import (  # noqa: F401
    PKG_DIR,
    LIB_DIR,
    SUBLAT_CUBIC,
    SUBLAT_DENSE,
    RULES_REFERENCE,
    RULES_BALANCED,
    ElpError,
    Schedule,
    autotune,
    Net,
    patch_moves,
    patch_features,
    tile_check,
    Batch,
    Engine,
    RingHost,
    LinHost,
    bench_copy,
    device_count,
    load_elp,
    exported_symbols,
    declared_symbols,
)
