"""pactae_amd — MI355X-native BambooSMP proposal engine.

The engine is a HIP/C++ library (pactae_amd/_lib/libsmp_hip.so) with a C ABI
(include/smp_hip.h).  ``pactae_amd.smp`` is the Python mirror of the reference's
engine interface over that ABI; importing it loads the library and raises if it
is not built.  ``pactae_amd.build.build()`` compiles it for gfx950.
"""

__all__ = ["build", "smp"]


def __getattr__(name):
    # Import the submodule by its full name: ``from . import smp`` would look the
    # attribute up on this package first and land here again.  The import binds
    # the submodule on the package, so this runs once per name.
    if name in ("smp", "build"):
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
