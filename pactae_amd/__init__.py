"""pactae_amd — MI355X-native BambooSMP proposal engine.

The engine is a HIP/C++ library (pactae_amd/_lib/libsmp_hip.so) with a C ABI
(include/smp_hip.h).  ``pactae_amd.smp`` is the Python mirror of the reference's
engine interface over that ABI; importing it loads the library and raises if it
is not built.  ``pactae_amd.build.build()`` compiles it for gfx950.
"""

__all__ = ["build", "smp"]


def __getattr__(name):
    if name == "smp":
        from . import smp
        return smp
    if name == "build":
        from .build import build
        return build
    raise AttributeError(name)
