"""MI355X (gfx950) rigid-body step: grid broadphase, SAT narrowphase, batched
contact solver and integration behind a C ABI (include/mi_rb.h)."""
from . import _lib  # noqa: F401
from .world import World, CLPhysicsDemo, reference_config  # noqa: F401
