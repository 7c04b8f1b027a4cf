"""MI355X (gfx950) particle substep of Lustrine: position-based fluids and
sand contact/friction, as hand-written HIP kernels behind a C ABI
(include/lustrine_gpu.h, lustrine_amd/liblustrine_gpu.so)."""
from . import scenes  # noqa: F401

__all__ = ["scenes", "gpu"]
