"""bluetruthengine_amd — MI355X-native wavefront path tracer.

The product is the HIP library ``libbte.so`` built from ``csrc/`` and exposed
through the C ABI declared in ``include/bte_api.h``; this package holds the
Python mirror of that interface, the POD layouts and the procedural scene
generators used by tests and the benchmark.
"""
from . import pod  # noqa: F401
