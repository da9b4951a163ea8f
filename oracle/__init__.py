"""CPU oracle for the smart-interpolation hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check (or, for the reference
walker, to time) the product.  ``biomedisa_amd`` never imports it.
"""
from .orc import *  # noqa: F401,F403
