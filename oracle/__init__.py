"""CPU oracle of the Navier-Stokes step: test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this package; the product (``aerolume_amd``) never does.
"""
