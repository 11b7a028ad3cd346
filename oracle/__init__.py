"""CPU oracle: a numpy FP64 restatement of py-cubed-sphere's advection step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker or as the timed CPU baseline.  The product (the HIP library and its
host code) never imports it.
"""
