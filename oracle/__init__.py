"""CPU FP64 oracle for the SEM forward/adjoint path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or as the timed CPU baseline -- never as the product path.  The product
path lives in ``mun_amd`` and runs on the GPU through the HIP C-ABI library.
"""
from .sem_oracle import *  # noqa: F401,F403
