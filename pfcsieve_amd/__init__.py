"""MI355X-native (gfx950 HIP) implementation of PFCSieve's GPU sieve.

The compute path is ``libpfcsieve.so`` (hand-written HIP kernels behind the C
ABI of ``include/pfc.h``); this package is the Python host binding used by
tests, the benchmark and the serving loop.
"""
from ._native import PfcError, lib  # noqa: F401
from .sieve import (COMPOSITORIAL, FACTORIAL, MODE_COMPOSITORIAL, MODE_FACTORIAL,  # noqa: F401
                    MODE_PRIMORIAL, PRIMORIAL, Batcher, Device, Job, SieveResult, TableLayout, footprint,
                    job_result, run_batch, run_sieve, setup_table_cost, setup_table_size)
