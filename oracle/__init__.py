"""The oracle: a CPU restatement of HCGrid's gridding path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()``
and bench.py's ``cpu_baseline`` leg as the checker; the product path never
imports it.  See ``oracle/hcg_oracle.h`` for what each function restates and
how it is pinned.
"""
from .binding import *  # noqa: F401,F403
