"""CPU oracle of the Sheen render path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, as the checker (or the timed CPU baseline); the product path
never calls it.  See sheen_oracle.cpp for what each function restates.
"""
from .binding import Oracle, build, lib  # noqa: F401
