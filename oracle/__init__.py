"""The oracle package (test infrastructure only): see oracle/oracle.py."""
from .oracle import *  # noqa: F401,F403
from .oracle import build, lib, ref, ref_available  # noqa: F401
