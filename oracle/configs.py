"""Run configurations (re-exported from the package for the oracle's build recipe)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from raishinsource_amd.configs import *  # noqa: E402,F401,F403
from raishinsource_amd.configs import ALL, Pram  # noqa: E402,F401
