"""Key layout of tests/golden/ntt_golden.npz (shared by the generator and the tests)."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from gen_golden import LDE, LOGS, case_key  # noqa: E402,F401
