"""Loads the package's ctypes binding (claude-vibe-code-raytracing_amd/lbm.py)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "claude-vibe-code-raytracing_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import lbm  # noqa: E402
import scenegen  # noqa: E402

HEADER = os.path.join(ROOT, "include", "lbm.h")
