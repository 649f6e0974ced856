#!/usr/bin/env python3
"""Top-level launcher: FD-q cache check, then the GPU solver (``-c config.yaml``)."""
import sys

from tailor_amd.cli import main

if __name__ == "__main__":
    sys.exit(main())
