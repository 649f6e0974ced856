"""``python -m tailor_amd.tailor -c config.yaml [options]``: the solver executable."""
import sys

from .cli import solver_main

if __name__ == "__main__":
    sys.exit(solver_main())
