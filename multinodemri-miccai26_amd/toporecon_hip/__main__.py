"""``python -m toporecon_hip reconstruct --algorithm tvme -- ARGS``."""

import sys

from .mains import cli_main

if __name__ == "__main__":
    sys.exit(cli_main())
