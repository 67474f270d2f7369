"""python -m rust-compress_amd [OPTIONS] INPUT... : the command-line driver (cli.py)."""
import sys

from .cli import main

sys.exit(main())
