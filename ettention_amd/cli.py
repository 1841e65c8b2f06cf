"""Command-line reconstruction, as the reference's `ettention` tool:

    python -m ettention_amd.cli --config reconstruction.xml [--key value ...]

The same program is built as the native executable ettention_amd/etn.
"""
from __future__ import annotations

import sys

from . import io


def run(argv=None):
    return io.main(sys.argv[1:] if argv is None else argv)


if __name__ == "__main__":
    sys.exit(run())
