# SPDX-License-Identifier: BSD-3-Clause
"""``python -m terraingen_amd backends``: each backend with its device or why it is unavailable,
then the default, as ``terraingen backends`` prints them (app/cli.cpp:192-205)."""
import sys

from . import native


def main(argv: list[str]) -> int:
    if argv[1:] != ["backends"]:
        print("usage: python -m terraingen_amd backends", file=sys.stderr)
        return 2
    text = native.backends_report()
    sys.stdout.write(text)
    return 1 if "error: " in text else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
