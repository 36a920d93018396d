"""Build step for the plbpd HIP library and C++ drop-in host library.

``python -m plbpd_amd.build`` (or __graft_entry__.build()) runs
``make -B -C plbpd_amd`` so that every source is recompiled for gfx950.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = True) -> None:
    cmd = ["make", f"-j{jobs}", "-C", ROOT]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv)
