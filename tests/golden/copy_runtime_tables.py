"""Copies two of the reference's published runtime tables (data files under
plotting/data/runtime/, written by RuntimeRecords::csvDump) into
tests/golden/runtime/ unchanged.  They are the known answers for the column
layout, number format and aggregation rules of trax_amd.runtime_record; run
once where the reference is present."""
import os
import shutil
import sys

SRC = "/root/reference/plotting/data/runtime"
FILES = ("runtime.bigOne.halfLocal.conf.tracks.test.cpu.csv", "runtime.bigOne.local.conf.events.100.test.gpu.csv")


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "runtime")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        src = os.path.join(SRC, f)
        if not os.path.exists(src):
            sys.exit(f"{src} not found")
        shutil.copyfile(src, os.path.join(out, f))
        print(f"copied {f}")


if __name__ == "__main__":
    main()
