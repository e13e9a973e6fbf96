"""History output in the reference's .hst text format (src/outputs/history.cpp:47-201).

The file `<basename>.<file_id>.hst` starts with the descriptor line and the column header;
every output appends one row: time and dt in the data format (default " %12.5e", outputs.cpp:
254-255), cycle and nbtotal as " %12d", then the package's history variables (sum operations
first, then max, then min). Rows written here and rows of a reference run are compared by
`benchmarks/burgers/burgers_diff.py` semantics: relative difference 2|a-b| / (a+b+1e-20).
"""
import os

import numpy as np

from . import burgers_package


class HistoryOutput:
    def __init__(self, basename="burgers", file_id="out0", data_format="%12.5e"):
        self.fname = f"{basename}.{file_id}.hst"
        self.fmt = " " + data_format
        self.file_number = 0

    def write(self, md, time, dt, cycle, nbtotal, labels=None, values=None):
        """Append one row; by default the Burgers package's octant masses of md."""
        if values is None:
            values = burgers_package.MassHistory(md)
            labels = [f"MS Mass {q}" for q in range(8)]
        mode = "a" if self.file_number > 0 else "w"
        with open(self.fname, mode) as f:
            if self.file_number == 0:
                f.write("#  History data\n")
                cols = ["# [1]=time    ", " [2]=dt      ", " [3]=cycle   ", " [4]=nbtotal "]
                cols += [" [%d]=%-8s" % (5 + q, lab) for q, lab in enumerate(labels)]
                f.write("".join(cols) + "\n")
            row = (self.fmt % time) + (self.fmt % dt) + (" %12d" % cycle) + (" %12d" % nbtotal)
            row += "".join(self.fmt % v for v in values)
            f.write(row + "\n")
        self.file_number += 1
        return values


def rel_diff(a, b):
    """burgers_diff.py's relative difference of two history tables."""
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return 2 * np.abs(a - b) / (a + b + 1e-20)


def load(fname):
    return np.loadtxt(fname) if os.path.exists(fname) else None
