"""The reference's restart files on the Python side.

A restart file holds one list-directed record of the 14 primitive fields
(PRIM_NAMES order) per point, i outer over -6..imax+6 and j inner over
-6..jmax+6, ghosts included: lector writes it (22_writer.f95:1116-1172),
init_restart reads it (03_restart_primitive.f95:50-61).  The reader below
follows list-directed input: a record may span several lines, values are
separated by blanks or commas, `r*c` repeats c r times, D or E marks the
exponent, and what is left of a record's last line is skipped.  It returns
host planes (14, jmax+13, imax+13), x fastest, ready for Solver.upload.

The Fortran driver writes these files (byte for byte the reference's when
built with the same compiler); Python only reads them.
"""
import numpy as np

NFIELDS = 14


def restart_file_name(hbck):
    """data_restart/all_variables_NNNN.dat as lector names dump hbck."""
    return "data_restart/all_variables_%04d.dat" % hbck if hbck < 1000 else \
        "data_restart/all_variables_%d.dat" % hbck


def _values(line):
    out = []
    for tok in line.replace(",", " ").split():
        if "*" in tok:
            r, c = tok.split("*", 1)
            out.extend([float(c.replace("D", "E").replace("d", "e"))] * int(r))
        else:
            out.append(float(tok.replace("D", "E").replace("d", "e")))
    return out


def read_restart_text(text, imax, jmax):
    """Host planes (14, jmax+13, imax+13) from the text of a restart file."""
    ni, nj = imax + 13, jmax + 13
    nrec = ni * nj
    recs = np.empty((nrec, NFIELDS))
    lines = iter(text.splitlines())
    for r in range(nrec):
        vals = []
        while len(vals) < NFIELDS:
            try:
                vals.extend(_values(next(lines)))
            except StopIteration:
                raise ValueError("restart file ends at record %d of %d" % (r, nrec)) from None
        recs[r] = vals[:NFIELDS]
    # record r = (i+6)*(jmax+13) + (j+6): i outer, j inner
    return np.ascontiguousarray(recs.reshape(ni, nj, NFIELDS).transpose(2, 1, 0))


def read_restart(path, imax, jmax):
    with open(path) as fh:
        return read_restart_text(fh.read(), imax, jmax)
