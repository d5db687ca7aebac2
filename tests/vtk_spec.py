"""A token-level validator of the legacy VTK file format (POLYDATA, ASCII),
written from the "Simple Legacy Formats" section of the VTK file-format
specification, independent of rbffd_amd.vtk.read_vtk."""
import numpy as np


TYPES = {"bit", "unsigned_char", "char", "unsigned_short", "short", "unsigned_int", "int", "unsigned_long",
         "long", "float", "double"}


def validate_legacy_polydata(text):
    lines = text.split("\n")
    assert lines[0].startswith("# vtk DataFile Version "), "header line"
    assert len(lines[1]) <= 256, "title longer than 256 characters"
    toks = " ".join(lines[2:]).split()
    pos = 0

    def take(n=1):
        nonlocal pos
        out = toks[pos:pos + n]
        assert len(out) == n, "file ends early"
        pos += n
        return out if n > 1 else out[0]

    assert take() == "ASCII"
    assert take(2) == ["DATASET", "POLYDATA"]
    assert take() == "POINTS"
    npts = int(take())
    assert take() in TYPES
    pts = np.array([float(v) for v in take(3 * npts)]).reshape(npts, 3)
    cells = []
    fields = {}
    while pos < len(toks):
        kw = take()
        if kw in ("VERTICES", "LINES", "POLYGONS", "TRIANGLE_STRIPS"):
            ncell, size = int(take()), int(take())
            used = 0
            for _ in range(ncell):
                m = int(take())
                idx = [int(v) for v in take(m)] if m > 1 else [int(take())]
                assert all(0 <= i < npts for i in idx), "cell index out of range"
                cells.append(idx)
                used += 1 + m
            assert used == size, "cell section size"
        elif kw == "POINT_DATA":
            assert int(take()) == npts
            while pos < len(toks):
                assert take() == "SCALARS"
                name = take()
                assert take() in TYPES
                ncomp = toks[pos]
                if ncomp.isdigit():
                    take()
                    assert int(ncomp) == 1
                assert take(2) == ["LOOKUP_TABLE", "default"]
                fields[name] = np.array([float(v) for v in take(npts)])
        else:
            raise AssertionError(f"unexpected keyword {kw}")
    return pts, cells, fields
