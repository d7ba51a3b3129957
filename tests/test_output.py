"""Trajectory and RDF writers against the reference's printf formats
(outputHandler.cpp:17-148), reproduced here with Python's printf-style
formatting, which rounds the exact binary value half to even as C does."""
import os

import numpy as np
import pytest

from readdygpu_amd import output


@pytest.fixture(scope="module")
def lib_built():
    from readdygpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()


def edge_positions(rng, n):
    x = rng.uniform(-1000, 1000, (n, 3)).astype(np.float32)
    special = np.array([0.0, -0.0, 0.0625, -0.0625, 0.0005, -0.0005, 0.0015, 1e-8, -1e-8, 999.9995, -999.9995,
                        123.4565, 2.5e-4, 1.0, -1.0, 99999.99], np.float32)
    x[:len(special), 0] = special
    x[:len(special), 1] = special[::-1]
    return x


def pdb_reference(frames):
    out = []
    for frame, x, t in frames:
        out.append("MODEL     %d\n" % frame)
        for a in range(len(t)):
            out.append("ATOM  %5d   %i    %i     1    " % (a + 1, t[a], t[a]))
            out.append("%8.3f%8.3f%8.3f 15.00  0.00\n" % (float(x[a, 0]), float(x[a, 1]), float(x[a, 2])))
        out.append("ENDMDL\n")
    return "".join(out)


def xml_reference(frames):
    out = ["<traj>\n"]
    for frame, x, t in frames:
        out.append('<tplgy_coords version="1.1" stepId="%d">\n' % frame)
        for a in range(len(t)):
            out.append('<p id="%d" type="%d" c="%58.3f,%58.3f,%58.3f"/>\n'
                       % (a + 1, t[a], float(x[a, 0]), float(x[a, 1]), float(x[a, 2])))
        out.append("</tplgy_coords>\n")
    out.append("</traj>\n")
    return "".join(out)


@pytest.mark.parametrize("n", [0, 7, 50_000])
def test_trajectory_text_equals_the_printf_formats(lib_built, tmp_path, n):
    rng = np.random.default_rng(n)
    frames = [(f, edge_positions(rng, n) if n >= 16 else rng.uniform(-50, 50, (n, 3)).astype(np.float32),
               rng.integers(0, 12, n).astype(np.int32)) for f in range(2)]
    for fmt, ref in ((output.PDB, pdb_reference), (output.XML, xml_reference)):
        p = str(tmp_path / f"t{fmt}.txt")
        with output.TrajectoryWriter(p, fmt) as w:
            for f, x, t in frames:
                w.write_frame(f, x, t)
        assert open(p).read() == ref(frames)


def test_rdf_table(lib_built, tmp_path):
    rng = np.random.default_rng(3)
    g = rng.uniform(0, 3, (2, 2, 5))
    p = str(tmp_path / "rdf.csv")
    output.write_rdf(p, g, 10.0, names=["A", "B"])
    lines = open(p).read().splitlines()
    assert lines[0] == "d,\tRDF(A,A),\tRDF(A,B),\tRDF(B,A),\tRDF(B,B)"
    for k in range(5):
        row = "%8.5f" % (k * 10.0 / 5) + "".join(", %8.3f" % g[i, j, k] for i in range(2) for j in range(2))
        assert lines[1 + k] == row


def test_bad_format_and_path(lib_built, tmp_path):
    with pytest.raises(OSError):
        output.TrajectoryWriter(str(tmp_path / "x.pdb"), 3)
    with pytest.raises(OSError):
        output.TrajectoryWriter(str(tmp_path / "no" / "such" / "dir.pdb"))
