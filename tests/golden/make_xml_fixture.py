"""Writes tests/golden/readdy_gpu_input/: the reference's GPU-core input files
(test/ReaDDyPaperReaDDyGPU/ReaDDy_input/) as data fixtures for the XML reader
and the JNI entry point.  The coordinate file is cut to its first 800
particles to stay small; the other files are copied unchanged.  Also writes
tests/golden/test_default_groups/: the group files of
test/testInputFiles/test_default/ (a dimer and a tetramer template and one
instance of each), the only reference inputs that hold group instances.

Run in a container that has the reference checkout:
    python tests/golden/make_xml_fixture.py /root/reference
"""
import os
import sys

FILES = ["param_global.xml", "param_particles.xml", "tplgy_potentials.xml", "param_groups.xml", "tplgy_groups.xml"]
KEEP = 800


def main(ref):
    src = os.path.join(ref, "test", "ReaDDyPaperReaDDyGPU", "ReaDDy_input")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "readdy_gpu_input")
    os.makedirs(dst, exist_ok=True)
    for f in FILES:
        with open(os.path.join(src, f), "rb") as a, open(os.path.join(dst, f), "wb") as b:
            b.write(a.read())
    out, kept = [], 0
    for line in open(os.path.join(src, "tplgy_coordinates.xml")):
        if line.lstrip().startswith("<p "):
            if kept >= KEEP:
                continue
            kept += 1
        out.append(line)
    with open(os.path.join(dst, "tplgy_coordinates.xml"), "w") as f:
        f.writelines(out)
    print(f"wrote {dst} ({kept} particles)")
    gsrc = os.path.join(ref, "test", "testInputFiles", "test_default")
    gdst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test_default_groups")
    os.makedirs(gdst, exist_ok=True)
    for f in ("param_groups.xml", "tplgy_groups.xml"):
        with open(os.path.join(gsrc, f), "rb") as a, open(os.path.join(gdst, f), "wb") as b:
            b.write(a.read())
    print(f"wrote {gdst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
