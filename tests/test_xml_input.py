"""XML input reader against the reference's GPU-core input files
(tests/golden/readdy_gpu_input, from test/ReaDDyPaperReaDDyGPU/ReaDDy_input)."""
import os
import shutil

import numpy as np
import pytest

from readdygpu_amd import BOX, HARMONIC, REPULSIVE, ATTRACTIVE, SPRING
from readdygpu_amd import xml_input

FIX = os.path.join(os.path.dirname(__file__), "golden", "readdy_gpu_input")


@pytest.fixture(scope="module")
def lib_built():
    from readdygpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()


def test_reads_reference_gpu_input(lib_built):
    x = xml_input.read_xml(xml_input.paths_in(FIX))
    s = x.spec
    # param_global.xml: dt 1e-7, dtOpenMM 1e-10, T 293.15, Kb, latticeBounds
    assert x.steps_per_frame == 1000 and x.frames == 100
    assert s.dt == pytest.approx(1e-10) and s.T == pytest.approx(293.15) and s.kB == pytest.approx(0.0083144621)
    assert s.box_lo == (-100.0, -100.0, -50.0) and s.box_hi == (100.0, 100.0, 150.0)
    # param_particles.xml: three types; D and default collision radii
    assert s.n_types == 3
    assert [t.D for t in s.types][:2] == pytest.approx([143.1e6, 71.6e6])
    assert [t.radius for t in s.types] == pytest.approx([1.5, 3.0, 3.12])
    # tplgy_potentials.xml: CUBE attractive k=10 + harmonic repulsion for all pairs
    assert len(s.pot1) == 1 and s.pot1[0].kind == BOX and s.pot1[0].subtype == ATTRACTIVE
    assert s.pot1[0].k == 10 and s.pot1[0].origin == (-50.0, -50.0, 0.0)
    assert s.pot1[0].extension == (100.0, 100.0, 100.0) and sorted(s.pot1[0].types) == [0, 1, 2]
    assert len(s.pot2) == 1 and s.pot2[0].kind == HARMONIC and s.pot2[0].subtype == REPULSIVE
    assert len(s.pot2[0].pairs) == 6          # "all": every unordered type pair
    # coordinates: first and last line of the fixture
    assert len(x.types) == 800 and x.readdy_ids[0] == 0 and x.readdy_ids[-1] == 799
    assert np.allclose(x.xyz[0], [-22.978307943118587, 4.877510590607036, 15.886088294498862])
    assert np.allclose(x.xyz[-1], [22.129760623915363, -6.773142742259554, 97.42074992892849])
    assert x.n_groups == 0


def test_number_lists_in_both_bracket_styles(lib_built, tmp_path):
    """"[[0,1];[1,1]]" (Java style) and "[0,1];[1,1]" (C++ style) give the same
    pairs; "[1,2]" lists two types (the reference's stringToIntVector loops on it)."""
    for f in os.listdir(FIX):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    pot = open(tmp_path / "tplgy_potentials.xml").read()
    pot = pot.replace('affectedParticleTypeIdPairs="all"', 'affectedParticleTypeIdPairs="[[0,1];[1,1]]"')
    pot = pot.replace('affectedParticleTypeIds="all"', 'affectedParticleTypeIds="[1,2]"')
    open(tmp_path / "tplgy_potentials.xml", "w").write(pot)
    x = xml_input.read_xml(xml_input.paths_in(str(tmp_path)))
    assert sorted(x.spec.pot2[0].pairs) == [(0, 1), (1, 1)]
    assert sorted(x.spec.pot1[0].types) == [1, 2]
    open(tmp_path / "tplgy_potentials.xml", "w").write(pot.replace("[[0,1];[1,1]]", "[0,1];[1,1]"))
    y = xml_input.read_xml(xml_input.paths_in(str(tmp_path)))
    assert sorted(y.spec.pot2[0].pairs) == [(0, 1), (1, 1)]


def test_malformed_and_missing_files_are_errors(lib_built, tmp_path):
    for f in os.listdir(FIX):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    p = xml_input.paths_in(str(tmp_path))
    bad = dict(p, param_global=str(tmp_path / "missing.xml"))
    with pytest.raises(ValueError, match="cannot read"):
        xml_input.read_xml(bad)
    open(tmp_path / "param_particles.xml", "w").write("<param_particles><particle><type>A</type>")
    with pytest.raises(ValueError, match="unclosed"):
        xml_input.read_xml(p)
    open(tmp_path / "param_particles.xml", "w").write(open(os.path.join(FIX, "param_particles.xml")).read())
    open(tmp_path / "tplgy_coordinates.xml", "a").write('<p id="9" type="7" c="0,0,0"/>\n')
    with pytest.raises(ValueError, match="unknown type"):
        xml_input.read_xml(p)


def test_group_files_become_bonds(lib_built, tmp_path):
    """The reference's test_default group files (test/testInputFiles/
    test_default): an A_DIMER template (HARMONIC k=1 on internal pair [0,1])
    and an A_TETRAMER template (HARMONIC k=1 on [0,1];[0,2];[1,3];[2,3]), one
    instance of each: <g type=1> over particles 0..3, <g type=0> over 4,5.
    Every affected internal pair of every instance is a bond; no subtype
    attribute means spring (P2_Harmonic.java:47-49)."""
    for f in os.listdir(FIX):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    gdir = os.path.join(os.path.dirname(__file__), "golden", "test_default_groups")
    for f in ("param_groups.xml", "tplgy_groups.xml"):
        shutil.copy(os.path.join(gdir, f), tmp_path / f)
    x = xml_input.read_xml(xml_input.paths_in(str(tmp_path)))
    s = x.spec
    assert x.n_groups == 2
    assert len(s.bond_pots) == 2 and all(p.kind == HARMONIC and p.subtype == SPRING and p.k == 1.0
                                         for p in s.bond_pots)
    bonds = sorted((int(a), int(b), int(p)) for (a, b), p in zip(s.bonds, s.bond_pot))
    assert bonds == [(0, 1, 1), (0, 2, 1), (1, 3, 1), (2, 3, 1), (4, 5, 0)]
    assert s.bond_r0 == [None, None]


def test_group_instance_of_unknown_template_is_an_error(lib_built, tmp_path):
    for f in os.listdir(FIX):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    open(tmp_path / "tplgy_groups.xml", "w").write(
        '<tplgy_groups version="1.1">\n<g id="0" type="0" internalAndParticleId="[[0,0];[1,1]]"/>\n</tplgy_groups>\n')
    with pytest.raises(ValueError, match="unknown group type"):
        xml_input.read_xml(xml_input.paths_in(str(tmp_path)))
