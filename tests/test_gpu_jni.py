"""The JNI entry point Java_readdy_impl_sim_top_TopGPU_cCreateSimulation driven
through a mock JNIEnv (ctypes function table at the JNI-specification
indices), on the reference's GPU input files."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

from readdygpu_amd import Simulation, _lib, output, xml_input

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(__file__), "golden", "readdy_gpu_input")

IDX = dict(DeleteLocalRef=23, GetObjectClass=31, GetMethodID=33, CallBooleanMethod=37, GetFieldID=94,
           GetObjectField=95, GetStringUTFChars=169, ReleaseStringUTFChars=170, GetArrayLength=171,
           NewFloatArray=181, GetPrimitiveArrayCritical=222, ReleasePrimitiveArrayCritical=223)
V = C.c_void_p


class MockJVM:
    """Java side of TopGPU: frameCallback records JPos, cReactions is scripted."""

    OBJ, CLS, MID, FID = 1, 2, 3, 4

    def __init__(self, strings, reactions_by_frame):
        self.strings = {100 + i: s.encode() + b"\0" for i, s in enumerate(strings)}
        self.bufs = {}
        self.arrays = {}
        self.next_id = 1000
        self.frames = []
        self.reactions_by_frame = reactions_by_frame
        self.creactions = self.new_array(np.array([0], np.float32))
        self.unexpected = []
        self._keep = []
        table = (V * 233)()
        trap = C.CFUNCTYPE(None)(lambda: self.unexpected.append("call"))
        self._keep.append(trap)
        for i in range(233):
            table[i] = C.cast(trap, V)

        def reg(name, restype, argtypes, f):
            cb = C.CFUNCTYPE(restype, *argtypes)(f)
            self._keep.append(cb)
            table[IDX[name]] = C.cast(cb, V)

        reg("GetStringUTFChars", V, [V, V, V], self.get_str)
        reg("ReleaseStringUTFChars", None, [V, V, V], lambda e, s, c: None)
        reg("GetObjectClass", V, [V, V], lambda e, o: self.CLS)
        reg("GetMethodID", V, [V, V, C.c_char_p, C.c_char_p],
            lambda e, c, n, s: self.MID if (n, s) == (b"frameCallback", b"(I[F)Z") else None)
        reg("GetFieldID", V, [V, V, C.c_char_p, C.c_char_p],
            lambda e, c, n, s: self.FID if (n, s) == (b"cReactions", b"[F") else None)
        reg("GetObjectField", V, [V, V, V], lambda e, o, f: self.creactions)
        reg("NewFloatArray", V, [V, C.c_int32], lambda e, n: self.new_array(np.zeros(n, np.float32)))
        reg("GetArrayLength", C.c_int32, [V, V], lambda e, a: len(self.arrays[a]))
        reg("GetPrimitiveArrayCritical", V, [V, V, V], lambda e, a, c: self.arrays[a].ctypes.data)
        reg("ReleasePrimitiveArrayCritical", None, [V, V, V, C.c_int32], lambda e, a, p, m: None)
        reg("DeleteLocalRef", None, [V, V], lambda e, o: None)
        reg("CallBooleanMethod", C.c_uint8, [V, V, V, C.c_int32, V], self.frame_callback)
        self.table = table
        self.env_slot = V(C.addressof(table))
        self.env = C.pointer(self.env_slot)

    def new_array(self, a):
        self.next_id += 1
        self.arrays[self.next_id] = a
        return self.next_id

    def get_str(self, env, s, iscopy):
        b = C.create_string_buffer(self.strings[s])
        self._keep.append(b)
        return C.addressof(b)

    def frame_callback(self, env, obj, mid, frame, jpos):
        a = self.arrays[jpos].copy()
        self.frames.append((frame, a))
        cr = self.reactions_by_frame.get(frame, [0])
        self.creactions = self.new_array(np.asarray(cr, np.float32))
        return 1


def fixture_dir(tmp_path, frames, dt):
    for f in os.listdir(FIX):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    g = open(tmp_path / "param_global.xml").read()
    g = g.replace("<nSimulationSteps>100</nSimulationSteps>", f"<nSimulationSteps>{frames}</nSimulationSteps>")
    g = g.replace("<dt>1e-7</dt>", f"<dt>{dt}</dt>")
    open(tmp_path / "param_global.xml", "w").write(g)
    return xml_input.paths_in(str(tmp_path))


def test_jni_entry_frames_and_reactions(tmp_path, monkeypatch):
    paths = fixture_dir(tmp_path, frames=3, dt="1e-9")       # 10 BD steps per frame
    monkeypatch.setenv("RDGPU_SEED", "5")
    lib = _lib.load()
    fn = lib.Java_readdy_impl_sim_top_TopGPU_cCreateSimulation
    fn.restype = None
    fn.argtypes = [V, V, C.c_uint8] + [V] * 6
    order = ["param_particles", "tplgy_potentials", "param_groups", "tplgy_coordinates", "tplgy_groups",
             "param_global"]
    # frame 1: delete index 0 (readdy id 0), create readdy id 5000 of type 1 at (1,2,3)
    jvm = MockJVM([paths[k] for k in order], {1: [1, 5000, 1, 1.0, 2.0, 3.0, 0]})
    fn(V(C.addressof(jvm.env_slot)), V(jvm.OBJ), 0, *[V(100 + i) for i in range(6)])
    assert not jvm.unexpected
    assert [f for f, _ in jvm.frames] == [0, 1, 2]
    n0 = int(jvm.frames[0][1][0])
    assert n0 == 800
    f0 = jvm.frames[0][1][1:].reshape(-1, 5)
    assert np.array_equal(f0[:, 3], np.arange(800)) and np.array_equal(f0[:, 4], np.arange(800))
    f2 = jvm.frames[2][1]
    assert int(f2[0]) == 800
    ids = f2[1:].reshape(-1, 5)[:, 4].astype(int)
    assert 0 not in ids and ids[-1] == 5000 and list(ids[:-1]) == list(range(1, 800))

    # every frame equals the sequential C-ABI loop (simulate, copy, react) with
    # the same seed: the JNI loop runs frame f+1 during the hand-off of frame f
    # and rewinds it when reactions arrive
    x = xml_input.read_xml(paths, seed=5)
    with Simulation(x.spec, x.xyz, x.types) as sim:
        for f in range(3):
            sim.simulate(x.steps_per_frame)
            xyz, _ = sim.copy_pos_from_device()
            assert np.array_equal(jvm.frames[f][1][1:].reshape(-1, 5)[:, :3], xyz), f"frame {f}"
            if f == 1:
                sim.apply_reactions(np.array([1, 5000, 1, 1.0, 2.0, 3.0, 0], np.float32))


def test_jni_entry_with_bonded_groups(tmp_path, monkeypatch):
    """The group files of the reference's test_default inputs (a tetramer over
    particles 0..3, a dimer over 4, 5): the JNI run bonds them exactly as the
    C ABI run of the parsed model does, and the bonds change the motion."""
    paths = fixture_dir(tmp_path, frames=1, dt="1e-9")
    gdir = os.path.join(os.path.dirname(__file__), "golden", "test_default_groups")
    for f in ("param_groups.xml", "tplgy_groups.xml"):
        shutil.copy(os.path.join(gdir, f), tmp_path / f)
    monkeypatch.setenv("RDGPU_SEED", "7")
    outdir = tmp_path / "out"
    outdir.mkdir()
    monkeypatch.setenv("RDGPU_OUTPUT", str(outdir))
    lib = _lib.load()
    fn = lib.Java_readdy_impl_sim_top_TopGPU_cCreateSimulation
    fn.restype = None
    fn.argtypes = [V, V, C.c_uint8] + [V] * 6
    order = ["param_particles", "tplgy_potentials", "param_groups", "tplgy_coordinates", "tplgy_groups",
             "param_global"]
    jvm = MockJVM([paths[k] for k in order], {})
    fn(V(C.addressof(jvm.env_slot)), V(jvm.OBJ), 0, *[V(100 + i) for i in range(6)])
    assert not jvm.unexpected and [f for f, _ in jvm.frames] == [0]
    f0 = jvm.frames[0][1][1:].reshape(-1, 5)
    x = xml_input.read_xml(paths, seed=7)
    assert len(x.spec.bonds) == 5
    with Simulation(x.spec, x.xyz, x.types) as sim:
        sim.simulate(x.steps_per_frame)
        xyz, _ = sim.copy_pos_from_device()
    assert np.array_equal(f0[:, :3], xyz)
    # trajectory files of the frame, as the stand-alone driver writes them
    for name, fmt in (("pdb.pdb", output.PDB), ("out_traj.xml", output.XML)):
        ref = str(tmp_path / ("ref_" + name))
        with output.TrajectoryWriter(ref, fmt) as w:
            w.write_frame(0, xyz, x.types)
        assert open(outdir / name).read() == open(ref).read()
    x.spec.bonds = None
    with Simulation(x.spec, x.xyz, x.types) as sim:
        sim.simulate(x.steps_per_frame)
        free, _ = sim.copy_pos_from_device()
    assert not np.array_equal(free[:6], xyz[:6])


def test_jni_entry_reports_bad_input(tmp_path, capfd):
    paths = fixture_dir(tmp_path, frames=1, dt="1e-9")
    open(paths["tplgy_potentials"], "w").write("<tplgy_potentials><pot type='NOPE'/></tplgy_potentials>")
    lib = _lib.load()
    fn = lib.Java_readdy_impl_sim_top_TopGPU_cCreateSimulation
    fn.restype = None
    fn.argtypes = [V, V, C.c_uint8] + [V] * 6
    order = ["param_particles", "tplgy_potentials", "param_groups", "tplgy_coordinates", "tplgy_groups",
             "param_global"]
    jvm = MockJVM([paths[k] for k in order], {})
    fn(V(C.addressof(jvm.env_slot)), V(jvm.OBJ), 0, *[V(100 + i) for i in range(6)])
    assert jvm.frames == []
    assert "unknown potential type NOPE" in capfd.readouterr().out
