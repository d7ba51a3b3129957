"""The host-only C++ of the library (XML reader, trajectory / RDF writers)
built with AddressSanitizer and UndefinedBehaviorSanitizer and run on the
reference's input fixtures, malformed inputs and edge-case frames."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GXX = shutil.which("g++")


@pytest.mark.skipif(GXX is None, reason="g++ not available")
def test_xml_reader_and_writers_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    cmd = [GXX, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-pthread", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "sanitize", "host_sanitize.cpp"),
           os.path.join(ROOT, "readdygpu_amd", "csrc", "rd_xml.cpp"),
           os.path.join(ROOT, "readdygpu_amd", "csrc", "rd_output.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    fix = os.path.join(ROOT, "tests", "golden", "readdy_gpu_input")
    groups = os.path.join(ROOT, "tests", "golden", "test_default_groups")
    r = subprocess.run([exe, fix, groups, str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]
    assert "0 failed checks" in out
