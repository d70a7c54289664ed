"""A compiled C99 program calls the library the way the reference's users do
(tests/c/caller.c: get_initialized_system, acceleration() with the pairwise
and Barnes-Hut methods, launch_simulation() with Leapfrog).  It is built with
gcc against this library's header and, where the reference sources are
present, against the reference's own headers (src/grav_sim.h), and linked to
libgrav_sim.so unchanged."""
import os
import struct
import subprocess

import numpy as np
import pytest

from grav_sim_amd import ic
from grav_sim_amd import lib as L
from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "c", "caller.c")
LIBDIR = os.path.join(REPO, "gravity-simulator_amd")
REF_SRC = "/root/reference/src"


def build(out, reference_headers=False):
    inc = ["-DUSE_REFERENCE_HEADERS", f"-I{REF_SRC}", "-I/root/reference/pcg"] if reference_headers \
        else [f"-I{os.path.join(REPO, 'include')}"]
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", *inc, SRC, f"-L{LIBDIR}", "-lgrav_sim",
           f"-Wl,-rpath,{LIBDIR}", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def write_input(path, x, v, m):
    with open(path, "wb") as f:
        f.write(struct.pack("i", len(m)) + np.ascontiguousarray(x).tobytes() + np.ascontiguousarray(v).tobytes()
                + np.ascontiguousarray(m).tobytes())


def test_c_caller_builds_links_and_reports_errors_as_error_status(tmp_path):
    """Without a GPU the program gets the library's ErrorStatus (no device)
    and prints its traceback; with one it succeeds."""
    exe = build(str(tmp_path / "caller"))
    x, v, m = ic.plummer(64, 3)
    write_input(tmp_path / "in.bin", x, v, m)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode in (0, 1), r.stderr
    if r.returncode == 1:
        assert "No HIP device available" in r.stdout + r.stderr


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")
def test_c_caller_builds_against_the_reference_headers(tmp_path):
    build(str(tmp_path / "caller_ref"), reference_headers=True)


@pytest.mark.gpu
def test_c_caller_results_on_the_gpu(tmp_path):
    exe = build(str(tmp_path / "caller"))
    x, v, m = ic.two_spheres(3000, 19)
    write_input(tmp_path / "in.bin", x, v, m)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(4, len(m), 3)
    assert np.array_equal(out[0], O.pairwise(x, m, eps=1e-3))
    assert np.array_equal(out[1], L.acceleration(x, m, 3, eps=1e-3, theta=0.5, leaf=1))
    xf, vf, steps = O.leapfrog(x, v, m, 1.0, 1e-3, 10 * 1e-3, eps=1e-3, method=1)
    assert steps == 10
    assert np.array_equal(out[2], xf) and np.array_equal(out[3], vf)
