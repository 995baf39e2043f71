"""Host-side pieces that need no GPU: the .ini grammar, the NS3D command
line's error paths and the .vti codec (CPU only)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from ns3d_amd import solver, vti

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NS3D = os.path.join(ROOT, "ns3d_amd", "NS3D")


def read_ini(tmp_path, text):
    f = tmp_path / "case.ini"
    f.write_text(text)
    p = solver.RunParams()
    buf = C.create_string_buffer(4096)
    rc = solver.lib().ns3d_read_ini(str(f).encode(), C.byref(p), buf, 4096)
    names = buf.raw.split(b"\0")[:5]
    return rc, p, [n.decode() for n in names]


def test_ini_defaults_and_grammar(tmp_path, capfd):
    rc, p, names = read_ini(tmp_path, """
# comment line
INPUT = data/my#image.vti   # inline comment after whitespace
Dx = 2e-6
mu=1e-3
rho = 1e3 # density
pcSolverTol = 1e-4
deltaRMR = 0.5
maxitSIMPLE = 30
method = simplerr
bogus = 1
no equals sign here
""")
    assert rc == 0
    assert names[0] == "data/my#image.vti" and names[1:] == ["none"] * 4
    assert p.dx == 2e-6 and p.dy == 2e-6 and p.dz == 2e-6      # dy/dz default to dx
    assert p.pcsolvertol == 1e-4 and p.maxitSIMPLE == 30
    assert p.method == solver.METHODS["SIMPLERR"]
    assert p.deltaRMR == 1e-4   # the reference never matches this key (navierstokes3D.cxx:342)
    assert p.vsolvertol == 1e-5 and p.psolvertol == 1e-6 and p.maxitsolver == 1000 and p.N1 == 2
    assert p.alpha == 0.8 and p.beta == 1.0 and p.p0 == 1e5 and p.gradp == 1e2 and p.save_step == 20
    err = capfd.readouterr().err
    assert "unknown key 'deltarmr'" in err and "unknown key 'bogus'" in err and "no '=' found" in err


@pytest.mark.parametrize("text,msg", [
    ("dx = 1\n", "required key 'input'"),
    ("input = a.vti\nmu = abc\n", "expects a number"),
    ("input = a.vti\nmaxitsolver = 1.5\n", "expects a integer"),
    ("input = a.vti\ndirection = 4\n", "'direction' must be"),
    ("input = a.vti\nmethod = PISO\n", "unknown method"),
    ("input = a.vti\nrho = 0\n", "'rho' must be > 0"),
])
def test_ini_fatal_errors(tmp_path, capfd, text, msg):
    rc, _, _ = read_ini(tmp_path, text)
    assert rc == 1
    assert msg in capfd.readouterr().err


@pytest.mark.parametrize("text,msg", [
    ("input = a.vti\nmu = 1e-400\n", "expects a number"),        # underflow is an error, as for std::stod
    ("input = a.vti\nmu = 1e999\n", "expects a number"),
    ("input = a.vti\nmaxitsolver = 99999999999\n", "expects a integer"),
    ("input = a.vti\nn1 = 4 5\n", "expects a integer"),
    ("input = a.vti\nmu =\n", "expects a number"),
    ("input = #a.vti\n", "key 'input' has an empty value"),       # '#' opening the value starts a comment
    ("input = a.vti\nmaxitSIMPLE = 0\n", "'maxitSIMPLE' must be > 0"),
    ("input = a.vti\nsave_step = -1\n", "'save_step' must be > 0"),
])
def test_ini_fatal_edge_cases(tmp_path, capfd, text, msg):
    rc, _, _ = read_ini(tmp_path, text)
    assert rc == 1
    assert msg in capfd.readouterr().err


def test_ini_crlf_tabs_and_summary(tmp_path, capfd):
    rc, p, names = read_ini(tmp_path, "\tinput\t=\tdir/x.vti\r\n  MU = 2.5\t# tab comment\r\n"
                                      "method=Ideal\r\nn2 = +3\r\ngradp = 0x10\r\nvxfield = v#1.vti\n")
    assert rc == 0
    assert names[0] == "dir/x.vti" and names[2] == "v#1.vti"
    assert p.mu == 2.5 and p.N2 == 3 and p.gradp == 16.0 and p.method == solver.METHODS["IDEAL"]
    out = capfd.readouterr().out
    assert "method         : IDEAL\n" in out and "mu (Pa.s)      : 2.5\n" in out
    assert "dx, dy, dz (m) : 1, 1, 1\n" in out and out.rstrip().endswith("======================")


def test_ini_porevalue_warning(tmp_path, capfd):
    rc, p, _ = read_ini(tmp_path, "input = a.vti\nporevalue = 7\n")
    assert rc == 0 and p.porevalue == 0
    assert "treating '7' as 0" in capfd.readouterr().err


def test_cli_usage_and_errors(tmp_path):
    r = subprocess.run([NS3D, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage: NS3D" in r.stdout
    r = subprocess.run([NS3D], capture_output=True, text=True)
    assert r.returncode == 1
    r = subprocess.run([NS3D, str(tmp_path / "missing.ini")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open input file" in r.stderr
    ini = tmp_path / "bad.ini"
    ini.write_text("input = nowhere.vti\n")
    r = subprocess.run([NS3D, str(ini)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "cannot open nowhere.vti" in r.stderr


def test_vti_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    a = (rng.random((7, 5, 3)) < 0.4).astype(np.uint8)
    vti.write_vti(str(tmp_path / "a.vti"), a)
    assert np.array_equal(vti.read_vti(str(tmp_path / "a.vti")), a)
    f = rng.standard_normal((6, 4, 5))
    vti.write_vti(str(tmp_path / "f.vti"), f)
    assert np.array_equal(vti.read_vti(str(tmp_path / "f.vti")), f)
    v = rng.standard_normal((3, 4, 2, 3))
    vti.write_vti(str(tmp_path / "v.vti"), v, compress=False)
    assert np.array_equal(vti.read_vti(str(tmp_path / "v.vti")), v)


REF_DATA = "/root/reference/data"


@pytest.mark.parametrize("name", ["spheres_64", "spheres_128"])
def test_product_reader_ingests_reference_images(name):
    """The host library's reader + binarisation (readvtkfile_uint8,
    vtkio.cxx:96-150) on the reference's own image files: shape and sha256 of
    the C-order voxel array equal the golden values, and porevalue 1 swaps
    fluid and solid."""
    import hashlib
    import json
    path = os.path.join(REF_DATA, name + ".vti")
    if not os.path.exists(path):
        pytest.skip("the reference's data files are not present here")
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "spheres128_simplec.json")))["lattice"][name]
    img = solver.read_image(path, 0)
    assert list(img.shape) == gold["shape"]
    assert hashlib.sha256(img.tobytes()).hexdigest() == gold["sha256"]
    assert int(img.sum()) == gold["solid_voxels"]
    assert np.array_equal(solver.read_image(path, 1), 1 - img)


def test_reader_round_trip_and_errors(tmp_path):
    """Files written by the Python codec (UInt32 headers, zlib, appended
    base64, x fastest) read back through the C++ reader; a missing file is an
    error with a message."""
    rng = np.random.default_rng(3)
    a = (rng.random((9, 7, 5)) < 0.4).astype(np.uint8)
    for compress in (True, False):
        f = tmp_path / f"a{int(compress)}.vti"
        vti.write_vti(str(f), a, compress=compress)
        assert np.array_equal(solver.read_image(f, 0), a)
    with pytest.raises(Exception, match="nowhere"):
        solver.read_image(tmp_path / "nowhere.vti", 0)
