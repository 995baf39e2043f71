"""The drop-in outputs through the NS3D command line (GPU).

The reference's own test case (tests/test_spheres.ini on data/spheres_128.vti,
regenerated bit for bit) runs through `ns3d_amd/NS3D` -- .ini reader, image
reader, host drivers, writers -- for 5 outer iterations with save_step 1:

* RES-1-spheres_128.csv: the header is the reference's; row 0 is the
  reference's row as printed (6 significant digits), it_pc aside (the
  reference prints uninitialised memory there, this build 0); later rows
  follow the reference's trajectory to the pressure-correction tolerance
  (1e-3) and are printed the same way (SIMPLEC.cxx:118, :173);
* INFO-1-spheres_128.log: its first 15 lines are the reference's, and the
  initial K agrees to 1e-9 (NSBase.cxx:107-123, SIMPLEC.cxx:109-133);
* P/V/Vx/Vy/Vz-1-spheres_128.vti: XML ImageData with UInt32 headers, zlib,
  appended base64, the reference files' extents, Float64 (V with 3
  components), one value per node (vtkio.cxx:161-292, NSBase.cxx:330-476);
* the SAVE-1..3 snapshots rotate (SIMPLEC.cxx:200-209): SAVE-1 is rewritten
  at the 4th save and then equals the final fields;
* a warm start from those files (pressurefield / vxfield / vyfield /
  vzfield, NSBase.cxx:2575-2690) equals the in-memory warm start from the
  same values, bit for bit.
"""
from __future__ import annotations

import json
import os
import re
import subprocess

import numpy as np
import pytest

from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver, vti

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NS3D = os.path.join(ROOT, "ns3d_amd", "NS3D")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "spheres128_simplec.json")))

INI = """input = spheres_128.vti
{fields}dx = 2e-6
dy = -1
dz = -1
mu = 1e-3
rho = 1e3
p0 = 1e5
gradp = 1e2
alpha = 0.8
beta = 1.0
vsolvertol = 1e-5
psolvertol = 1e-6
pcsolvertol = 1e-3
simpletol = 1e-4
maxitsolver = 500
maxitSIMPLE = {maxit}
save_step = 1
N1 = 2
N2 = 2
porevalue = 0
direction = 1
border = 0
method = SIMPLEC
"""


def run(d, maxit, fields=""):
    (d / "case.ini").write_text(INI.format(maxit=maxit, fields=fields))
    print(f"[outputs] NS3D in {d} (maxitSIMPLE {maxit})", flush=True)
    r = subprocess.run([NS3D, "case.ini"], cwd=d, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


@pytest.fixture(scope="module")
def lattice():
    img = generators.sphere_lattice(128, 12, 5)
    assert GOLD["lattice"]["spheres_128"]["solid_voxels"] == int(img.sum())
    return img


@pytest.fixture(scope="module")
def first_run(tmp_path_factory, lattice):
    d = tmp_path_factory.mktemp("ns3d_run")
    vti.write_vti(str(d / "spheres_128.vti"), lattice)
    run(d, 5)
    return d


def res_rows(path):
    lines = open(path).read().splitlines()
    return lines[0], [line.split(",") for line in lines[1:] if line.strip()]


def test_res_file(first_run):
    header, rows = res_rows(first_run / "RES-1-spheres_128.csv")
    assert header == GOLD["res_header"]
    assert len(rows) == 5 and [int(r[0]) for r in rows] == list(range(5))
    ref = [r.split(",") for r in GOLD["res_rows"]]
    # row 0 as the reference printed it, it_pc aside
    assert rows[0][:13] == ref[0][:13]
    assert rows[0][13] == "0"
    # every field printed as the reference prints it: 6 significant digits,
    # ", " before maxdiv and it_u
    for r in rows:
        assert r[9].startswith(" ") and r[10].startswith(" ")
        for x in r[1:10]:
            assert x.strip() == f"{float(x):g}", x
    for r, q in zip(rows[1:], ref[1:5]):
        got, want = np.array([float(x) for x in r[1:10]]), np.array([float(x) for x in q[1:10]])
        assert np.allclose(got, want, rtol=2e-2), (r[0], got, want)


def test_info_file(first_run):
    lines = open(first_run / "INFO-1-spheres_128.log").read().splitlines()
    head = GOLD["info_head"]
    # the reference's lines, with this run's cap of 5 outer iterations
    assert lines[:11] == head[:11]
    assert head[11] == "Max number of SIMPLE iterations: 30" and lines[11] == "Max number of SIMPLE iterations: 5"
    assert lines[12:15] == head[12:15]
    m = re.match(r"Initial Permeability K \(m²\): (\S+)", lines[15])
    assert m and head[15].startswith("Initial Permeability K")
    assert abs(float(m.group(1)) / float(head[15].split(": ")[1]) - 1) <= 1e-9


def header_of(path):
    raw = open(path, "rb").read(4096).decode("latin-1")
    return raw[:raw.index("<AppendedData") + 40]


@pytest.mark.parametrize("name,shape,comps", [("P", (128, 128, 128), 1), ("Vx", (129, 128, 128), 1),
                                              ("Vy", (128, 128, 128), 1), ("Vz", (128, 128, 128), 1),
                                              ("V", (128, 128, 128), 3)])
def test_vti_files(first_run, name, shape, comps):
    path = first_run / f"{name}-1-spheres_128.vti"
    h = header_of(path)
    assert 'type="ImageData"' in h and 'header_type="UInt32"' in h and 'compressor="vtkZLibDataCompressor"' in h
    assert 'format="appended"' in h and 'encoding="base64"' in h and 'type="Float64"' in h
    ext = re.search(r'WholeExtent="([^"]+)"', h).group(1)
    if name in GOLD["vti_heads"]:
        assert ext == re.search(r'WholeExtent="([^"]+)"', GOLD["vti_heads"][name]).group(1)
    a = vti.read_vti(str(path))
    assert a.shape == (shape if comps == 1 else shape + (3,))
    counts = dict(P=GOLD["info"]["np"], Vx=GOLD["info"]["nu"], Vy=GOLD["info"]["nv"], Vz=GOLD["info"]["nw"])
    if name in counts:
        assert np.count_nonzero(a) == counts[name]
        assert np.all(np.isfinite(a))


def test_save_rotation(first_run):
    for k in (1, 2, 3):
        assert (first_run / f"P-1-spheres_128-SAVE-{k}.vti").exists(), k
    assert not (first_run / "P-1-spheres_128-SAVE-4.vti").exists()
    for name in ("P", "Vx"):
        final = vti.read_vti(str(first_run / f"{name}-1-spheres_128.vti"))
        s1 = vti.read_vti(str(first_run / f"{name}-1-spheres_128-SAVE-1.vti"))
        s2 = vti.read_vti(str(first_run / f"{name}-1-spheres_128-SAVE-2.vti"))
        s3 = vti.read_vti(str(first_run / f"{name}-1-spheres_128-SAVE-3.vti"))
        assert np.array_equal(s1, final), name   # SAVE-1 rewritten at iteration 4, the last
        assert not np.array_equal(s2, s3) and not np.array_equal(s3, final), name


def test_file_warm_start_equals_memory_warm_start(first_run, tmp_path, lattice):
    for n in ("P", "Vx", "Vy", "Vz"):
        (tmp_path / f"{n}.vti").write_bytes((first_run / f"{n}-1-spheres_128.vti").read_bytes())
    vti.write_vti(str(tmp_path / "spheres_128.vti"), lattice)
    run(tmp_path, 2, fields="pressurefield = P.vti\nvxfield = Vx.vti\nvyfield = Vy.vti\nvzfield = Vz.vti\n")
    info = open(tmp_path / "INFO-1-spheres_128.log").read()
    k_file = [float(x) for x in re.findall(r"Permeability K \(m²\): (\S+)", info)]
    # the same values handed over in memory
    with G.NavierStokesGPU(lattice) as g:
        pg, ug, vg, wg = g.index()
    grids = {n: vti.read_vti(str(tmp_path / f"{n}.vti")) for n in ("P", "Vx", "Vy", "Vz")}

    def nodes(grid, idx):
        x = np.zeros(int(idx.max()) + 1)
        m = idx >= 0
        x[idx[m]] = grid[m]
        return x
    prm = dict(dx=2e-6, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2, alpha=0.8, beta=1.0, vsolvertol=1e-5,
               psolvertol=1e-6, pcsolvertol=1e-3, simpletol=1e-4, maxitsolver=500, maxitSIMPLE=2, device=0,
               method="SIMPLEC", save_step=1)
    s = solver.Solver(lattice, **prm)
    s.warm_start(nodes(grids["P"], pg), nodes(grids["Vx"], ug), nodes(grids["Vy"], vg), nodes(grids["Vz"], wg))
    rows, summ = s.run()
    s.close()
    k_mem = [float(f"{k:.15g}") for k in rows[:, 4]]
    # INFO: the initial K, then one K per outer iteration (15 digits)
    assert k_file[1:1 + len(k_mem)] == k_mem, (k_file, k_mem)
    # and the warm start is a warm start: row 0 continues where the first
    # run stopped, not from the cold initial K
    _, rows1 = res_rows(first_run / "RES-1-spheres_128.csv")
    assert abs(k_mem[0] / float(rows1[-1][4]) - 1) < 0.2
    assert abs(k_mem[0] / float(rows1[0][4]) - 1) > 0.5
