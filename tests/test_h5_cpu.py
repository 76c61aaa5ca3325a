"""HDF5 / XDMF output (include/mpdx_h5.h) against the reference's own output
files of the revolving-sphere runs (tests/paper_2015_GMD/4_revolving_sphere_3d/
refdata/<variant>/, written by output::hdf5_xdmf): the .xmf files byte for byte,
the const.h5 headers and attributes (h5dump -A) and the raw float32 bytes of
its X, Y, Z and T, and the psi datasets read back.  The fields written are the
reference's own (golden fixtures), so these checks need no GPU."""
import ctypes
import gzip
import hashlib
import math
import os
import re
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from libmpdataxx_amd import build as B
from libmpdataxx_amd import h5 as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = os.path.join(B.HDF5_ROOT, "bin", "h5dump")
H5DIFF = os.path.join(B.HDF5_ROOT, "bin", "h5diff")
REFDATA = "/root/reference/tests/paper_2015_GMD/4_revolving_sphere_3d/refdata"
FCT, IGA = 1 << 0, 1 << 5
# variant -> (opts, n_iters) (revolving_sphere_3d.cpp:104-110)
VARIANTS = {"upwind": (0, 1), "iga": (IGA, 2), "basic": (0, 2), "iga_fct": (IGA | FCT, 2), "fct": (FCT, 2)}
NX, NT = 59, 556
DX = 100 / (NX - 1)
DT = 0.018 * 2 * math.pi

pytestmark = pytest.mark.skipif(not B.have_hdf5(), reason="HDF5 is not installed")


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(H.H5LIB_PATH):
        pytest.fail("libmpdx_h5.so not built (python -m libmpdataxx_amd.build)")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mpdx_h5.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpdx_h5_[a-z_]+)\s*\(", text)))


def test_library_exports_declared_symbols():
    assert declared_symbols() == sorted(H.EXPORTED)
    lib = ctypes.CDLL(H.H5LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def golden_field(golden_dir, name):
    raw = gzip.open(os.path.join(golden_dir, "revolving_sphere_3d", name)).read()
    return np.frombuffer(raw, dtype="<f4").reshape(NX, NX, NX)


def golden_const_sums(golden_dir):
    out = {}
    for line in open(os.path.join(golden_dir, "revolving_sphere_3d_h5", "const_data.sha256")):
        digest, key = line.split()
        out[key] = digest
    return out


def sphere_setup(variant, real_bytes=8):
    opts, iters = VARIANTS[variant]
    return H.make_setup((NX, NX, NX), real_bytes, (DX, DX, DX), DT, iters, opts, outfreq=NT)


def write_sphere(outdir, variant, golden_dir):
    s = sphere_setup(variant)
    H.start(s, outdir, NT)
    H.record(s, outdir, 0, 0.0, {"psi": golden_field(golden_dir, "initial_0.f32.gz")})
    # time = timestep * dt (solver_common.hpp:348)
    H.record(s, outdir, NT, NT * DT, {"psi": golden_field(golden_dir, f"{variant}_556.f32.gz")}, [0])


def h5dump_attrs(directory, name="const.h5"):
    return subprocess.run([H5DUMP, "-A", name], cwd=directory, capture_output=True, check=True).stdout


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_revolving_sphere_layout_matches_refdata(tmp_path, golden_dir, variant):
    out = tmp_path / variant
    write_sphere(out, variant, golden_dir)
    g = os.path.join(golden_dir, "revolving_sphere_3d_h5")
    for name in ("timestep0000000000.xmf", "timestep0000000556.xmf", "temp.xmf"):
        assert open(out / name, "rb").read() == open(os.path.join(g, name), "rb").read(), name
    assert h5dump_attrs(out) == open(os.path.join(g, f"const_{variant}.txt"), "rb").read()
    sums = golden_const_sums(golden_dir)
    shapes = {"X": (NX + 1,) * 3, "Y": (NX + 1,) * 3, "Z": (NX + 1,) * 3, "T": (2,)}
    for ds, shape in shapes.items():
        a = H.read(str(out / "const.h5"), ds, shape)
        assert hashlib.sha256(a.astype("<f4").tobytes()).hexdigest() == sums[f"{variant}/const.h5:/{ds}"], ds
    np.testing.assert_array_equal(H.read(str(out / "timestep0000000000.h5"), "psi", (NX,) * 3),
                                  golden_field(golden_dir, "initial_0.f32.gz"))
    np.testing.assert_array_equal(H.read(str(out / "timestep0000000556.h5"), "psi", (NX,) * 3),
                                  golden_field(golden_dir, f"{variant}_556.f32.gz"))


@pytest.mark.skipif(not os.path.isdir(REFDATA), reason="reference refdata not present")
@pytest.mark.parametrize("variant", ["fct", "upwind"])
def test_h5diff_against_reference_files(tmp_path, golden_dir, variant):
    """h5diff (data and attributes) of the files written here against the
    reference's own output files: no differences."""
    out = tmp_path / variant
    write_sphere(out, variant, golden_dir)
    for name in ("const.h5", "timestep0000000000.h5", "timestep0000000556.h5"):
        r = subprocess.run([H5DIFF, str(out / name), os.path.join(REFDATA, variant, name)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (name, r.stdout, r.stderr)


def test_record_dataset_properties(tmp_path, golden_dir):
    """hdf5.hpp:109, 157-164, 271-278: float32, chunk = the whole grid,
    deflate level 5."""
    write_sphere(tmp_path, "fct", golden_dir)
    txt = subprocess.run([H5DUMP, "-p", "-H", "timestep0000000556.h5"], cwd=tmp_path, capture_output=True,
                         text=True, check=True).stdout
    assert "H5T_IEEE_F32LE" in txt
    assert f"CHUNKED ( {NX}, {NX}, {NX} )" in txt
    assert "COMPRESSION DEFLATE { LEVEL 5 }" in txt


def schedule(nt, outfreq, outstart, outwindow):
    """Timesteps recorded by one advance(nt) from timestep 0, enumerated from
    output_common's hooks: hook_ante_loop records once, hook_post_step once
    for every t in [0, outwindow) with (timestep - t) % outfreq == 0 and
    timestep - t >= outstart (output_common.hpp:47-62, 145-156)."""
    rec = [0]
    for ts in range(1, nt + 1):
        for t in range(outwindow):
            if (ts - t) % outfreq == 0 and ts - t >= outstart:
                rec.append(ts)
    return rec


@pytest.mark.parametrize("nt,outfreq,outstart,outwindow", [
    (10, 3, 0, 1), (9, 3, 0, 1), (10, 3, 0, 2), (9, 3, 0, 2), (11, 3, 0, 3), (10, 3, 3, 1), (12, 4, 8, 1),
    (556, 556, 0, 1), (7, 1, 0, 1), (20, 5, 0, 4),
])
def test_time_axis_matches_record_schedule(tmp_path, nt, outfreq, outstart, outwindow):
    """The T axis of const.h5 lists the time of every record an advance(nt)
    makes (the reference computes it in closed form, hdf5.hpp:199-215)."""
    dt = 0.25
    s = H.make_setup((4, 5), 8, (1.0, 1.0), dt, 2, FCT, outfreq=outfreq, outstart=outstart, outwindow=outwindow)
    H.start(s, tmp_path, nt)
    rec = schedule(nt, outfreq, outstart, outwindow)
    assert H.n_out(s, nt) == len(rec)
    np.testing.assert_array_equal(H.read(str(tmp_path / "const.h5"), "T", (len(rec),)),
                                  np.array(rec, dtype=np.float64) * dt)


def test_float_solver_coordinates_and_times(tmp_path):
    """With real_t = float the coordinates are d * index and the times
    dt * k computed in float, then stored (hdf5.hpp:166-215)."""
    d, dt = 0.1, 0.3
    s = H.make_setup((5, 4, 3), 4, (d, 2 * d, 3 * d), dt, 2, FCT, outfreq=2)
    H.start(s, tmp_path, 6)
    i = np.arange(6, dtype=np.float32)
    x = H.read(str(tmp_path / "const.h5"), "X", (6, 5, 4))
    np.testing.assert_array_equal(x, np.broadcast_to((np.float32(d) * i)[:, None, None], (6, 5, 4)))
    z = H.read(str(tmp_path / "const.h5"), "Z", (6, 5, 4))
    np.testing.assert_array_equal(z, np.broadcast_to((np.float32(3 * d) * i[:4])[None, None, :], (6, 5, 4)))
    t = H.read(str(tmp_path / "const.h5"), "T", (4,))
    want = (np.float64(np.float32(dt)) * np.array([0, 2, 4, 6])).astype(np.float32)
    np.testing.assert_array_equal(t, want)


def test_two_dimensional_markup(tmp_path):
    """xdmf_writer.hpp:50-75: 2DSMesh topology and X_Y geometry over n+1
    corner points; cell attributes over n cells, sorted by name."""
    s = H.make_setup((6, 4), 8, (0.5, 0.25), 0.1, 3, IGA | FCT)
    H.start(s, tmp_path, 1)
    rng = np.random.default_rng(5)
    f = {"rho": rng.random((6, 4)), "chi": rng.random((6, 4))}
    H.record(s, tmp_path, 0, 0.0, f)
    H.record(s, tmp_path, 1, 0.1, f, [0])
    root = ET.parse(tmp_path / "timestep0000000001.xmf").getroot()
    grid = root.find("Domain/Grid")
    assert grid.find("Time").get("Value") == "0.100000"
    assert grid.find("Topology").get("TopologyType") == "2DSMesh"
    assert grid.find("Topology").get("Dimensions") == "7 5 "
    geo = grid.find("Geometry")
    assert geo.get("GeometryType") == "X_Y"
    assert [e.text for e in geo.findall("DataItem")] == ["const.h5:/X", "const.h5:/Y"]
    attrs = grid.findall("Attribute")
    assert [a.get("Name") for a in attrs] == ["chi", "rho"]
    assert [a.find("DataItem").text for a in attrs] == ["timestep0000000001.h5:/chi", "timestep0000000001.h5:/rho"]
    assert all(a.find("DataItem").get("Dimensions") == "6 4 " for a in attrs)
    temp = ET.parse(tmp_path / "temp.xmf").getroot()
    incl = [e.get("href") for e in temp.iter("{http://www.w3.org/2001/XInclude}include")]
    assert incl == ["timestep0000000000.xmf", "timestep0000000001.xmf"]
    for name, a in f.items():
        np.testing.assert_array_equal(H.read(str(tmp_path / "timestep0000000001.h5"), name, (6, 4)),
                                      a.astype(np.float32))
    hdr = h5dump_attrs(tmp_path).decode()
    assert '"Z"' not in hdr and '"dk"' not in hdr and '"dj"' in hdr
    assert '"iga|fct"' in hdr


def test_hdf5_only_writes_no_markup(tmp_path):
    s = H.make_setup((4, 4, 4), 8, xdmf=False)
    H.start(s, tmp_path, 1)
    H.record(s, tmp_path, 0, 0.0, {"psi": np.ones((4, 4, 4))})
    assert sorted(os.listdir(tmp_path)) == ["const.h5", "timestep0000000000.h5"]


def test_errors_are_reported(tmp_path):
    with pytest.raises(RuntimeError, match="only 2D and 3D"):
        H.start(H.make_setup((8,)), tmp_path, 1)
    s = H.make_setup((4, 4))
    with pytest.raises(ValueError, match="shape"):
        H.record(s, tmp_path, 0, 0.0, {"psi": np.ones((4, 5))})
    H.start(s, tmp_path, 1)
    with pytest.raises(RuntimeError, match="no dataset"):
        H.read(str(tmp_path / "const.h5"), "psi", (4, 4))
    with pytest.raises(RuntimeError, match="holds 25 values"):
        H.read(str(tmp_path / "const.h5"), "X", (4, 4))
    with pytest.raises(RuntimeError, match="cannot create"):
        H.start(s, tmp_path / "a" / "b", 1)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(H, "H5LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(H, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        H.lib()
