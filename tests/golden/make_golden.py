"""Regenerate the golden fixtures under tests/golden/ from the reference's own
test data (run in a container that has the reference tree and h5dump).

Fixtures written (all data, no code):
  revolving_sphere_3d/<variant>_556.f32.gz   float32 psi[x][y][z], 59^3, after
                                              556 steps (refdata/<variant>/
                                              timestep0000000556.h5, dataset psi)
  revolving_sphere_3d/initial_0.f32.gz       the t=0 field (refdata/common/...)
  revolving_sphere_3d/stats_<variant>.txt    refdata/stats_<variant>.txt.gz
  rotating_cone_2d/stats_<variant>_.txt      refdata/stats_<variant>_.txt.gz
  coupled_harmosc/stats.txt                  6_coupled_harmosc/refdata/stats.txt.gz
  revolving_sphere_3d_h5/*.xmf               refdata/fct/*.xmf (the same in every variant)
  revolving_sphere_3d_h5/const_<variant>.txt h5dump -A of refdata/<variant>/const.h5
  revolving_sphere_3d_h5/const_data.sha256   sha256 of the raw float32 X, Y, Z, T of
                                              each refdata/<variant>/const.h5
  SHA256SUMS                                 checksums of the above

The HDF5 datasets are float32 little-endian, C order (output/hdf5.hpp:56,
327-357); h5dump -b LE writes their raw bytes.
"""
import gzip
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

REF = "/root/reference/tests/paper_2015_GMD"
H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))
SPHERE_VARIANTS = ["upwind", "basic", "iga", "iga_fct", "fct"]
CONE_VARIANTS = ["basic", "fct", "iga_fct", "iga_tot_fct", "iters3_tot_fct"]


def dump_psi(h5path, out_gz):
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "psi.bin")
        subprocess.run([H5DUMP, "-d", "/psi", "-b", "LE", "-o", raw, h5path],
                       check=True, stdout=subprocess.DEVNULL, cwd=tmp)
        data = open(raw, "rb").read()
    assert len(data) == 59 ** 3 * 4, len(data)
    with gzip.GzipFile(out_gz, "wb", mtime=0) as f:
        f.write(data)


def h5_fixtures(rs):
    """The output layout of the revolving-sphere runs (output::hdf5_xdmf):
    the XDMF files (the same bytes in every variant's directory), the
    header and attribute dump of each variant's const.h5 (h5dump -A), and
    checksums of the raw float32 bytes of its X, Y, Z and T datasets."""
    out = os.path.join(HERE, "revolving_sphere_3d_h5")
    os.makedirs(out, exist_ok=True)
    xmfs = ["timestep0000000000.xmf", "timestep0000000556.xmf", "temp.xmf"]
    for name in xmfs:
        data = open(os.path.join(rs, "fct", name), "rb").read()
        for v in SPHERE_VARIANTS:
            assert open(os.path.join(rs, v, name), "rb").read() == data, (v, name)
        open(os.path.join(out, name), "wb").write(data)
    sums = []
    for v in SPHERE_VARIANTS:
        vdir = os.path.join(rs, v)
        txt = subprocess.run([H5DUMP, "-A", "const.h5"], cwd=vdir, capture_output=True, check=True).stdout
        open(os.path.join(out, f"const_{v}.txt"), "wb").write(txt)
        with tempfile.TemporaryDirectory() as tmp:
            for ds in ("X", "Y", "Z", "T"):
                raw = os.path.join(tmp, ds + ".bin")
                subprocess.run([H5DUMP, "-d", "/" + ds, "-b", "LE", "-o", raw, os.path.join(vdir, "const.h5")],
                               check=True, stdout=subprocess.DEVNULL, cwd=tmp)
                sums.append(f"{hashlib.sha256(open(raw, 'rb').read()).hexdigest()}  {v}/const.h5:/{ds}")
    with open(os.path.join(out, "const_data.sha256"), "w") as f:
        f.write("\n".join(sums) + "\n")


def advscheme_opts_records():
    """1_advscheme_opts/refdata/log-orig.gz: the same gnuplot layout, 20
    blocks (options abs, iga, iga|tot, iga|fct, iga|tot|fct; per option the
    two equations at t/dt = 0 and 1200) of the 601 cell values.  Kept: one
    line per block, in the log's order."""
    out = os.path.join(HERE, "advscheme_opts")
    os.makedirs(out, exist_ok=True)
    text = gzip.open(os.path.join(REF, "1_advscheme_opts", "refdata", "log-orig.gz"), "rt").read()
    blocks, cur, data = [], [], False
    for l in text.split("\n"):
        if l.startswith("splot") or l.startswith("plot"):
            data = True
            continue
        if not data:
            continue
        if l == "e":
            # histeps plots print every value twice
            doubled = len(cur) % 2 == 0 and cur[0::2] == cur[1::2]
            blocks.append(cur[::2] if doubled else cur)
            cur = []
        elif l.strip() and not l.startswith("set") and not l.startswith("reset"):
            try:
                float(l)
            except ValueError:
                data = False
                continue
            cur.append(l.strip())
    with open(os.path.join(out, "records.txt"), "w") as f:
        for r in blocks:
            f.write(" ".join(r) + "\n")


def basic_example_records():
    """0_basic_example/refdata/log-orig.gz is the gnuplot script of the
    example's output: six records (t/dt = 0, 20, ..., 100) of the 101 cell
    values, each printed twice (histeps), blocks ended by 'e'.  Kept: one
    line per record with the 101 values as printed."""
    out = os.path.join(HERE, "basic_example")
    os.makedirs(out, exist_ok=True)
    text = gzip.open(os.path.join(REF, "0_basic_example", "refdata", "log-orig.gz"), "rt").read()
    body = text.split("\n")
    start = next(k for k, l in enumerate(body) if l.startswith("splot")) + 1
    records, cur = [], []
    for l in body[start:]:
        if l == "e":
            records.append(cur[::2])
            cur = []
        elif l.strip():
            cur.append(l.strip())
    with open(os.path.join(out, "records.txt"), "w") as f:
        for r in records:
            f.write(" ".join(r) + "\n")


def main():
    sphere = os.path.join(HERE, "revolving_sphere_3d")
    cone = os.path.join(HERE, "rotating_cone_2d")
    os.makedirs(sphere, exist_ok=True)
    os.makedirs(cone, exist_ok=True)
    rs = os.path.join(REF, "4_revolving_sphere_3d", "refdata")
    dump_psi(os.path.join(rs, "common", "timestep0000000000.h5"),
             os.path.join(sphere, "initial_0.f32.gz"))
    for v in SPHERE_VARIANTS:
        dump_psi(os.path.join(rs, v, "timestep0000000556.h5"),
                 os.path.join(sphere, f"{v}_556.f32.gz"))
        with gzip.open(os.path.join(rs, f"stats_{v}.txt.gz"), "rb") as f, \
                open(os.path.join(sphere, f"stats_{v}.txt"), "wb") as g:
            shutil.copyfileobj(f, g)
    rc = os.path.join(REF, "3_rotating_cone_2d", "refdata")
    for v in CONE_VARIANTS:
        with gzip.open(os.path.join(rc, f"stats_{v}_.txt.gz"), "rb") as f, \
                open(os.path.join(cone, f"stats_{v}_.txt"), "wb") as g:
            shutil.copyfileobj(f, g)
    osc = os.path.join(HERE, "coupled_harmosc")
    os.makedirs(osc, exist_ok=True)
    with gzip.open(os.path.join(REF, "6_coupled_harmosc", "refdata", "stats.txt.gz"), "rb") as f, \
            open(os.path.join(osc, "stats.txt"), "wb") as g:
        shutil.copyfileobj(f, g)
    h5_fixtures(rs)
    basic_example_records()
    advscheme_opts_records()
    lines = []
    for sub in ("revolving_sphere_3d", "revolving_sphere_3d_h5", "rotating_cone_2d", "coupled_harmosc",
                "basic_example", "advscheme_opts"):
        for name in sorted(os.listdir(os.path.join(HERE, sub))):
            path = os.path.join(HERE, sub, name)
            digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
            lines.append(f"{digest}  {sub}/{name}")
    with open(os.path.join(HERE, "SHA256SUMS"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    sys.exit(main())
