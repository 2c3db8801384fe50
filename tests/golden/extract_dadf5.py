"""Extract the golden DADF5 results shipped with the reference into small
fixtures (run in the build container only; needs /root/reference and h5dump).

Sources (data files held by the reference's own tests):

* ``python/tests/resources/Result/6grains6x7x8_tensionY_singlePhase.hdf5`` --
  DAMASK_grid spectral_basic, FCC phenopowerlaw, 6x7x8 cells, tension along y,
  40 increments, output every 4th (used by ``python/tests/test_Result.py``);
* ``tests/integration/resources/plastic_detect_changes/17grains_tensionX_phenopowerlawbcc.hdf5``
  -- cI phenopowerlaw with non-Schmid terms, 10x16x12 cells, 150 increments
  (used by ``tests/integration/test_plastic_detect_changes.py``).

Writes ``fcc_6grains6x7x8.npz`` and ``bcc_17grains.npz`` plus the embedded
setup files (.vti, material/load YAML) next to this script.
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference"
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.dirname(os.path.abspath(__file__))


def dataset(path, name, dtype="<f8"):
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        tmp = f.name
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", tmp, path], check=True,
                   stdout=subprocess.DEVNULL)
    a = np.fromfile(tmp, dtype)
    os.unlink(tmp)
    return a


def text_dataset(path, name):
    out = subprocess.run([H5DUMP, "-d", name, path], check=True, capture_output=True, text=True).stdout
    m = re.search(r'\(0\): "(.*)"\s*\n\s*\}\s*\n\s*ATTRIBUTE', out, re.S) or re.search(r'\(0\): "(.*)"\s*\n\s*\}', out, re.S)
    body = m.group(1)
    lines = body.split("\n")
    # h5dump indents continuation lines by 11 spaces
    return "\n".join([lines[0]] + [ln[11:] if ln.startswith(" " * 11) else ln.lstrip() for ln in lines[1:]])


def attr_t(path, inc):
    out = subprocess.run([H5DUMP, "-a", f"/increment_{inc}/t/s", path], capture_output=True, text=True).stdout
    m = re.search(r"\(0\): ([-+0-9.eE]+)", out)
    return float(m.group(1)) if m else np.nan


def cell_to_entries(path, name):
    out = subprocess.run([H5DUMP, "-d", name, path], check=True, capture_output=True, text=True).stdout
    return np.array([int(x) for x in re.findall(r'"\s*,\s*(-?\d+)\s*\n', out)], np.int64)


def fcc():
    path = f"{REF}/python/tests/resources/Result/6grains6x7x8_tensionY_singlePhase.hdf5"
    d = {}
    incs = [0, 4, 8, 12, 16, 20, 24, 28, 32, 36, 40]
    d["increments"] = np.array(incs)
    N = 6 * 7 * 8
    for inc in incs:
        base = f"/increment_{inc}/phase/pheno_fcc/mechanical/"
        for k, w in (("F", 9), ("P", 9), ("F_e", 9), ("F_p", 9), ("L_p", 9), ("O", 4), ("xi_sl", 12)):
            d[f"{k}_{inc}"] = dataset(path, base + k).reshape(N, w)
        d[f"u_p_{inc}"] = dataset(path, f"/increment_{inc}/geometry/u_p").reshape(N, 3)
        d[f"u_n_{inc}"] = dataset(path, f"/increment_{inc}/geometry/u_n").reshape(-1, 3)
    d["cells"] = dataset(path, "/geometry/cells", "<i4").astype(np.int64)
    d["size"] = dataset(path, "/geometry/size")
    d["origin"] = dataset(path, "/geometry/origin")
    d["cell_to_phase_entry"] = cell_to_entries(path, "/cell_to/phase")
    d["cell_to_homogenization_entry"] = cell_to_entries(path, "/cell_to/homogenization")
    np.savez_compressed(os.path.join(OUT, "fcc_6grains6x7x8.npz"), **d)
    for name, fname in (("/setup/6grains6x7x8.vti", "fcc_6grains6x7x8.vti"),
                        ("/setup/singlePhase.yaml", "fcc_singlePhase.yaml"),
                        ("/setup/tensionY.yaml", "fcc_tensionY.yaml")):
        with open(os.path.join(OUT, fname), "w") as f:
            f.write(text_dataset(path, name))


def bcc():
    base = f"{REF}/tests/integration/resources/plastic_detect_changes"
    path = f"{base}/17grains_tensionX_phenopowerlawbcc.hdf5"
    out = subprocess.run(["/opt/conda/bin/h5ls", "-r", path], capture_output=True, text=True).stdout
    incs = sorted({int(m) for m in re.findall(r"/increment_(\d+)/", out)})
    d = {"increments": np.array(incs)}
    N = 10 * 16 * 12
    ph = re.findall(r"/increment_\d+/phase/([^/\s]+)/mechanical/xi_sl", out)[0]
    for inc in incs:
        m = f"/increment_{inc}/phase/{ph}/mechanical/"
        d[f"xi_sl_{inc}"] = dataset(path, m + "xi_sl").reshape(N, -1)
        d[f"gamma_sl_{inc}"] = dataset(path, m + "gamma_sl").reshape(N, -1)
        d[f"u_p_{inc}"] = dataset(path, f"/increment_{inc}/geometry/u_p").reshape(N, 3)
        d[f"u_n_{inc}"] = dataset(path, f"/increment_{inc}/geometry/u_n").reshape(-1, 3)
    np.savez_compressed(os.path.join(OUT, "bcc_17grains.npz"), **d)
    for src, dst in (("17grains.vti", "bcc_17grains.vti"), ("tensionX.yaml", "bcc_tensionX.yaml"),
                     ("numerics.yaml", "bcc_numerics.yaml")):
        with open(f"{base}/{src}") as f, open(os.path.join(OUT, dst), "w") as g:
            g.write(f.read())
    # only the phenopowerlawbcc phase and the materials of material.yaml
    import yaml
    with open(f"{base}/material.yaml") as f:
        mat = yaml.safe_load(f)
    keep = {"homogenization": mat["homogenization"], "phase": {ph: mat["phase"][ph]},
            "material": [{**m, "constituents": [{**c, "phase": ph} for c in m["constituents"]]}
                         for m in mat["material"]]}
    with open(os.path.join(OUT, "bcc_material.yaml"), "w") as g:
        yaml.safe_dump(keep, g, sort_keys=False)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference tree not available")
    fcc()
    bcc()
