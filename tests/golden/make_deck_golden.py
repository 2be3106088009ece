"""Writes the deck fixtures under tests/golden/decks/ by running the reference's
own compiled mesh and solver units (oracle/_ref/ref_deck, built from
/root/reference by ``make -C oracle ref``) on the reference's input decks.
Run in the build container only:

    python tests/golden/make_deck_golden.py [deck ...]

The decks are read here with Python's xml.etree (an XML reader independent of
the product's own) into the InputParameters fields the reference's parser
produces (src/InputParameters.cpp:91-524: ix *= refinement + 1, scatter
matrices dest-major, transposed when transpose="true").  Per deck the script
stores

  decks/<deck>.xml  the deck's elements re-serialised (comments dropped), as
                    input data for the GPU tests;
  decks/<deck>.npz  the reference's scalar flux (all of it, or a seeded sample
                    with per-group sums for large meshes), its success flag and
                    SI/MG iteration counts, and the SHA-256 and line count of
                    every text file the reference wrote (<base>_dims.txt,
                    _blocks.txt, _meshx/y/z.txt, _g<g>_flux.txt).

For 1D decks only the mesh is built (the reference's 1D solver needs Eigen, which
is absent): the mesh files are pinned, the flux is left to the oracle.

Eigenvalue decks (solve_type="eigen") run the reference's solveEigenvalue; its
log is printed at 17 significant digits, from which k and the power-iteration
count are stored (k_eff, pi_iterations).  Decks whose name ends in "_coarse" are
the reference deck with fewer cells per region (COARSE below) so that the
reference's CPU solver finishes in seconds; the deck's physics is unchanged.
"""
from __future__ import annotations

import hashlib
import json
import re
import subprocess
import sys
import tempfile
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent))
from deck_xml import read_deck  # noqa: E402

ROOT = Path(__file__).resolve().parents[2]
REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "decks"
HARNESS = ROOT / "oracle" / "_ref" / "ref_deck"

# deck name -> path under the reference tree; all fixed-source decks it holds
# that the reference's 2D/3D solver (or 1D mesh) runs in a few minutes here.
DECKS = {
    "sample_params": "sample_params.xml",
    "2D_watanabe": "cases/2D_watanabe.xml",
    "2D_test": "cases/2D_test.xml",
    "2D_assembly_fs": "cases/2D_assembly_fs.xml",
    "3D_kobayashi_1": "cases/3D_kobayashi_1.xml",
    "3D_kobayashi_3": "cases/3D_kobayashi_3.xml",
    "1D_reed": "cases/1D_reed.xml",
    "1D_NPRE_555_mp2": "cases/1D_NPRE_555_mp2.xml",
    "1D_NPRE_555_mp3_3g_src": "cases/1D_NPRE_555_mp3_3g_src.xml",
    "1D_NPRE_555_mp3_3g_eig": "cases/1D_NPRE_555_mp3_3g_eig.xml",
    "1D_NPRE_555_mp3_3g_src_dsa": "cases/1D_NPRE_555_mp3_3g_src_dsa.xml",
    "1D_NPRE_555_mp3_3g_eig_dsa": "cases/1D_NPRE_555_mp3_3g_eig_dsa.xml",
    "1D_NPRE_555_mp1": "cases/1D_NPRE_555_mp1.xml",
    "2D_assembly_tr_short": "cases/2D_assembly_transient/2D_assembly_tr.xml",
    "2D_assembly_tr_remove": "cases/2D_assembly_transient/2D_assembly_tr.xml",
    "2D_assembly_eig": "cases/2D_assembly_eig.xml",
    "2D_assembly_eig_coarse": "cases/2D_assembly_eig.xml",
    "3D_assembly_fs_coarse": "cases/3D_assembly_fs.xml",
    "3D_assembly_eig_coarse": "cases/3D_assembly_eig.xml",
}
# mesh attribute -> new value for the "_coarse" decks
_ASSEMBLY_XY = "8, " + "2, " * 11 + "8"
COARSE = {
    "2D_assembly_eig_coarse": {"refinement": ("mesh", "0")},
    # the transient deck's first 5 of 1000 steps (same dt = 0.01) on the unrefined mesh
    "2D_assembly_tr_short": {"refinement": ("mesh", "0"), "solver": ("t1", "0.05"),
                             "solver:t_steps": ("t_steps", "5")},
    # 4 steps with the step source removed after the third (t = 0, .01, .02 on)
    "2D_assembly_tr_remove": {"refinement": ("mesh", "0"), "solver": ("t1", "0.04"),
                              "solver:t_steps": ("t_steps", "4"),
                              "source_step": ("remove_at", "0.02")},
    "3D_assembly_fs_coarse": {"ix": ("vals", _ASSEMBLY_XY), "iy": ("vals", _ASSEMBLY_XY),
                              "iz": ("vals", "4, 1, 4")},
    "3D_assembly_eig_coarse": {"ix": ("vals", _ASSEMBLY_XY), "iy": ("vals", _ASSEMBLY_XY),
                               "iz": ("vals", "8")},
}
FULL_LIMIT = 40000  # store the whole flux up to this many values
THREADS = 8


def write_spec(p, path, run):
    lines = [f"{p['dims']} {p['G']} {p['scheme']} {p['np']} {p['na']} {p['fs_max_it']} "
             f"{p['mg_max_it']} {THREADS} {run}",
             f"{p['fs_tol']!r} {p['mg_tol']!r} {p['k_tol']!r}"]
    if run == 3:
        lines.append(f"{p['t0']!r} {p['t1']!r} {p['t_steps']}")
    for ax in "xyz"[:p["dims"]]:
        d, n = p["d" + ax], p["i" + ax]
        lines.append(f"{len(d)} " + " ".join(repr(v) for v in d) + " " + " ".join(map(str, n)))
    lines.append(f"{len(p['blocks'])} " + " ".join(map(str, p["blocks"])))
    lines.append(str(len(p["materials"])))
    for bid, d in sorted(p["materials"].items()):
        flags = [d["src"] is not None, d["chi"] is not None, d["prod"] is not None]
        nd = d["nd"]
        vals = d["sigt"] + d["sigs"] + (d["src"] or []) + (d["chi"] or []) + \
            (d["prod"] or []) + (d["beta"] + d["chi_d"] + d["lambda"] if nd else []) + \
            (d["inv_v"] or [])
        lines.append(f"{bid} " + " ".join(str(int(f)) for f in flags) +
                     f" {nd} {int(d['inv_v'] is not None)} " + " ".join(repr(v) for v in vals))
    lines.append(str(len(p["steps"])))
    for bid, st in sorted(p["steps"].items()):
        lines.append(f"{bid} {st['type']} {st['insert']!r} {st['remove']!r} " +
                     " ".join(repr(v) for v in st["src"]))
    Path(path).write_text("\n".join(lines) + "\n")


def _files(base, dims, G):
    names = ["dims", "blocks", "meshx"] + (["meshy"] if dims > 1 else []) + \
        (["meshz"] if dims > 2 else []) + [f"g{g}_flux" for g in range(G)]
    return {n: Path(f"{base}_{n}.txt") for n in names}


def _coarsen(name, src, tmp):
    """The reference deck with the COARSE mesh attributes replaced."""
    tree = ET.parse(src)
    root = tree.getroot()
    mesh = root.find("mesh")
    for tag, (attr, val) in COARSE[name].items():
        tag = tag.split(":")[0]
        if tag == "refinement":
            mesh.set("refinement", val)
        elif tag == "solver":
            root.find("solver").set(attr, val)
        elif tag == "source_step":
            root.find("transients").find("source_step").set(attr, val)
        else:
            mesh.find(tag).set(attr, val)
    out = Path(tmp) / f"{name}.src.xml"
    tree.write(out, encoding="unicode")
    return out


def make(name, tmp):
    src = REF / DECKS[name]
    if name in COARSE:
        src = _coarsen(name, src, tmp)
    root, p = read_deck(src)
    OUT.mkdir(exist_ok=True)
    ET.ElementTree(root).write(OUT / f"{name}.xml", encoding="unicode")
    run = 0 if p["dims"] == 1 else {"steady": 1, "eigen": 2, "transient": 3}[p["mode"]]
    spec, base = Path(tmp) / f"{name}.spec", str(Path(tmp) / name)
    write_spec(p, spec, run)
    subprocess.run([str(HARNESS), str(spec), base], check=True)
    raw = Path(base + ".bin").read_bytes()
    ok = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    nphi = p["G"] * int(np.prod([sum(p["i" + ax]) for ax in "xyz"[:p["dims"]]]))
    phi = np.frombuffer(raw[4:4 + 8 * nphi], dtype=np.float64)
    files = _files(base, p["dims"], p["G"])
    if run == 3:
        # the transient's initial dump holds the zero initial condition; its
        # per-step outputs are <base>_t<step>_g<g>_flux.txt and _d<d>_dnps.txt
        rest = raw[4 + 8 * nphi:]
        nd = int(np.frombuffer(rest[:4], dtype=np.int32)[0])
        dnps = np.frombuffer(rest[4:], dtype=np.float64).reshape(nd, -1)
        for step in range(p["t_steps"]):
            for g in range(p["G"]):
                files[f"t{step}_g{g}_flux"] = Path(f"{base}_t{step}_g{g}_flux.txt")
            for d in range(nd):
                files[f"t{step}_d{d}_dnps"] = Path(f"{base}_t{step}_d{d}_dnps.txt")
    if not run:
        files = {k: v for k, v in files.items() if not k.endswith("_flux")}
    arrays = dict(ref_success=ok, dims=p["dims"], groups=p["G"])
    meta = {"success": ok, "files": {}}
    for k, f in files.items():
        data = f.read_bytes()
        arrays[f"sha256_{k}"] = hashlib.sha256(data).hexdigest()
        arrays[f"lines_{k}"] = data.count(b"\n")
        meta["files"][k] = arrays[f"lines_{k}"]
    if run:
        N = phi.size // p["G"]
        phi = phi.reshape(p["G"], N)
        # The reference writes each value with operator<< at precision 6, which
        # is printf's %.6g: check that equivalence once here, on its own files.
        last = f"t{p['t_steps'] - 1}_" if run == 3 else ""
        for g in range(p["G"]):
            text = files[f"{last}g{g}_flux"].read_text().splitlines()
            assert text == ["%.6g" % v for v in phi[g]], (name, g)
        if run == 3:
            arrays["dnps"] = dnps
            arrays["t_steps"] = p["t_steps"]
            # the reference's per-step text files, for a line-by-line comparison
            for k, f in files.items():
                if k.startswith("t"):
                    arrays[f"text_{k}"] = np.array(f.read_text().splitlines())
        if phi.size <= FULL_LIMIT:
            arrays["phi"] = phi
        else:
            rng = np.random.default_rng(7)
            idx = np.sort(rng.choice(N, size=4096, replace=False)).astype(np.int64)
            arrays["sample_idx"] = idx
            arrays["phi_sample"] = phi[:, idx]
            arrays["phi_group_l2"] = np.sqrt((phi ** 2).sum(axis=1))
        log = Path(base + ".log").read_text()
        sys.path.insert(0, str(Path(__file__).resolve().parent))
        from make_golden import parse_log  # noqa: E402
        meta["log"] = parse_log(log)
        arrays["si_iterations"] = np.array([a for a, _, _ in meta["log"]["si"]], dtype=np.int64)
        meta["log"]["si"] = meta["log"]["si"][-p["G"]:]  # the manifest keeps the last pass
        if run == 2:
            k = re.findall(r"Converged value of k = (\S+)", log)
            pi = re.findall(r"PI (?:converged|failed to converge) after (\d+) iterations", log)
            arrays["pi_iterations"] = int(pi[-1])
            if k:
                arrays["k_eff"] = float(k[-1])
                meta["k_eff"] = float(k[-1])
            meta["pi_iterations"] = int(pi[-1])
    np.savez_compressed(OUT / f"{name}.npz", **arrays)
    return meta


# Decks too large for the reference to solve here, re-serialised as benchmark
# inputs only (tests/golden/bench_decks/<name>.xml, no fixture).
BENCH_DECKS = {
    "3D_assembly_eig": "cases/3D_assembly_eig.xml",
    "3D_assembly_fs": "cases/3D_assembly_fs.xml",
    "2D_assembly_tr": "cases/2D_assembly_transient/2D_assembly_tr.xml",
}


def write_bench_decks():
    out = OUT.parent / "bench_decks"
    out.mkdir(exist_ok=True)
    for name, rel in BENCH_DECKS.items():
        root, _ = read_deck(REF / rel)
        ET.ElementTree(root).write(out / f"{name}.xml", encoding="unicode")


def main(argv):
    if argv == ["--bench-decks"]:
        write_bench_decks()
        return
    names = argv or list(DECKS)
    OUT.mkdir(exist_ok=True)
    man_path = OUT / "manifest.json"
    manifest = json.loads(man_path.read_text()) if man_path.exists() else {}
    with tempfile.TemporaryDirectory() as tmp:
        for n in names:
            print("deck", n, flush=True)
            manifest[n] = make(n, tmp)
            print("  ", manifest[n]["success"], flush=True)
    man_path.write_text(json.dumps(manifest, indent=1, sort_keys=True))


if __name__ == "__main__":
    main(sys.argv[1:])
