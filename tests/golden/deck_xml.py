"""Reads a <doma_inp> deck with Python's xml.etree into the InputParameters
fields the reference's parser produces (src/InputParameters.cpp:91-524):
ix *= refinement + 1, scatter matrices dest-major (transposed when
transpose="true"), vectors split on commas with whitespace removed
(InputParameters.cpp:9-49).  Test infrastructure: an XML reader independent of
the product's own, used to check it and to write fixtures.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np


def _vals(text, kind=float):
    s = "".join(text.split())
    return [kind(v) for v in s.split(",")]


def read_deck(path):
    """InputParameters fields of a deck, as plain Python values."""
    root = ET.parse(path).getroot()
    s, aq, m = root.find("solver"), root.find("angular_quad"), root.find("mesh")
    G = int(s.get("num_groups"))
    p = dict(
        mode=s.get("solve_type"), scheme=1 if s.get("scheme") == "tw_dd" else 0, G=G,
        fs_tol=float(s.get("fs_tol")), mg_tol=float(s.get("mg_tol")),
        fs_max_it=int(s.get("fs_max_it")), mg_max_it=int(s.get("mg_max_it")),
        dsa=s.get("dsa") == "true", np=int(aq.get("np")), na=int(aq.get("na")),
        k_tol=float(s.get("k_tol")) if s.get("solve_type") == "eigen" else 0.0,
        dims=int(m.get("dims")), refinement=int(m.get("refinement")))
    r = p["refinement"] + 1
    for ax in "xyz"[:p["dims"]]:
        p["d" + ax] = _vals(m.find("d" + ax).get("vals"))
        p["i" + ax] = [v * r for v in _vals(m.find("i" + ax).get("vals"), int)]
    p["blocks"] = _vals(m.find("blocks").get("vals"), int)
    mats = {}
    for mat in root.find("materials").findall("mat"):
        bid = int(mat.get("block"))
        # delayed groups are read in transient mode only (InputParameters.cpp:291-294)
        nd = int(mat.get("delayed_groups", "0")) if p["mode"] == "transient" else 0
        d = {"src": None, "chi": None, "prod": None, "beta": [], "nd": nd}
        for rx in mat.findall("reaction"):
            t = rx.get("type")
            v = _vals(rx.get("mgxs"))
            if t == "total":
                d["sigt"] = v
            elif t == "scatter":
                a = np.array(v).reshape(G, G)
                if rx.get("transpose", "false") == "true":
                    a = a.T
                d["sigs"] = a.reshape(-1).tolist()
            elif t == "source":
                d["src"] = v
            elif t == "chi_p":
                d["chi"] = v
            elif t == "nu_sigma_f":
                d["prod"] = v
            elif t == "beta":
                d["beta"] = v
        mats[bid] = d
    p["materials"] = mats
    return root, p


def fission_arrays(p, ids):
    """chi, prod, prompt [M, G], has_chi, has_prod [M] of the materials `ids`
    (prompt = 1 - sum_d beta[g][d], TransportSolver3D.cpp:380-386), or None when
    no material carries fission data."""
    G = p["G"]
    mats = [p["materials"][i] for i in ids]
    if not any(m["chi"] or m["prod"] for m in mats):
        return None
    chi = np.array([m["chi"] or [0.0] * G for m in mats])
    prod = np.array([m["prod"] or [0.0] * G for m in mats])
    prompt = np.ones((len(ids), G))
    for r, m in enumerate(mats):
        if m["nd"] > 0:
            b = np.array(m["beta"]).reshape(G, m["nd"])
            for g in range(G):
                gb = 0.0
                for d in range(m["nd"]):
                    gb += b[g, d]
                prompt[r, g] = 1.0 - gb
    has_chi = np.array([m["chi"] is not None for m in mats], dtype=np.uint8)
    has_prod = np.array([m["prod"] is not None for m in mats], dtype=np.uint8)
    return chi, prod, prompt, has_chi, has_prod
