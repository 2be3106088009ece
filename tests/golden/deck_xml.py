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
    """InputParameters fields of a fixed-source deck, as plain Python values."""
    root = ET.parse(path).getroot()
    s, aq, m = root.find("solver"), root.find("angular_quad"), root.find("mesh")
    G = int(s.get("num_groups"))
    p = dict(
        mode=s.get("solve_type"), scheme=1 if s.get("scheme") == "tw_dd" else 0, G=G,
        fs_tol=float(s.get("fs_tol")), mg_tol=float(s.get("mg_tol")),
        fs_max_it=int(s.get("fs_max_it")), mg_max_it=int(s.get("mg_max_it")),
        dsa=s.get("dsa") == "true", np=int(aq.get("np")), na=int(aq.get("na")),
        dims=int(m.get("dims")), refinement=int(m.get("refinement")))
    r = p["refinement"] + 1
    for ax in "xyz"[:p["dims"]]:
        p["d" + ax] = _vals(m.find("d" + ax).get("vals"))
        p["i" + ax] = [v * r for v in _vals(m.find("i" + ax).get("vals"), int)]
    p["blocks"] = _vals(m.find("blocks").get("vals"), int)
    mats = {}
    for mat in root.find("materials").findall("mat"):
        bid = int(mat.get("block"))
        d = {"src": None}
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
        mats[bid] = d
    p["materials"] = mats
    return root, p
