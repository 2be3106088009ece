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
        t0=float(s.get("t0", "0")), t1=float(s.get("t1", "1")),
        t_steps=int(s.get("t_steps", "1")), ics=s.get("ics", "zero"),
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
        d = {"src": None, "chi": None, "prod": None, "beta": [], "nd": nd, "chi_d": [],
             "lambda": [], "inv_v": None}
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
            elif t == "chi_d":
                d["chi_d"] = v
            elif t == "lambda_d":
                d["lambda"] = v
            elif t == "inv_vel":
                d["inv_v"] = v
        mats[bid] = d
    p["materials"] = mats
    steps = {}
    tr = root.find("transients")
    if tr is not None and p["mode"] == "transient":
        for node in tr.findall("source_step"):
            ins, rem = node.get("insert_at"), node.get("remove_at")
            kind = 0 if ins and rem else (1 if ins else 2)
            steps[int(node.get("block"))] = dict(
                type=kind, insert=float(ins) if ins else 0.0, remove=float(rem) if rem else 0.0,
                src=_vals(node.get("mgxs")))
    p["steps"] = steps
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


def problem_from_deck(p):
    """A doma_amd Problem (with fission, kinetics and step sources where the deck
    has them) for the parsed deck `p`: cells k-major, i fastest; block of a cell =
    blocks[(rz * nry + ry) * nrx + rx] (BrickMesh3D.cpp:86-102); materials are the
    block ids in ascending order.  Quadratures: GC(2na, 2np) in 2D/3D, GL(2np) in 1D."""
    from doma_amd import capi, workloads

    dims, G = p["dims"], p["G"]
    widths, regions = [], []
    for ax in "xyz":
        if ax in "xyz"[:dims]:
            widths.append(np.concatenate([np.full(n, d / n) for d, n in zip(p["d" + ax], p["i" + ax])]))
            regions.append(np.repeat(np.arange(len(p["i" + ax])), p["i" + ax]))
        else:
            widths.append(np.ones(1))
            regions.append(np.zeros(1, dtype=int))
    nrx, nry = len(p["ix"]), len(p["iy"]) if dims > 1 else 1
    rz, ry, rx = np.meshgrid(regions[2], regions[1], regions[0], indexing="ij")
    blocks = np.array(p["blocks"])[((rz * nry + ry) * nrx + rx).reshape(-1)]
    ids = sorted(p["materials"])
    mat = np.searchsorted(ids, blocks).astype(np.int32)
    mats = [p["materials"][i] for i in ids]
    sigt = np.array([m["sigt"] for m in mats])
    sigs = np.array([m["sigs"] for m in mats]).reshape(len(ids), G, G)
    src = np.array([m["src"] or [0.0] * G for m in mats])
    has = np.array([m["src"] is not None for m in mats], dtype=np.uint8)
    if dims == 1:
        mu, eta, xi, w = workloads.quadrature_1d(p["np"])
    elif dims == 2:
        mu, eta, xi, w = workloads.quadrature_2d(p["np"], p["na"])
    else:
        mu, eta, xi, w = workloads.quadrature_3d(p["np"], p["na"])
    prob = capi.Problem(dims, widths[0], widths[1], widths[2], G, p["scheme"], mat, sigt, sigs,
                        src, has, mu, eta, xi, w)
    f = fission_arrays(p, ids)
    if f is not None:
        prob.fission = capi.Fission(*f)
    if p["mode"] == "transient":
        K = capi.MAX_DELAYED
        M = len(ids)
        inv_v = np.array([m["inv_v"] for m in mats])
        nd = np.array([m["nd"] for m in mats], dtype=np.int32)
        chi_d, beta, lam = np.zeros((M, G, K)), np.zeros((M, G, K)), np.zeros((M, K))
        for r, m in enumerate(mats):
            n = m["nd"]
            if n:
                chi_d[r, :, :n] = np.array(m["chi_d"]).reshape(G, n)
                beta[r, :, :n] = np.array(m["beta"]).reshape(G, n)
                lam[r, :n] = m["lambda"]
        prob.kinetics = capi.Kinetics(inv_v, nd, chi_d, beta, lam)
        if p["steps"]:
            ty = np.full(M, capi.STEP_NONE, dtype=np.int32)
            t_in, t_out, s_src = np.zeros(M), np.zeros(M), np.zeros((M, G))
            for bid, st in p["steps"].items():
                r = ids.index(bid)
                ty[r], t_in[r], t_out[r] = st["type"], st["insert"], st["remove"]
                s_src[r] = st["src"]
            prob.steps = capi.StepSources(ty, t_in, t_out, s_src)
    return prob
