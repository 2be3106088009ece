"""Host-side checks of the C ABI (no GPU needed): the library loads and
exports every entry point include/doma_c.h declares, device entry points fail
with DOMA_ERR_DEVICE instead of falling back to the CPU, and the seeded media
generator follows its documented rules."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest
import torch

import doma_amd
from doma_amd import capi, workloads

HEADER = Path(__file__).resolve().parents[1] / "include" / "doma_c.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(doma_\w+)\s*\(", text)))


def test_header_symbols_exported():
    names = _declared()
    assert len(names) >= 20
    lib = capi.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding knows every one of them
    bound = {n for n, _, _ in capi.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_version_string():
    assert capi.lib().doma_version().decode().startswith("doma_amd")


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_create_without_gpu_is_an_error():
    p = workloads.random_blocks(16, 1, 1, 1)
    with pytest.raises(doma_amd.DomaError) as e:
        doma_amd.Context(p)
    assert e.value.code == capi.DOMA_ERR_DEVICE


def _spec(**kw):
    base = dict(n=16, block=4, groups=3, sigt_log=0, sigt_lo=0.5, sigt_hi=2.0, c_lo=0.3,
                c_hi=0.9, scatter_rule=capi.SCATTER_DOWN1, source_rule=capi.SOURCE_CENTRAL,
                src_lo=6, src_hi=10, source_fraction=0.0, cell_width=1.0, seed=123)
    base.update(kw)
    return capi.MediaSpec(**base)


def test_media_deterministic_and_seeded():
    a = capi.media_generate(_spec())
    b = capi.media_generate(_spec())
    c = capi.media_generate(_spec(seed=124))
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1], c[1])


def test_media_down1_rule():
    mat, sigt, sigs, src, has = capi.media_generate(_spec())
    M, G = sigt.shape
    assert M == 2 * (16 // 4) ** 3
    assert (sigt >= 0.5).all() and (sigt <= 2.0).all()
    for m in range(M):
        for g in range(G):
            c_in = sigs[m, g, g] / (0.7 * sigt[m, g])
            assert 0.3 <= c_in <= 0.9
            if g + 1 < G:
                assert sigs[m, g + 1, g] == pytest.approx(0.3 * c_in * sigt[m, g], rel=1e-14)
        # nothing else is non-zero: no up-scatter, no skipped groups
        mask = np.eye(G, dtype=bool) | np.eye(G, k=-1, dtype=bool)
        assert (sigs[m][~mask] == 0).all()
    # the unit source sits in group 0 on [6, 10)^3 only
    n = 16
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inside = ((i >= 6) & (i < 10) & (j >= 6) & (j < 10) & (k >= 6) & (k < 10)).ravel()
    q = np.where(has[mat][:, None], src[mat], 0.0)
    assert (q[inside, 0] == 1.0).all()
    assert (q[inside, 1:] == 0).all()
    assert (q[~inside] == 0).all()


def test_media_full_rule_rows():
    mat, sigt, sigs, src, has = capi.media_generate(
        _spec(groups=8, sigt_log=1, sigt_lo=0.2, sigt_hi=5.0, c_lo=0.5, c_hi=0.95,
              scatter_rule=capi.SCATTER_FULL, source_rule=capi.SOURCE_BLOCKS,
              source_fraction=0.05))
    M, G = sigt.shape
    assert (sigt >= 0.2 - 1e-12).all() and (sigt <= 5.0 + 1e-12).all()
    assert has.sum() >= 1
    # column sums (out of source group g) equal c_g sigma_t,g with c in [0.5, 0.95]
    out = sigs.sum(axis=1)
    c = out / sigt
    assert (c >= 0.5 - 1e-12).all() and (c <= 0.95 + 1e-12).all()
    # up-scatter only from the last group into the one above it
    for m in range(M):
        up = np.triu(sigs[m], k=1)
        nz = np.argwhere(up != 0)
        assert all((d, s) == (G - 2, G - 1) for d, s in nz)


def test_media_rejects_bad_spec():
    with pytest.raises(ValueError):
        capi.media_generate(_spec(block=5))
