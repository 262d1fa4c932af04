"""Member state I/O on the device: NPY field export and snapshot/restore
(resume is bitwise identical to never having stopped)."""
import os

import numpy as np
import pytest

from nimbus_amd import lib as L
from tests.helpers import small_case

pytestmark = pytest.mark.gpu


def _member(ctx, precision, mode, seed=0):
    g, base, st = small_case(nx=20, ny=18, nz=14, cloudy=True, wind_noise=2.0,
                             dtype=np.float32 if precision == L.PRECISION_MIXED else np.float64)
    cfg = L.step_config(2.5, 6, acoustic_mode=mode, sponge={"z_d": 4000.0, "alpha": 0.05, "w_cells": 4,
                                                              "w_alpha": 0.2, "lateral_cells": 3,
                                                              "lateral_alpha": 0.05}, sgs=L.sgs_preset())
    m = ctx.member(g, cfg, base, precision)
    m.upload(st)
    return m


@pytest.mark.parametrize("tiling", [None, (32, 2, 1)])
@pytest.mark.parametrize("precision,mode", [(L.PRECISION_MIXED, L.ACOUSTIC_HEVI),
                                            (L.PRECISION_FP64, L.ACOUSTIC_EXPLICIT)])
def test_snapshot_resume_is_bitwise(hip_ctx, tmp_path, precision, mode, tiling):
    """Snapshot after two steps, three more steps on the original and on the
    restored member (from memory and from a file): identical results. The
    launch geometry travels with the snapshot (the default linear order, or
    32 x 8 column tiles)."""
    print(f"case: snapshot precision {precision} mode {mode} tiling {tiling}", flush=True)
    a = _member(hip_ctx, precision, mode)
    try:
        if tiling is not None:
            a.set_tiling(*tiling)
        hip_ctx.step([a], 2)
        blob = a.snapshot()
        a.save(str(tmp_path / "m.snap"))
        hip_ctx.step([a], 3)
        want, want_diag = a.download(), a.diag()
    finally:
        a.destroy()
    for src in (blob, str(tmp_path / "m.snap")):
        b = hip_ctx.restore(src)
        try:
            assert b.diag()["steps"] == 2
            hip_ctx.step([b], 3)
            got, got_diag = b.download(), b.diag()
        finally:
            b.destroy()
        for f in L.FIELDS:
            assert np.array_equal(got[f], want[f]), f
        assert np.array_equal(got["rain_mm"], want["rain_mm"])
        assert got_diag == want_diag


def test_corrupted_or_truncated_snapshot_is_refused(hip_ctx):
    a = _member(hip_ctx, L.PRECISION_MIXED, L.ACOUSTIC_HEVI)
    try:
        blob = bytearray(a.snapshot())
    finally:
        a.destroy()
    bad = bytearray(blob)
    bad[len(bad) // 2] ^= 0x40
    with pytest.raises(L.NimbusError, match="checksum"):
        hip_ctx.restore(bytes(bad))
    with pytest.raises(L.NimbusError, match="length"):
        hip_ctx.restore(bytes(blob[:-8]))
    bad = bytearray(blob)
    bad[0:8] = b"NOTASNAP"
    with pytest.raises(L.NimbusError, match="not a member snapshot"):
        hip_ctx.restore(bytes(bad))


@pytest.mark.parametrize("precision", [L.PRECISION_MIXED, L.PRECISION_FP64])
def test_export_npy_matches_download(hip_ctx, tmp_path, precision):
    print(f"case: export precision {precision}", flush=True)
    m = _member(hip_ctx, precision, L.ACOUSTIC_HEVI)
    try:
        hip_ctx.step([m], 2)
        st = m.download()
        m.export_npy(str(tmp_path))
    finally:
        m.destroy()
    names = {"u": "u", "v": "v", "w": "w", "rho": "rho", "p": "p", "theta": "theta", "qv": "qv", "qc": "qc",
             "qr": "qr"}
    for f, n in names.items():
        a = np.load(str(tmp_path / f"{n}.npy"))
        assert a.dtype == np.float32 and a.shape == st[f].shape
        assert np.array_equal(a, st[f].astype(np.float32)), f
    r = np.load(str(tmp_path / "accumulated_rainfall.npy"))
    assert r.dtype == np.float32 and r.shape == st["rain_mm"].shape
    assert np.array_equal(r, st["rain_mm"].astype(np.float32))
    assert sorted(os.listdir(str(tmp_path))) == sorted([n + ".npy" for n in names.values()] +
                                                       ["accumulated_rainfall.npy"])
