"""The ctypes mirrors of the C ABI structs (nimbus_amd/lib.py) have the C
layout of include/nimbus_hip.h: a C program compiled against the header
prints sizeof and every field's offsetof, which must equal ctypes'."""
import ctypes
import os
import subprocess

import pytest

from nimbus_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "nimbus_grid": L.Grid,
    "nimbus_step_config": L.StepConfig,
    "nimbus_base_state": L.BaseState,
    "nimbus_host_state": L.HostState,
    "nimbus_member_desc": L.MemberDesc,
    "nimbus_acoustic_substep_fused_dispatch": L.AcousticFusedDispatch,
    "nimbus_acoustic_substeps_batched_dispatch": L.AcousticBatchedDispatch,
    "nimbus_slow_tendencies_dispatch": L.SlowTendenciesDispatch,
    "nimbus_slow_rk2_dispatch": L.SlowRk2Dispatch,
    "nimbus_advect_scalars_dispatch": L.AdvectScalarsDispatch,
    "nimbus_kessler_dispatch": L.KesslerDispatch,
    "nimbus_cartesian_cgrid_slow_tendencies_dispatch": L.CgridSlowTendenciesDispatch,
    "nimbus_slow_apply_dispatch": L.SlowApplyDispatch,
    "nimbus_state_relaxation_dispatch": L.StateRelaxationDispatch,
    "nimbus_microphysics_apply_dispatch": L.MicrophysicsApplyDispatch,
    "nimbus_kessler_pointwise_dispatch": L.KesslerPointwiseDispatch,
    "nimbus_kessler_sedimentation_dispatch": L.KesslerSedimentationDispatch,
    "nimbus_conservative_cartesian_cgrid_transport_dispatch": L.ConservativeTransportDispatch,
    "nimbus_diffusion_batch_dispatch": L.DiffusionBatchDispatch,
}


def test_ctypes_structs_match_the_c_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "nimbus_hip.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = str(tmp_path / "layout")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), str(src), "-o",
                        exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            parts = line.split()
            got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
