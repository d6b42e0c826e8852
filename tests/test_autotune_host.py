"""Host side of the build-time tuner: the tuned header it reads and rewrites, and
the build that compiles it in (no GPU needed)."""
import os
import shutil

from fmd_amd import autotune as A
from fmd_amd import build as B


def test_tuned_header_defines_every_parameter():
    vals = A.read_tuned()
    assert {m for m, _, _ in A.PARAMS} <= set(vals)
    assert vals["FMDX_EAM_LDS"] == 0   # the LDS-table passes are measured slower: off
    for m, cands, _ in A.PARAMS:
        assert vals[m] > 0
        assert len(cands) >= 2


def test_write_then_read_round_trip(tmp_path):
    p = str(tmp_path / "fmdx_tuned.h")
    shutil.copy(A.TUNED_H, p)
    vals = A.read_tuned(p)
    new = dict(vals, FMDX_TILE64=256, FMDX_EAM_NP32_1=6)
    A.write_tuned(new, [("FMDX_TILE64", 256, [(192, 4.1), (224, 3.9), (256, 3.8)])], p)
    assert A.read_tuned(p) == new
    text = open(p).read()
    assert "256: 3.800 ms" in text
    # a second run replaces the table instead of stacking it
    A.write_tuned(vals, [("FMDX_TILE32", 384, [(320, 3.0), (384, 2.9)])], p)
    assert A.read_tuned(p) == vals
    text = open(p).read()
    assert text.count("Last tuning run") == 1 and "256: 3.800 ms" not in text


def test_force_kernels_take_the_tuned_header():
    src = open(os.path.join(B.DEV_SRC, A.TUNED_FILE)).read()
    assert '#include "fmdx_tuned.h"' in src
    for m, _, _ in A.PARAMS:
        assert m in src, m


def test_builder_key_names_every_parameter():
    key = A.Builder.key(A.read_tuned())
    assert key.count("_") >= len(A.PARAMS) - 1
    assert "tile64" in key and "np32_1" in key


def test_every_workload_is_named_and_candidates_hold_the_incumbent():
    vals = A.read_tuned()
    for m, cands, work in A.PARAMS:
        assert work in A.WORKLOAD_NAME
        assert vals[m] in cands, m


def test_potential_energy_guard():
    assert A.agrees(-1.0000001, -1.0, 1e-6)
    assert not A.agrees(-1.00001, -1.0, 1e-6)
    assert not A.agrees(float("nan"), -1.0, 1e-6)


def test_lds_table_variant_fits_one_workgroup_per_cu(tmp_path):
    """The FP32 EAM passes with the spline table in LDS (FMDX_EAM_LDS, off by default) still
    compile, and 16 waves with 184-candidate tiles leave room for a 5002-knot table (80 KB)
    beside the pair parameters, within the 160 KB of one CU, at no more than 128 VGPRs (the
    force pass with the pair energy keeps a few words of its live state in scratch at that
    limit; the variant is a measured alternative, not the path in use)."""
    import re
    import subprocess
    out = str(tmp_path / "k_force_lds.s")
    subprocess.run([B.HIPCC] + B.HIP_FLAGS + ["-DFMDX_EAM_LDS=16", "-DFMDX_EAM_LDS_TILE=184", "--cuda-device-only",
                    "-S", os.path.join(B.DEV_SRC, A.TUNED_FILE), "-o", out], check=True, capture_output=True)
    text = open(out).read()
    meta = text[text.index("amdhsa.kernels:"):]
    found = 0
    for blk in re.split(r"\n  - ", meta)[1:]:
        name = re.search(r"\.name:\s+(\S+)", blk)
        if not name or "k_cell_lds" not in name.group(1):
            continue
        name = name.group(1)
        get = lambda k: int(re.search(r"\." + k + r":\s+(\d+)", blk).group(1))
        found += 1
        assert get("max_flat_workgroup_size") == 16 * 64, name
        assert get("group_segment_fixed_size") + 48 + 5002 * 16 <= 160 * 1024, name
        assert get("vgpr_count") <= 128, name
        assert get("private_segment_fixed_size") <= 64, name
    assert found >= 3   # density pass, force pass with and without the pair energy


def test_each_pair_once_variant_compiles(tmp_path):
    """The FP32 LJ and Morse passes that evaluate each pair once (FMDX_N3, off by default:
    measured slower than the two-sided passes) still compile, within the registers and LDS
    that let 15 workgroups share a CU, and take their partner sums with LDS atomics."""
    import re
    import subprocess
    out = str(tmp_path / "k_force_n3.s")
    subprocess.run([B.HIPCC] + B.HIP_FLAGS + ["-DFMDX_N3=1", "--cuda-device-only", "-S",
                    os.path.join(B.DEV_SRC, A.TUNED_FILE), "-o", out], check=True, capture_output=True)
    text = open(out).read()
    meta = text[text.index("amdhsa.kernels:"):]
    found = 0
    for blk in re.split(r"\n  - ", meta)[1:]:
        name = re.search(r"\.name:\s+(\S+)", blk)
        # the pair kernels with each pair once; not the batch's looping launch for the cells
        # of a compacted list beyond its grid (last template argument true), which rarely runs
        if not name or not re.search(r"(k|kb)_cellIfLi0ELi[23]ELb[01]ELb1E(Ev|Lb0E)", name.group(1)):
            continue
        get = lambda k: int(re.search(r"\." + k + r":\s+(\d+)", blk).group(1))
        found += 1
        assert get("vgpr_count") <= 128, name.group(1)
        assert get("group_segment_fixed_size") * 15 <= 160 * 1024, name.group(1)
    assert found >= 8   # LJ and Morse, with and without the energy, single system and batch
    assert "ds_add_f32" in text
