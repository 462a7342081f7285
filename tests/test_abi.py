"""The C ABI library loads and exports every function include/empi_hip.h
declares (no device call is made here)."""
import os
import re
import subprocess

import pytest

import empi_amd as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "empi_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(empi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared()
    assert "empi_plan_run" in names and "empi_ws_find_best" in names
    assert sorted(E.EXPORTS) == names


def test_library_exports_all_symbols():
    lib = E.lib()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(empi_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared()) <= exported


def test_strerror_codes_match_reference_strings():
    lib = E.lib()
    for code, name in E.STATUS_NAMES.items():
        assert lib.empi_strerror(code).decode() == name
    assert lib.empi_strerror(99).decode() == "internalError"


def test_host_side_argument_checks():
    with pytest.raises(E.EmpiError) as ei:
        E.dictionary(128.0, 1280, 1.5)
    assert ei.value.status == 4
    scales, entries = E.dictionary(128.0, 1280, 0.01)
    assert entries == 5008073 and len(scales) == 34


def test_plan_create_rejects_bad_items_before_any_device_call():
    """empi_plan_create validates every item (the settings checks of
    src/empi.cpp:59-112) and the kernels' size limit before touching a
    device, so these run on a host without a GPU."""
    bad = [
        E.ItemSpec(0, 1280, 128.0, E.SMP, 10, 99.0, 0.01),      # no channels
        E.ItemSpec(1, 1, 128.0, E.SMP, 10, 99.0, 0.01),         # one sample
        E.ItemSpec(1, 1280, 0.0, E.SMP, 10, 99.0, 0.01),        # fs
        E.ItemSpec(1, 1280, 128.0, E.SMP, 0, 99.0, 0.01),       # iterationMax
        E.ItemSpec(1, 1280, 128.0, 7, 10, 99.0, 0.01),          # variant
        E.ItemSpec(1, 1280, 128.0, E.SMP, 10, 101.0, 0.01),     # energyPercent > 100
        E.ItemSpec(1, 1280, 128.0, E.SMP, 10, 0.0, 0.01),       # residualEnergy 1
        E.ItemSpec(1, 1280, 128.0, E.SMP, 10, 99.0, 1.5),       # energyError outside (0, 1)
    ]
    for spec in bad:
        with pytest.raises(E.EmpiError) as ei:
            E.Plan([spec])
        assert ei.value.status == 4, spec
    # beyond 2^24 samples the kernels' 24-bit index arithmetic is not exact
    with pytest.raises(E.EmpiError) as ei:
        E.Plan([E.ItemSpec(1, (1 << 24) + 1, 256.0, E.SMP, 10, 99.0, 0.01)])
    assert ei.value.status == 8
