"""The C-ABI library loads and exports every entry point include/mcbiome.h declares (no compute)."""
import ctypes as C
import os
import re

import mcbiome

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mcbiome.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(mc[gb]_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 20
    lib = C.CDLL(mcbiome.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_helpers_without_device():
    names = mcbiome.biome_names()
    assert len(names) == 67 and names == sorted(names)
    L = mcbiome.lib()
    assert L.mcg_version_from_name(b"26.3") == 2 and L.mcg_version_from_name(b"1.21") == -1
    assert L.mcg_biome_id(b"minecraft:plains") == names.index("plains")


def test_no_cpu_fallback_without_device():
    """Without a HIP device every compute entry fails loudly instead of computing on the CPU."""
    if mcbiome.device_count() > 0:
        return
    try:
        mcbiome.Engine("26.3", "overworld")
    except mcbiome.EngineError as e:
        assert "device" in str(e)
    else:
        raise AssertionError("engine opened without a device")
