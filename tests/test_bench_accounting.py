"""The bench's roofline accounting covers every kernel class the library
times: each profile slot named in cube_api.hip (kSlotNames) either has
algorithmic bytes in bench.algorithmic_bytes or is listed here as timed
without them, so a renamed or added slot cannot drop out of the roofline
choice unnoticed."""
import importlib.util
import os
import re

from cube_nufluid_amd.params import CONFIGS

ROOT = os.path.join(os.path.dirname(__file__), "..")
# slots timed without an algorithmic-bytes figure: on-chip sorts, the small
# coarse FFT, and the fluid (reported by the c4 leg with its own bytes)
UNACCOUNTED = {"drift_sort_big", "fft_coarse", "fluid", "fluid_density", "fluid_gravity"}


def slot_names():
    with open(os.path.join(ROOT, "cube_nufluid_amd", "csrc", "cube_api.hip")) as f:
        src = f.read()
    m = re.search(r"kSlotNames\[PS_NSLOT\]\s*=\s*\{([^}]*)\}", src)
    assert m, "kSlotNames not found"
    return re.findall(r'"([a-z0-9_]+)"', m.group(1))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_every_slot_is_accounted():
    bench = load_bench()
    cfg = CONFIGS["C3"]
    names = slot_names()
    assert len(names) == len(set(names))
    for n in names:
        b = bench.algorithmic_bytes(n, cfg, 1 << 30)
        assert (b is not None) != (n in UNACCOUNTED), n

