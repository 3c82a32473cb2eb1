"""Checkpoint files on the host: the round trip of checkpoint.f90:36-47 and
particle_initialization.f90:34-82, and the reference's stops when a
checkpoint or the redshift list cannot be used (zip-format mismatch,
particle_initialization.f90:41-45; an empty z_checkpoint.txt,
initialize.f90:55-61; truncated files)."""
import os

import numpy as np
import pytest

from cube_nufluid_amd._lib import SimHeader
from cube_nufluid_amd.checkpoint import BoxState, output_name, read_checkpoint, read_z_list, write_checkpoint


def small_state(nc=4, n=50, izipx=2, izipv=1, seed=3):
    rng = np.random.default_rng(seed)
    sim = SimHeader()
    sim.nplocal, sim.izipx, sim.izipv = n, izipx, izipv
    sim.a, sim.t, sim.tau = 0.5, 0.25, -2.0
    cnt = np.bincount(rng.integers(0, nc ** 3, n), minlength=nc ** 3).astype(np.int32)
    xd, vd = (np.int8 if izipx == 1 else np.int16), (np.int8 if izipv == 1 else np.int16)
    xp = rng.integers(np.iinfo(xd).min, np.iinfo(xd).max, (n, 3)).astype(xd)
    vp = rng.integers(np.iinfo(vd).min, np.iinfo(vd).max, (n, 3)).astype(vd)
    pid = rng.permutation(n).astype(np.int32) + 1
    vfield = rng.standard_normal((nc ** 3, 3)).astype(np.float32)
    return BoxState(sim, xp, vp, pid, cnt, vfield)


def test_round_trip_is_exact(tmp_path):
    st = small_state()
    write_checkpoint(st, str(tmp_path) + "/", 1.5)
    back = read_checkpoint(str(tmp_path) + "/", 1.5, 4, izipx=2, izipv=1)
    assert back.sim.to_bytes() == st.sim.to_bytes()
    for a, b in ((back.xp, st.xp), (back.vp, st.vp), (back.pid, st.pid), (back.rhoc, st.rhoc),
                 (back.vfield, st.vfield)):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    # file names follow z2str: '1.500_xp_1.bin' in image1/
    assert os.path.exists(os.path.join(str(tmp_path), "image1", "1.500_xp_1.bin"))


@pytest.mark.parametrize("want", [(1, 1), (2, 2)])
def test_zip_format_mismatch_stops(tmp_path, want):
    """A checkpoint written with izipx/izipv = 2/1 is refused when the run
    expects another pair, as the reference stops with 'zip format
    incompatable' instead of reading particles of the wrong width."""
    st = small_state(izipx=2, izipv=1)
    write_checkpoint(st, str(tmp_path) + "/", 0.0)
    with pytest.raises(ValueError, match="zip format"):
        read_checkpoint(str(tmp_path) + "/", 0.0, 4, izipx=want[0], izipv=want[1])


def test_truncated_and_inconsistent_files_stop(tmp_path):
    st = small_state()
    d = str(tmp_path) + "/"
    write_checkpoint(st, d, 0.0)
    fn = output_name(d, 0.0, "xp")
    with open(fn, "r+b") as f:
        f.truncate(os.path.getsize(fn) - 6)
    with pytest.raises(ValueError, match="truncated"):
        read_checkpoint(d, 0.0, 4)
    # a cell count that does not add up to nplocal
    st2 = small_state()
    st2.rhoc = st2.rhoc.copy()
    st2.rhoc[0] += 1
    write_checkpoint(st2, d, 1.0)
    with pytest.raises(ValueError, match="nplocal"):
        read_checkpoint(d, 1.0, 4)
    # a header of the wrong size
    with open(output_name(d, 1.0, "info"), "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(ValueError, match="224 bytes"):
        read_checkpoint(d, 1.0, 4)


def test_z_list_parse_empty_and_missing(tmp_path):
    """A blank record reads 0 under (f8.4), so only a file without records
    is empty; text the format cannot read stops the run."""
    p = tmp_path / "z_checkpoint.txt"
    p.write_text("  49.0000\n  10.0000\n\n   0.0000\n")
    assert read_z_list(str(p)) == [49.0, 10.0, 0.0, 0.0]
    b = tmp_path / "blank.txt"
    b.write_text("\n\n")
    assert read_z_list(str(b)) == [0.0, 0.0]
    e = tmp_path / "empty.txt"
    e.write_text("")
    with pytest.raises(ValueError, match="empty"):
        read_z_list(str(e))
    bad = tmp_path / "bad.txt"
    bad.write_text("49.0\n0.5 extra\n")
    with pytest.raises(ValueError, match="f8.4"):
        read_z_list(str(bad))
    with pytest.raises(FileNotFoundError):
        read_z_list(str(tmp_path / "missing.txt"))


def test_host_reads_both_redshift_lists_from_files(tmp_path):
    """Cube takes z_checkpoint.txt and z_halofind.txt as paths, read as the
    reference reads them (initialize.f90:55-71)."""
    from cube_nufluid_amd.driver import Cube
    zc, zh = tmp_path / "z_checkpoint.txt", tmp_path / "z_halofind.txt"
    zc.write_text("  49.0000\n   0.0000\n")
    zh.write_text("   1.0000\n   0.5000\n")
    host = Cube.__new__(Cube)
    host._init_time(np.float32(0.02), str(zc), zh)
    assert host.z_checkpoints == [49.0, 0.0] and host.z_halofind == [1.0, 0.5]
    assert host.cur_checkpoint == 2 and host.cur_halofind == 3
    zh.write_text("")
    with pytest.raises(ValueError, match="z_halofind.txt empty"):
        Cube.__new__(Cube)._init_time(np.float32(0.02), str(zc), str(zh))
