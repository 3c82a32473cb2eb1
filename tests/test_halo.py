"""Runtime halo finder (halofind.f90:18-598).

CPU tests pin the restatement (oracle/halo_oracle.c) against brute force:
the radius-sorted offset table, the cell-centre CIC density, the peak and
mesh-mass stage, and for isolated and overlapping clumps the overdensity
membership, the density criterion and the means.  GPU tests compare
cube_halofind with the restatement: counts, catalogue order, halo_pid
stream and hid bit for bit, float fields within FP32 rounding.
"""
import numpy as np
import pytest

import oracle_lib as ol
from cube_nufluid_amd.halo import HaloParams, HaloCatalog, read_halo_file, write_halo_files, HALO_FIELDS
from cube_nufluid_amd.ic import encode_particles, rank_to_cell
from cube_nufluid_amd.params import BoxConfig

NCELL = 4


def clumps_state(cfg, specs, n_bg, seed, vel_scale=0.5):
    """Gaussian clumps (centre in coarse units, n, sigma in coarse units)
    over a uniform background."""
    rng = np.random.default_rng(seed)
    parts = [(np.asarray(c) + rng.normal(0, s, size=(n, 3))) for c, n, s in specs]
    parts.append(rng.uniform(0, cfg.nc, size=(n_bg, 3)))
    pos = np.concatenate(parts) % cfg.nc
    vel = rng.normal(0, vel_scale, size=pos.shape)
    return encode_particles(cfg, pos, vel, np.arange(1, len(pos) + 1, dtype=np.int32), np.float32(0.5))


def fine_positions(cfg, st):
    """Global fine-cell positions (float64) of the particles of st."""
    cells = rank_to_cell(np.repeat(np.arange(cfg.nc ** 3), st.rhoc), cfg.nt, cfg.nnt)
    cells = np.stack(cells, axis=1) if isinstance(cells, tuple) else np.asarray(cells)
    u = st.xp.view(np.uint16 if cfg.izipx == 2 else np.uint8).astype(np.float64)
    off = (u + 0.5) / (65536.0 if cfg.izipx == 2 else 256.0)
    return NCELL * (cells.astype(np.float64) + off), cells


def test_offset_table_is_radius_sorted_in_generation_order():
    R = 5
    t = ol.halo_offset_table(R)
    rows = []
    for i in range(-R, R + 1):
        for j in range(-R, R + 1):
            for k in range(-R, R + 1):
                r2 = i * i + j * j + k * k
                if r2 <= R * R:
                    rows.append((r2, len(rows), i, j, k))
    rows.sort()
    want = np.array([(i, j, k, r2) for r2, _, i, j, k in rows], dtype=np.int32)
    np.testing.assert_array_equal(t, want)
    # the shell test of halofind.f90:192 compares neighbouring entries: r2 is
    # non-decreasing and every r2 <= R^2 appears
    assert np.all(np.diff(t[:, 3]) >= 0) and len(t) == len(rows)


@pytest.mark.parametrize("nnt", [1, 2])
def test_density_is_cell_centre_cic(nnt):
    cfg = BoxConfig(nc=8 * nnt, nnt=nnt, ncb=6)
    st = clumps_state(cfg, [((3.3, 4.1, 5.7), 300, 0.4)], 2000, seed=4)
    hp = HaloParams().for_box(cfg)
    rho = ol.halo_density(cfg, st, hp)
    x, _ = fine_positions(cfg, st)
    nf = cfg.nf
    m = float(st.sim.mass_p_cdm)
    want = np.zeros((nf, nf, nf))
    p = x - 0.5
    i1 = np.floor(p).astype(np.int64)
    w1 = (i1 + 1) - p
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = ((w1[:, 0] if dx == 0 else 1 - w1[:, 0]) * (w1[:, 1] if dy == 0 else 1 - w1[:, 1]) *
                     (w1[:, 2] if dz == 0 else 1 - w1[:, 2]))
                np.add.at(want, ((i1[:, 2] + dz) % nf, (i1[:, 1] + dy) % nf, (i1[:, 0] + dx) % nf), w * m)
    assert np.max(np.abs(rho - want)) < 1e-5 * np.max(want)
    assert rho.sum(dtype=np.float64) == pytest.approx(m * st.sim.nplocal, rel=1e-6)


def test_peak_stage_matches_brute_force():
    cfg = BoxConfig(nc=16, nnt=2, ncb=6)
    specs = [((3.2, 4.4, 12.1), 1500, 0.25), ((7.9, 8.1, 8.0), 900, 0.2), ((13.0, 2.0, 6.0), 2500, 0.35)]
    st = clumps_state(cfg, specs, 12000, seed=11)
    hp = HaloParams().for_box(cfg)
    rho = ol.halo_density(cfg, st, hp)
    nf, nft, nfb, nnt = cfg.nf, NCELL * cfg.nt, NCELL * cfg.ncb, cfg.nnt
    tab = ol.halo_offset_table(hp.nc_halo_max)
    mx = rho.copy()
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                mx = np.maximum(mx, np.roll(rho, (-dz, -dy, -dx), axis=(0, 1, 2)))
    peaks = np.argwhere((mx == rho) & (rho > hp.den_peak_cutoff))  # (z, y, x)
    m_p = np.float32(st.sim.mass_p_cdm)
    thr = m_p * np.float32(hp.min_halo_particles) / np.float32(2.0)
    nvalid = 0
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                T0 = np.array([tx, ty, tz]) * nft
                for z, y, x in peaks:
                    g = np.array([x, y, z])
                    # tile-local 1-based node in -3..nft+4, through the periodic images
                    loc = None
                    for img in np.ndindex(3, 3, 3):
                        l = g + (np.array(img) - 1) * nf - T0 + 1
                        if np.all(l >= 1 - NCELL) and np.all(l <= nft + NCELL):
                            if loc is None:
                                loc = []
                            loc.append(l)
                    for l in loc or []:
                        m = np.float32(0)
                        for ii in range(1, len(tab) + 1):
                            dx, dy, dz, r2 = tab[ii - 1]
                            ix, iy, iz = l[0] + dx, l[1] + dy, l[2] + dz
                            if min(ix, iy, iz) <= 1 - nfb or max(ix, iy, iz) >= nft + nfb - 1:
                                continue
                            gg = (np.array([ix, iy, iz]) - 1 + T0) % nf
                            m = np.float32(m + rho[gg[2], gg[1], gg[0]])
                            if ii < len(tab) and tab[ii][3] == r2:
                                continue
                            if ii >= 19 and np.float32(m / np.float32(ii)) < np.float32(hp.den_odc):
                                break
                        nvalid += int(m > thr)
    _, _, _, counts = ol.halofind(cfg, st, hp)
    assert nvalid > 0 and counts[2] == nvalid


def halo_member_check(cfg, st, hp, h, members, excluded=()):
    """Brute force of halofind.f90:375-418 for an isolated halo: the members
    are the i_odc particles nearest to the refined centre (excluding those
    already claimed), and the enclosed density crosses den_odc there."""
    x, cells = fine_positions(cfg, st)
    nf = cfg.nf
    d = x - h[HALO_FIELDS["hpos"]].astype(np.float64)
    d -= nf * np.round(d / nf)
    r = np.sqrt((d * d).sum(1))
    keep = ~np.isin(st.pid, np.asarray(excluded, dtype=np.int32))
    idx = np.nonzero(keep)[0]
    order = idx[np.argsort(r[idx], kind="stable")]
    n = int(h[HALO_FIELDS["mass_odc"]])
    assert n == len(members)
    assert set(st.pid[order[:n]].tolist()) == set(members.tolist())
    rs = r[order]
    m = float(st.sim.mass_p_cdm)
    den = 0.75 / np.pi * np.arange(1, len(rs) + 1) * m / rs ** 3
    assert den[n] < hp.den_odc                       # ii = n+1 falls below
    assert np.all(den[1:n] >= hp.den_odc * (1 - 1e-5))  # ii = 2..n do not
    # x_mean: mean of the members' positions (images nearest the centre)
    xm = np.mod(h[HALO_FIELDS["hpos"]].astype(np.float64) + d[order[:n]].mean(0), nf)
    np.testing.assert_allclose(h[HALO_FIELDS["x_mean"]], xm, atol=2e-4)
    return order[:n]


def test_isolated_clump_membership_and_criterion():
    cfg = BoxConfig(nc=16, nnt=2, ncb=6)
    st = clumps_state(cfg, [((6.3, 9.8, 5.1), 2500, 0.3)], 15000, seed=21)
    hp = HaloParams().for_box(cfg)
    halos, words, hid, counts = ol.halofind(cfg, st, hp, nhid=st.sim.nplocal)
    cat = HaloCatalog(hp.den_odc, halos, words, hid, int(counts[2]), int(counts[3]))
    assert cat.nhalo == 1
    mem = cat.members()[0]
    sel = halo_member_check(cfg, st, hp, halos[0], mem)
    # hid marks exactly the members with halo number 1
    assert np.array_equal(np.nonzero(hid)[0] + 1, np.sort(mem))
    assert np.all(hid[mem - 1] == 1)
    # v_mean, v_disp from the decoded velocities
    sig = st.sim.sigma_vi
    cells_r = np.repeat(np.arange(cfg.nc ** 3), st.rhoc)
    v = np.array([[ol.vdecode(int(st.vp[i, e]), sig, cfg.izipv) + st.vfield[cells_r[i], e] for e in range(3)]
                  for i in sel])
    np.testing.assert_allclose(halos[0, HALO_FIELDS["v_mean"]], v.mean(0), rtol=1e-5, atol=1e-6)
    vd = np.sqrt(((v - v.mean(0)) ** 2).sum() / (len(v) - 1))
    assert halos[0, HALO_FIELDS["v_disp"]] == pytest.approx(vd, rel=1e-5)


def test_overlapping_clumps_claim_disjoint_members():
    """The denser clump is processed first and tags its members; the second
    halo is the overdensity of the remaining particles (halofind.f90:421,
    :276)."""
    cfg = BoxConfig(nc=16, nnt=1, ncb=6)
    specs = [((8.0, 8.0, 8.0), 3000, 0.25), ((8.9, 8.3, 7.8), 1500, 0.3)]
    st = clumps_state(cfg, specs, 15000, seed=5)
    hp = HaloParams().for_box(cfg)
    halos, words, hid, counts = ol.halofind(cfg, st, hp)
    cat = HaloCatalog(hp.den_odc, halos, words, None, int(counts[2]), int(counts[3]))
    assert cat.nhalo == 2
    a, b = cat.members()
    assert len(np.intersect1d(a, b)) == 0
    halo_member_check(cfg, st, hp, halos[0], a)
    halo_member_check(cfg, st, hp, halos[1], b, excluded=a)


def test_halo_files_round_trip(tmp_path):
    cfg = BoxConfig(nc=16, nnt=2, ncb=6)
    st = clumps_state(cfg, [((4.2, 4.2, 4.2), 2000, 0.3), ((11.0, 12.5, 3.0), 1800, 0.25)], 12000, seed=8)
    hp = HaloParams().for_box(cfg)
    halos, words, hid, counts = ol.halofind(cfg, st, hp, nhid=cfg.nf ** 3)
    cat = HaloCatalog(hp.den_odc, halos, words, hid, int(counts[2]), int(counts[3]))
    prefix = str(tmp_path / "image1" / "0.000_")
    write_halo_files(prefix, cat)
    ntot, n, den, recs = read_halo_file(prefix + "halo_1.bin")
    assert ntot == n == cat.nhalo and den == 200.0
    np.testing.assert_array_equal(recs, halos)
    w = np.fromfile(prefix + "halo_pid_1.bin", dtype=np.int32)
    np.testing.assert_array_equal(w, words)
    assert (tmp_path / "image1" / "0.000_hid_1.bin").stat().st_size == 2 * cfg.nf ** 3
    # a file cut short, or holding more records than it declares, is rejected
    raw = open(prefix + "halo_1.bin", "rb").read()
    for bad in (raw[:-4], raw[:-4 * HALO_FIELDS["s_mean"].stop], raw + raw[12:12 + 168], raw[:6]):
        open(prefix + "bad.bin", "wb").write(bad)
        with pytest.raises(ValueError, match="halo file"):
            read_halo_file(prefix + "bad.bin")


# ------------------------------------------------------------------ GPU
def gpu_catalog(cfg, st, hp, nhid=0):
    from cube_nufluid_amd.driver import Cube
    cube = Cube(cfg, st)
    try:
        return cube.halofind(hp, nhid=nhid, max_halos=4096)
    finally:
        cube.close()


def compare(cat, halos, words, hid, nf):
    """Catalogue order, counts, halo_pid stream and hid bit for bit; float
    fields within FP32 rounding of the sums (positions in fine cells)."""
    assert cat.nhalo == len(halos)
    np.testing.assert_array_equal(cat.pid_words, words)
    if hid is not None:
        np.testing.assert_array_equal(cat.hid, hid)
    g, o = cat.halos.astype(np.float64), halos.astype(np.float64)
    np.testing.assert_array_equal(g[:, 3], o[:, 3])                       # mass_odc
    for f in ("hpos", "x_mean", "s_mean", "q_mean"):
        dd = g[:, HALO_FIELDS[f]] - o[:, HALO_FIELDS[f]]
        dd -= nf * np.round(dd / nf)
        assert np.max(np.abs(dd)) <= 2e-4, f
    for f in ("radius_odc", "v_disp", "v_mean", "var_x", "ang_mom", "inertia", "inertia_q"):
        a, b = g[:, HALO_FIELDS[f]], o[:, HALO_FIELDS[f]]
        scale = np.maximum(np.abs(b), 1e-3 * np.max(np.abs(b)) + 1e-30)
        assert np.max(np.abs(a - b) / scale) < 2e-5, f


@pytest.mark.gpu
@pytest.mark.parametrize("nnt", [1, 2, 4])
def test_gpu_halofind_matches_restatement(nnt):
    cfg = BoxConfig(nc=16 if nnt < 4 else 32, nnt=nnt, ncb=6, box=100.0)
    rng = np.random.default_rng(30 + nnt)
    specs = []
    for k in range(14):
        c = rng.uniform(0, cfg.nc, 3)
        if k % 4 == 0:
            c[0] = (cfg.nt * (1 + k % cfg.nnt)) % cfg.nc + rng.uniform(-0.6, 0.6)  # straddles a tile face
        specs.append((c, int(rng.integers(400, 3000)), float(rng.uniform(0.12, 0.35))))
    specs.append((specs[1][0] + np.array([0.7, -0.4, 0.3]), 1200, 0.25))   # overlaps clump 1
    st = clumps_state(cfg, specs, 20000, seed=40 + nnt)
    hp = HaloParams().for_box(cfg)
    nhid = st.sim.nplocal
    halos, words, hid, counts = ol.halofind(cfg, st, hp, nhid=nhid)
    print("oracle: %d halos, %d peaks" % (len(halos), counts[2]), flush=True)
    cat = gpu_catalog(cfg, st, hp, nhid=nhid)
    assert cat.npeaks >= counts[2] and cat.nprocessed == counts[3]
    compare(cat, halos, words, hid, cfg.nf)
    assert len(halos) >= 8


@pytest.mark.gpu
def test_gpu_halofind_is_deterministic():
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=200.0)
    rng = np.random.default_rng(77)
    specs = [(rng.uniform(0, cfg.nc, 3), int(rng.integers(300, 4000)), float(rng.uniform(0.1, 0.4))) for _ in range(60)]
    st = clumps_state(cfg, specs, 100000, seed=78)
    hp = HaloParams().for_box(cfg)
    a = gpu_catalog(cfg, st, hp, nhid=st.sim.nplocal)
    b = gpu_catalog(cfg, st, hp, nhid=st.sim.nplocal)
    print("gpu: %d halos, %d peaks" % (a.nhalo, a.npeaks), flush=True)
    assert a.nhalo > 20
    np.testing.assert_array_equal(a.halos, b.halos)
    np.testing.assert_array_equal(a.pid_words, b.pid_words)
    np.testing.assert_array_equal(a.hid, b.hid)
    halos, words, hid, counts = ol.halofind(cfg, st, hp, nhid=st.sim.nplocal)
    compare(a, halos, words, hid, cfg.nf)
