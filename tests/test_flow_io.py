"""libdgpu_io.so on the CPU: the flow file of WriteHDF5 / ReadHDF5
(for/phdf5.f90:918-1456), the g2gf / gf2g maps (for/tools.f90:255-315) and the
statistics series (for/phdf5.f90:7-278), checked against the numpy restatement
(oracle/flow_io.py), closed forms, and the file as the HDF5 C library itself
reads it (tests/h5raw.py).  No reference file exists to compare with; the layout
is pinned by the reference's own calls (dataset names, dimension order, chunking,
attribute types) as read back here."""
import re

import numpy as np
import pytest

import h5raw
from oracle import flow_io as ofi

pytestmark = pytest.mark.skipif(not h5raw.available(), reason="HDF5 C library not in the image")

NX, NY, NZ = 8, 7, 6


def fields(seed, nth=2):
    rng = np.random.default_rng(seed)
    shp = (NY + 2, NZ + 2, NX + 2)
    return {"u1": rng.standard_normal(shp), "u2": rng.standard_normal(shp), "u3": rng.standard_normal(shp),
            "th": [rng.standard_normal(shp) for _ in range(nth)], "p": rng.standard_normal(shp)}


def io():
    from diablo3_amd import flow_io
    return flow_io


def test_g2gf_matches_restatement_and_closed_forms():
    fio = io()
    v = fields(1)["u2"]
    a = v.copy()
    fio.g2gf(a, NX, NY, NZ)
    assert np.array_equal(a, ofi.g2gf(v, NX, NY, NZ))
    # closed form: a profile linear in the row index moves by half a row inside,
    # and the wall rows keep the first and last interior values
    lin = np.zeros_like(v)
    lin[:, :NZ, :NX] = (3.0 + 0.25 * np.arange(NY + 2))[:, None, None]
    b = lin.copy()
    fio.g2gf(b, NX, NY, NZ)
    rows = np.arange(NY + 2)
    assert np.allclose(b[2:NY, :NZ, :NX], (3.0 + 0.25 * (rows[2:NY] + 0.5))[:, None, None], rtol=0, atol=1e-15)
    assert np.all(b[1, :NZ, :NX] == lin[2, :NZ, :NX]) and np.all(b[NY, :NZ, :NX] == lin[NY, :NZ, :NX])
    # x, z padding and row 0 are not touched
    assert np.array_equal(b[:, NZ:, :], lin[:, NZ:, :]) and np.array_equal(b[0], lin[0])


def test_gf2g_inverts_g2gf_and_sets_the_wall_rows():
    fio = io()
    v = fields(2)["u2"]
    a = v.copy()
    fio.g2gf(a, NX, NY, NZ)
    fio.gf2g(a, NX, NY, NZ)
    assert np.array_equal(a, ofi.gf2g(ofi.g2gf(v, NX, NY, NZ), NX, NY, NZ))
    # interior rows come back (to rounding), the wall rows carry zero mass flux
    assert np.allclose(a[2:NY + 1, :NZ, :NX], v[2:NY + 1, :NZ, :NX], rtol=0, atol=1e-14)
    assert np.array_equal(a[1], -a[2]) and np.array_equal(a[NY + 1], -a[NY])


def test_write_read_round_trip(tmp_path):
    fio = io()
    f = fields(3)
    u2f = f["u2"].copy()
    fio.g2gf(u2f, NX, NY, NZ)
    path = str(tmp_path / "out.h5")
    t = {"time": 1.25, "save_flow_time": 2.5, "save_stats_time": 0.125, "save_movie_time": 3.0, "time_step": 42}
    fio.write_physical(path, NX, NY, NZ, f["u1"], u2f, f["u3"], f["th"], f["p"], t)
    r = fio.read_physical(path, NX, NY, NZ, 2)
    for k, src in (("u1", f["u1"]), ("u2f", u2f), ("u3", f["u3"]), ("p", f["p"])):
        assert np.array_equal(r[k], ofi.swapyz(ofi.swapzy(src, NX, NY, NZ), NX, NY, NZ)), k
    for n in range(2):
        assert np.array_equal(r["th"][n], ofi.swapyz(ofi.swapzy(f["th"][n], NX, NY, NZ), NX, NY, NZ))
    assert r["times"] == t


def test_file_layout_as_hdf5_reads_it(tmp_path):
    fio = io()
    f = fields(4)
    u2f = f["u2"].copy()
    fio.g2gf(u2f, NX, NY, NZ)
    path = str(tmp_path / "layout.h5")
    t = {"time": 0.5, "save_flow_time": 1.0, "save_stats_time": 0.75, "save_movie_time": 2.0, "time_step": 7}
    fio.write_physical(path, NX, NY, NZ, f["u1"], u2f, f["u3"], f["th"], f["p"], t)
    h = h5raw.File(path)
    try:
        res, same, dims = h.attr("/", "Resolution", "i32")
        assert list(res) == [NX, NY, NZ] and same and dims == (3,)
        date, _, ddims = h.attr("/", "Date", "str")
        assert ddims == (20,) and re.fullmatch(rb"\d\d-[A-Z]{3}-\d{4} \d\d:\d\d:\d\d", date), date
        for name in ("Time", "Save_Flow_Time", "Save_Stats_Time", "Save_Movie_Time"):
            v, same, d = h.attr("/Timestep", name, "f64")
            key = {"Time": "time"}.get(name, name.lower())
            assert same and d == () and v[0] == t[key], name
        v, same, d = h.attr("/Timestep", "Time_Step", "i32")
        assert same and d == () and v[0] == 7
        # (Nx, Nyp, Nzp) Fortran dimensions = C order (Nz, Ny, Nx); one y plane per chunk
        for name, src in (("U", f["u1"]), ("W", u2f), ("V", f["u3"]), ("TH1", f["th"][0]), ("TH2", f["th"][1]),
                          ("P", f["p"])):
            data, chunk, f64 = h.dataset("/Timestep/" + name)
            assert data.shape == (NZ, NY, NX) and chunk == (NZ, 1, NX) and f64, name
            # element (x+1, y, z+1) of the Fortran dataset = field(x, z, y)
            assert np.array_equal(data, ofi.swapzy(src, NX, NY, NZ)), name
            assert data[2, 0, 3] == src[1, 2, 3]
        assert not h.exists("/Timestep/TH3")
    finally:
        h.close()


def test_read_checks_resolution_and_masks_scalars(tmp_path):
    fio = io()
    f = fields(5)
    path = str(tmp_path / "m.h5")
    fio.write_physical(path, NX, NY, NZ, f["u1"], f["u2"], f["u3"], f["th"], None, {"time_step": 1})
    with pytest.raises(RuntimeError, match="different resolution"):
        fio.read_physical(path, NX + 2, NY, NZ, 2)
    r = fio.read_physical(path, NX, NY, NZ, 2, th_mask=0b01)
    assert r["p"] is None
    assert np.array_equal(r["th"][0], ofi.swapyz(ofi.swapzy(f["th"][0], NX, NY, NZ), NX, NY, NZ))
    assert not r["th"][1].any()  # scalar 2 not read: its array untouched (zeros here)


def test_statistics_series(tmp_path):
    fio = io()
    path = str(tmp_path / "mean.h5")
    profs = [np.linspace(0, 1, NY) * (k + 1) for k in range(3)]
    for k, pr in enumerate(profs):
        fio.append_profile(path, "ume", pr, 0.1 * (k + 1))
        fio.append_real(path, "time", 0.1 * (k + 1))
    h = h5raw.File(path)
    try:
        n, same, _ = h.attr("/ume", "SAMPLES", "i32")
        assert n[0] == 3 and same
        for k in range(3):
            data, _, f64 = h.dataset(f"/ume/{k + 1:04d}")
            assert f64 and np.array_equal(data, profs[k])
            tv, _, _ = h.attr(f"/ume/{k + 1:04d}", "Time", "f64")
            assert tv[0] == 0.1 * (k + 1)
            v, _, _ = h.dataset(f"/time/{k + 1:04d}")
            assert v.shape == (1,) and v[0] == 0.1 * (k + 1)
        assert h.attr("/time", "SAMPLES", "i32")[0][0] == 3
    finally:
        h.close()


def test_mean_h5_sample(tmp_path):
    """One save_stats sample in mean.h5 with the reference's group names (v and w
    swapped in the names, for/diagnostics.f90:140-235, 453-477) and series
    counters; the z* groups and BPE (1838-1850)."""
    from diablo3_amd import stats_io
    from diablo3_amd.stats import PROFILES, PER_SCALAR
    ny, nth, nb = 9, 2, 12
    rng = np.random.default_rng(8)
    prof = {k: rng.standard_normal(ny + 2) for k in PROFILES}
    prof.update({f"{b}{n + 1:02d}": rng.standard_normal(ny + 2) for n in range(nth) for b in PER_SCALAR})
    prof["u1z_x0"] = 0.375
    z = {k: rng.standard_normal(nb) for k in stats_io.ZSTAR_GROUPS}
    z.update({f"{b}{n + 1:02d}": rng.standard_normal(nb) for n in range(nth) for b in ("thw_zstar", "chi_zstar")})
    z.update({"th1PDF": rng.random(nb), "th1bin": np.linspace(-1, 1, nb + 1), "BPE": -0.125})
    gyf = np.linspace(-1, 1, ny + 2)
    path = str(tmp_path / "mean.h5")
    stats_io.write_mean(path, 0.5, prof, gyf, ny, nth, z)
    h = h5raw.File(path)
    try:
        J = slice(1, ny + 1)
        assert np.array_equal(h.dataset("/wme/0001")[0], prof["vme"][J])
        assert np.array_equal(h.dataset("/vrms/0001")[0], prof["wrms"][J])
        assert np.array_equal(h.dataset("/ww_dwdz/0001")[0], prof["vv_dvdy"][J])
        assert np.array_equal(h.dataset("/thw02_m/0001")[0], prof["thv_m02"][J])
        assert np.array_equal(h.dataset("/dthdz01/0001")[0], prof["dthdy01"][J])
        assert np.array_equal(h.dataset("/gzf/0001")[0], gyf[J])
        assert h.dataset("/time/0001")[0][0] == 0.5 and h.dataset("/u1z_x0/0001")[0][0] == 0.375
        assert h.dataset("/BPE/0001")[0][0] == -0.125
        assert np.array_equal(h.dataset("/th1bin/0001")[0], z["th1bin"][:nb])
        assert np.array_equal(h.dataset("/thw_zstar/0002")[0], z["thw_zstar02"])
        assert h.attr("/thw_zstar", "SAMPLES", "i32")[0][0] == nth
        assert h.attr("/ume", "SAMPLES", "i32")[0][0] == 1
        assert h.attr("/ume/0001", "Time", "f64")[0][0] == 0.5
    finally:
        h.close()


def test_plane_series(tmp_path):
    """WriteHDF5_XYplane / XZplane / ZYplane: var2d(n1, n2) stored with C
    dimensions (n2, n1), one sample per call, SAMPLES and Time as for profiles;
    write_movie writes the groups save_stats_chan(movie) writes."""
    fio = io()
    path = str(tmp_path / "movie.h5")
    a = np.arange(NY * NX, dtype=np.float64).reshape(NY, NX)   # [Nyp][Nx]: var2d(Nx, Nyp)
    b = np.arange(NZ * NX, dtype=np.float64).reshape(NZ, NX) * 0.5
    fio.append_plane(path, "u_xz", a, 0.25)
    fio.append_plane(path, "u_xz", 2 * a, 0.5)
    fio.append_plane(path, "u_xy", b, 0.5)
    h = h5raw.File(path)
    try:
        assert h.attr("/u_xz", "SAMPLES", "i32")[0][0] == 2
        d, _, f64 = h.dataset("/u_xz/0002")
        assert f64 and d.shape == (NY, NX) and np.array_equal(d, 2 * a)
        assert h.attr("/u_xz/0001", "Time", "f64")[0][0] == 0.25
        assert np.array_equal(h.dataset("/u_xy/0001")[0], b)
    finally:
        h.close()
    # the groups of one movie sample
    from diablo3_amd import configs, stats_io
    cfg = configs.front_les(NX, NY, NZ, 5000.0, 2e-4, 1e-2)
    cfg.use_les = False
    planes = {k: np.full((NY, NX), float(i)) for i, k in enumerate(stats_io.XZ_PLANES)}
    slices = {}
    for f in stats_io.movie_fields(1):
        slices[f + "_xz"] = np.zeros((NY, NX))
        slices[f + "_yz"] = np.zeros((NY, NZ))
    mv, mx = str(tmp_path / "mv.h5"), str(tmp_path / "mean_xz.h5")
    stats_io.write_movie(mv, mx, 1.0, planes, slices, cfg)
    h = h5raw.File(mv)
    try:
        assert h.exists("/th1_yz") and h.exists("/omegaZ_xz") and not h.exists("/nu_t_xz") and not h.exists("/u_xy")
        assert h.dataset("/epsilon_yz/0001")[0].shape == (NY, NZ)
    finally:
        h.close()
    h = h5raw.File(mx)
    try:
        for k in stats_io.XZ_PLANES:
            assert h.exists("/" + k), k   # finite front, Nz > 1: every plane
        assert h.dataset("/chi_xz/0001")[0][0, 0] == stats_io.XZ_PLANES.index("chi_xz")
    finally:
        h.close()


def test_grid_file_layout_and_metrics(tmp_path):
    """grid.h5 as create_grid.m writes it (/grids/y, Ny + 1 doubles G(2..Ny+2)) and
    ReadGridHDF5 reads it: read back by the HDF5 library independently of the
    writer, and the metrics built from the file's values are the problem's."""
    from diablo3_amd import configs
    from diablo3_amd.grid import create_grid, metrics
    from diablo3_amd.problem import Problem
    fio = io()
    cfg = configs.dns_64()
    gy = create_grid(cfg.ny, cfg.ly, cfg.grid_cs, cfg.grid_type)
    path = str(tmp_path / "grid.h5")
    fio.write_grid(path, gy)
    h = h5raw.File(path)
    try:
        data, chunk, f64 = h.dataset("/grids/y")
    finally:
        h.close()
    assert data.shape == (cfg.ny + 1,) and f64 and chunk is None
    assert np.array_equal(data, gy)
    assert np.array_equal(fio.read_grid(path), gy)
    # the run's metrics are built from the file's values (for/domain.f90:276-300, 420-500)
    p = Problem(cfg)
    met = metrics(fio.read_grid(path))
    for k in ("gy", "gyf", "dy", "dyf"):
        assert np.array_equal(met[k], p.met[k]), k
    with pytest.raises(RuntimeError):
        fio.read_grid(str(tmp_path / "missing.h5"))
