"""NetCDFGrid3DWriter / NetCDFGrid3DReader (ocustorage/gridnetcdf.h:67-183,
gridnetcdf.cpp) on the CPU.

The file format is pinned by an independent implementation: scipy's
netcdf_file reads what the writer produces (dimensions, global attributes,
coordinate variables, every record of every variable, fill values for
records never written), and the reader reads what scipy writes (classic
and 64-bit offset variants).  Staggered variables take the face
coordinates; a grid's layout (ghosts, padding) never reaches the file."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from fields import uniform
from opencurrent_amd import netcdf as N
from opencurrent_amd._types import HostGrid

FILL_F = np.float32(9.9692099683868690e+36)


def padded(n, dtype, seed, g=1, pad=(1, 2, 3)):
    h = HostGrid(*n, g, g, g, *pad, dtype=dtype)
    h.buf[:] = -77  # ghosts and padding: must not leak into the file
    vals = (1000 * (2 * uniform(seed, n) - 1)).astype(dtype)
    h.interior()[:] = vals
    return h, vals


def test_writer_output_reads_back_through_scipy(tmp_path):
    f = tmp_path / "grid.nc"
    nx, ny, nz = 5, 4, 6
    w = N.NetCDFGrid3DWriter(f, nx, ny, nz, .5, .25, 2.0)
    assert w.define_variable("p", N.NC_DOUBLE, N.GS_CENTER_POINT)
    assert w.define_variable("u", N.NC_FLOAT, N.GS_U_FACE)
    assert w.define_variable("w", N.NC_FLOAT, N.GS_W_FACE)
    assert w.define_variable("flag", N.NC_INT, N.GS_CENTER_POINT)
    assert not w.define_variable("p", N.NC_DOUBLE, N.GS_CENTER_POINT)  # duplicate
    p0, pv0 = padded((nx, ny, nz), np.float64, 1)
    u0, uv0 = padded((nx + 1, ny, nz), np.float32, 2)
    w0, wv0 = padded((nx, ny, nz + 1), np.float32, 3, g=2)
    fl, flv = padded((nx, ny, nz), np.int32, 4, pad=(0, 0, 0))
    assert w.add_data("p", p0, 0)
    assert w.add_data("u", u0, 0)
    assert w.add_data("w", w0, 0)
    assert w.add_data("flag", fl, 0)
    assert not w.add_data("u", p0, 0)            # wrong size
    assert w.add_time_level(.75) == 1
    p1, pv1 = padded((nx, ny, nz), np.float64, 5)
    assert w.add_data("p", p1, 1)
    assert w.close()

    with netcdf_file(str(f), "r", mmap=False) as d:
        assert d.version_byte == 2
        assert d.dimensions["Time"] is None
        assert (d.nx, d.ny, d.nz) == (nx, ny, nz)
        assert (np.float32(d.hx), np.float32(d.hy), np.float32(d.hz)) == (np.float32(.5), np.float32(.25),
                                                                          np.float32(2.0))
        assert list(d.variables["Time"][:]) == [0.0, .75]
        # coordinate variables: i*h + offset (the offset is not scaled by h)
        assert np.array_equal(d.variables["XAxisCentered"][:], np.arange(nx, dtype=np.float32) * .5 + .5)
        assert np.array_equal(d.variables["XAxisUFace"][:], np.arange(nx + 1, dtype=np.float32) * .5)
        assert np.array_equal(d.variables["ZAxisWFace"][:], np.arange(nz + 1, dtype=np.float32) * 2)
        assert d.variables["p"].dimensions == ("Time", "XAxisCentered", "YAxisCentered", "ZAxisCentered")
        assert d.variables["u"].dimensions == ("Time", "XAxisUFace", "YAxisCentered", "ZAxisCentered")
        assert np.array_equal(d.variables["p"][0], pv0)
        assert np.array_equal(d.variables["p"][1], pv1)
        assert np.array_equal(d.variables["u"][0], uv0)
        assert np.array_equal(d.variables["w"][0], wv0)
        assert np.array_equal(d.variables["flag"][0], flv)
        # record 1 of u was never written: the library's fill value
        assert np.all(d.variables["u"][1] == FILL_F)
        assert np.all(d.variables["flag"][1] == -2147483647)


@pytest.mark.parametrize("version", [1, 2])
def test_reader_reads_files_scipy_writes(tmp_path, version):
    f = tmp_path / f"scipy{version}.nc"
    nx, ny, nz = 4, 3, 5
    rng = np.random.default_rng(8800)
    with netcdf_file(str(f), "w", version=version) as d:
        d.createDimension("Time", None)
        d.createDimension("XAxisCentered", nx)
        d.createDimension("YAxisCentered", ny)
        d.createDimension("YAxisVFace", ny + 1)
        d.createDimension("ZAxisCentered", nz)
        d.nx, d.ny, d.nz = np.int32(nx), np.int32(ny), np.int32(nz)
        d.hx, d.hy, d.hz = np.float32(.1), np.float32(.2), np.float32(.3)
        t = d.createVariable("Time", "f4", ("Time",))
        t[:] = np.array([0, .5, 1.5], dtype=np.float32)
        a = d.createVariable("temp", "f8", ("Time", "XAxisCentered", "YAxisCentered", "ZAxisCentered"))
        v = d.createVariable("v", "f4", ("Time", "XAxisCentered", "YAxisVFace", "ZAxisCentered"))
        av = rng.standard_normal((3, nx, ny, nz))
        vv = rng.standard_normal((3, nx, ny + 1, nz)).astype(np.float32)
        a[:] = av
        v[:] = vv
    r = N.NetCDFGrid3DReader(f)
    assert (r.nx, r.ny, r.nz, r.num_time_levels) == (nx, ny, nz, 3)
    assert r.hy == np.float32(.2)
    assert [r.get_time(q) for q in range(3)] == [0.0, .5, 1.5]
    assert r.variables() == {"temp": (N.NC_DOUBLE, (nx, ny, nz)), "v": (N.NC_FLOAT, (nx, ny + 1, nz))}
    for q in range(3):
        assert np.array_equal(r.read("temp", q).interior(), av[q])
        g = HostGrid(nx, ny + 1, nz, 2, 2, 2, 3, 0, 1, dtype=np.float32)
        g.buf[:] = -5
        assert r.read_variable("v", g, q)
        assert np.array_equal(g.interior(), vv[q])
        # the layout outside the interior is left alone
        assert np.count_nonzero(g.buf == -5) == g.buf.size - vv[q].size
    assert not r.read_variable("temp", HostGrid(nx, ny, nz, dtype=np.float32), 0)   # wrong type
    assert not r.read_variable("temp", HostGrid(nx, ny, nz), 3)                      # no such record
    r.close()


def test_round_trip_many_records(tmp_path):
    f = tmp_path / "series.nc"
    n = (7, 3, 9)
    w = N.NetCDFGrid3DWriter(f, *n, 1.0, 1.0, 1.0)
    w.define_variable("a", N.NC_FLOAT, N.GS_CENTER_POINT)
    series = []
    for t in range(6):
        if t:
            assert w.add_time_level(.1 * t) == t
        g, v = padded(n, np.float32, 100 + t)
        assert w.add_data("a", g, t)
        series.append(v)
    # a variable defined after data was written (the reference's redef)
    w.define_variable("b", N.NC_DOUBLE, N.GS_V_FACE)
    gb, vb = padded((n[0], n[1] + 1, n[2]), np.float64, 200)
    assert w.add_data("b", gb, 4)
    assert w.close()
    r = N.NetCDFGrid3DReader(f)
    assert r.num_time_levels == 6
    for t in range(6):
        assert np.array_equal(r.read("a", t).interior(), series[t])
    assert np.array_equal(r.read("b", 4).interior(), vb)
    with netcdf_file(str(f), "r", mmap=False) as d:
        assert np.array_equal(d.variables["a"][:], np.stack(series))
        assert np.array_equal(d.variables["b"][4], vb)


def test_errors(tmp_path):
    with pytest.raises(Exception):
        N.NetCDFGrid3DReader(tmp_path / "missing.nc")
    (tmp_path / "junk.nc").write_bytes(b"not a netcdf file at all")
    with pytest.raises(Exception):
        N.NetCDFGrid3DReader(tmp_path / "junk.nc")
    w = N.NetCDFGrid3DWriter(tmp_path / "x.nc", 2, 2, 2)
    assert not w.define_variable("s", 3)   # NC_SHORT is not a grid type
    assert not w.add_data("nothing", HostGrid(2, 2, 2), 0)
    assert w.close()


@pytest.mark.gpu
def test_ns_fields_written_from_the_gpu_read_back(gpu, tmp_path):
    """After GPU NS steps, u, v, w, T and p (staggered as the reference's
    apps write them) go to a file and come back bit-identical."""
    from test_gpu_ns_fused import CASES, initial, make_params
    from opencurrent_amd.solvers import IncompressibleNS3D
    n, vd, tbc, fbc, nu, kappa, scheme, ts = CASES["closed_vx_ab2"]
    p = make_params(n, np.float64, vd, tbc(), fbc(), nu, kappa, scheme, ts)
    ns = IncompressibleNS3D(p, *initial(n, np.float64, 8700))
    f = tmp_path / "ns.nc"
    w = N.NetCDFGrid3DWriter(f, *n, p.hx, p.hy, p.hz)
    names = [("u", N.GS_U_FACE), ("v", N.GS_V_FACE), ("w", N.GS_W_FACE), ("T", N.GS_CENTER_POINT),
             ("p", N.GS_CENTER_POINT)]
    for name, st in names:
        assert w.define_variable(name, N.NC_DOUBLE, st)
    fields = []
    for step in range(2):
        # the step's pressure solve need not meet its tolerance on this small
        # closed box; the fields are what they are, and they go to the file
        ns.advance_one_step(2e-3)
        if step:
            assert w.add_time_level(step * 2e-3) == step
        fs = [ns.field(q) for q in range(5)]
        for (name, _), g in zip(names, fs):
            assert w.add_data(name, g, step)
        fields.append([g.interior().copy() for g in fs])
    assert w.close()
    r = N.NetCDFGrid3DReader(f)
    for step in range(2):
        for q, (name, _) in enumerate(names):
            assert np.array_equal(r.read(name, step).interior(), fields[step][q])
