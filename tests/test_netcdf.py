"""NetCDF classic parcel and field files against an independent
implementation of the format (scipy.io.netcdf_file, which reads and writes
CDF-1 and CDF-2), with the reference's schema (parcel_netcdf.f90:95-309,
856-975; field_netcdf.f90:68-266, 465-561; netcdf_writer.f90:453-594).

No GPU: the handle-free epic_nc_* entry points of the library."""
import math

import numpy as np
import pytest
from scipy.io import netcdf_file

from epic_amd import core

GRID = (8, 6, 4, (-0.5 * math.pi, -1.0, 0.0), (math.pi, 2.0, 1.5))
PARCEL_UNITS = {"x_position": "m", "y_position": "m", "z_position": "m", "volume": "m^3", "buoyancy": "m/s^2",
                "humidity": "1", "x_vorticity": "1/s", "y_vorticity": "1/s", "z_vorticity": "1/s",
                "B11": "m^2", "B12": "m^2", "B13": "m^2", "B22": "m^2", "B23": "m^2"}


def _parcels(n, seed=0):
    rng = np.random.default_rng(seed)
    return {"position": rng.uniform(-1, 1, (n, 3)), "vorticity": rng.normal(size=(n, 3)),
            "B": rng.uniform(0.1, 1, (n, 5)), "volume": rng.uniform(0.5, 1, n), "buoyancy": rng.normal(size=n),
            "humidity": rng.uniform(0, 1, n)}


def _text(v):
    return v.decode() if isinstance(v, bytes) else v


def test_parcel_file_reads_back_with_an_independent_reader(tmp_path):
    p = _parcels(37)
    path = tmp_path / "a_0000000001_parcels.nc"
    core.nc_write_parcels(path, GRID, p, 2.5, phys={"q0": 0.015, "lambda_c": 0.5})
    raw = open(path, "rb").read(4)
    assert raw == b"CDF\x02"
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions["n_parcels"] == 37 and f.dimensions["world%size"] == 1
        assert f.dimensions["t"] is None   # the unlimited dimension
        assert _text(f.file_type) == "parcels" and _text(f.Conventions) == "CF-1.8"
        assert list(f.ncells) == [8, 6, 4]
        assert np.allclose(f.extent, GRID[4]) and np.allclose(f.origin, GRID[3])
        assert getattr(f, "physical_quantities.saturation_specific_humidity_at_ground_level") == 0.015
        assert getattr(f, "physical_quantities.scale_height") == 2.0
        t = f.variables["t"]
        assert t[:].tolist() == [2.5]
        assert _text(t.units) == "seconds since 1970-01-01" and _text(t.calendar) == "proleptic_gregorian"
        for name, unit in PARCEL_UNITS.items():
            v = f.variables[name]
            assert v.dimensions == ("t", "n_parcels") and _text(v.units) == unit, name
        assert np.array_equal(f.variables["x_position"][0], p["position"][:, 0])
        assert np.array_equal(f.variables["z_vorticity"][0], p["vorticity"][:, 2])
        assert np.array_equal(f.variables["B13"][0], p["B"][:, 2])
        assert np.array_equal(f.variables["humidity"][0], p["humidity"])
        assert f.variables["start_index"][:].tolist() == [1]
        assert f.variables["xhi"][:].tolist() == [7] and f.variables["yhi"][:].tolist() == [5]
    back, grid, t = core.nc_read_parcels(path)
    assert t == 2.5 and grid[:3] == (8, 6, 4) and np.allclose(grid[4], GRID[4])
    for k in p:
        assert np.array_equal(back[k], p[k]), k


@pytest.mark.parametrize("version", [1, 2])
def test_reads_files_of_an_independent_writer(tmp_path, version):
    """A file written by scipy in the reference's schema: float32 and
    missing optional datasets, three records of which the last is read."""
    n = 11
    rng = np.random.default_rng(version)
    path = tmp_path / "ext.nc"
    with netcdf_file(path, "w", version=version) as f:
        f.createDimension("t", None)   # scipy wants the unlimited dimension first
        f.createDimension("n_parcels", n)
        f.ncells = np.array([8, 6, 4], dtype=np.int32)
        f.extent = np.array(GRID[4])
        f.origin = np.array(GRID[3])
        tv = f.createVariable("t", "d", ("t",))
        vals = {}
        for name in ("B11", "B12", "B13", "B22", "B23", "x_position", "y_position", "z_position", "volume",
                     "x_vorticity"):
            v = f.createVariable(name, "f" if name == "volume" else "d", ("t", "n_parcels"))
            for r in range(3):
                v[r] = rng.normal(size=n)
            vals[name] = np.array(v[2], dtype=np.float64)
        tv[:] = [0.0, 0.5, 1.0]
    back, grid, t = core.nc_read_parcels(path)
    assert t == 1.0 and grid[:3] == (8, 6, 4)
    assert np.array_equal(back["position"][:, 1], vals["y_position"])
    assert np.array_equal(back["B"][:, 4], vals["B23"])
    assert np.array_equal(back["volume"], vals["volume"])   # float32 widened exactly
    assert np.array_equal(back["vorticity"][:, 0], vals["x_vorticity"])
    assert not back["vorticity"][:, 1:].any() and not back["buoyancy"].any() and not back["humidity"].any()


def test_missing_required_dataset_and_foreign_files(tmp_path):
    path = tmp_path / "bad.nc"
    with netcdf_file(path, "w") as f:
        f.createDimension("n_parcels", 3)
        f.createVariable("x_position", "d", ("n_parcels",))[:] = [1, 2, 3]
    with pytest.raises(core.EpicError, match="The parcel B11 must be provided"):
        core.nc_read_parcels(path)
    h5 = tmp_path / "netcdf4.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(core.EpicError, match="NetCDF classic"):
        core.nc_read_parcels(h5)


def test_field_file_layout(tmp_path):
    nx, ny, nz = GRID[:3]
    H = (nx + 3) * (ny + 3) * (nz + 3)
    rng = np.random.default_rng(7)
    names = ("velog", "vtend", "vortg", "tbuoyg", "dbuoyg", "humg", "volg", "nparg", "nsparg")
    fields = {k: rng.normal(size=(3, H) if k in ("velog", "vtend", "vortg") else H) for k in names}
    fields["nparg"] = rng.integers(0, 20, H).astype(np.float64)
    path = tmp_path / "a_fields.nc"
    core.nc_write_fields(path, GRID, fields, 0.75, phys={"glat": 4.0})

    def interior(a):   # halo layout f(-1:nz+1, -1:ny+1, -1:nx+1) -> [z][y][x] of the nodes
        return a.reshape(nx + 3, ny + 3, nz + 3)[1:nx + 1, 1:ny + 1, 1:nz + 2].transpose(2, 1, 0)

    with netcdf_file(path, "r", mmap=False) as f:
        assert (f.dimensions["x"], f.dimensions["y"], f.dimensions["z"]) == (nx, ny, nz + 1)
        assert _text(f.file_type) == "fields"
        dx = GRID[4][0] / nx
        assert np.allclose(f.variables["x"][:], GRID[3][0] + dx * np.arange(nx))
        assert np.allclose(f.variables["z"][:], GRID[3][2] + GRID[4][2] / nz * np.arange(nz + 1))
        assert _text(f.variables["z"].axis) == "Z"
        u = f.variables["y_velocity"]
        assert u.dimensions == ("t", "z", "y", "x") and _text(u.units) == "m/s"
        assert np.array_equal(u[0], interior(fields["velog"][1]))
        assert np.array_equal(f.variables["volume"][0], interior(fields["volg"]))
        assert f.variables["nparg"].typecode() == "i"
        assert np.array_equal(f.variables["nparg"][0], interior(fields["nparg"]).astype(np.int32))
        lwc = f.variables["liquid_water_content"][0]
        assert np.allclose(lwc, (interior(fields["tbuoyg"]) - interior(fields["dbuoyg"])) / 4.0)
        assert f.variables["t"][:].tolist() == [0.75]


def _cdf5_file(path, n, data, t):
    """A CDF-5 (64-bit data) parcel file built byte by byte from the
    published format description: 8-byte counts, sizes and offsets."""
    import struct

    def name(s):
        b = s.encode()
        return struct.pack(">Q", len(b)) + b + b"\0" * ((-len(b)) % 4)

    names = list(data)
    hdr = b"CDF\x05" + struct.pack(">Q", 1)
    hdr += struct.pack(">IQ", 0x0A, 2) + name("t") + struct.pack(">Q", 0) + name("n_parcels") + struct.pack(">Q", n)
    hdr += struct.pack(">IQ", 0, 0)   # no global attributes
    nv = len(names) + 1
    var_hdr_len = None
    # two passes: the header length fixes the offsets
    for _ in range(2):
        body = struct.pack(">IQ", 0x0B, nv)
        off = (len(hdr) + (var_hdr_len or 0))
        body += name("t") + struct.pack(">Q", 1) + struct.pack(">Q", 0) + struct.pack(">IQ", 0, 0)
        body += struct.pack(">I", 6) + struct.pack(">Q", 8) + struct.pack(">Q", off)
        off += 8
        for k in names:
            body += name(k) + struct.pack(">Q", 2) + struct.pack(">QQ", 0, 1) + struct.pack(">IQ", 0, 0)
            body += struct.pack(">I", 6) + struct.pack(">Q", 8 * n) + struct.pack(">Q", off)
            off += 8 * n
        var_hdr_len = len(body)
    blob = hdr + body + struct.pack(">d", t)
    for k in names:
        blob += np.asarray(data[k], dtype=">f8").tobytes()
    path.write_bytes(blob)


def test_reads_cdf5(tmp_path):
    n = 5
    rng = np.random.default_rng(11)
    data = {k: rng.normal(size=n) for k in ("B11", "B12", "B13", "B22", "B23", "x_position", "y_position",
                                             "z_position", "volume", "buoyancy")}
    path = tmp_path / "cdf5.nc"
    _cdf5_file(path, n, data, 3.25)
    back, grid, t = core.nc_read_parcels(path)
    assert t == 3.25 and grid[:3] == (0, 0, 0)   # no box attributes
    assert np.array_equal(back["position"][:, 2], data["z_position"])
    assert np.array_equal(back["B"][:, 1], data["B12"])
    assert np.array_equal(back["buoyancy"], data["buoyancy"])


def _field_file(path, grid, fields, times, file_type="fields"):
    """A field file written by scipy (CDF-1): x/y/z_vorticity, buoyancy and
    humidity on (t, z, y, x), several records."""
    nx, ny, nz = grid[:3]
    with netcdf_file(path, "w") as f:
        f.file_type = file_type
        f.ncells = np.array([nx, ny, nz], dtype=np.int32)
        f.extent = np.array(grid[4])
        f.origin = np.array(grid[3])
        f.createDimension("t", None)   # scipy wants the unlimited dimension first
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        f.createDimension("z", nz + 1)
        tv = f.createVariable("t", "d", ("t",))
        for r, t in enumerate(times):
            tv[r] = t
        for name, vals in fields.items():
            v = f.createVariable(name, "d", ("t", "z", "y", "x"))
            v[:] = vals


def test_field_file_reader_and_file_info(tmp_path):
    """read_netcdf_fields (field_netcdf.f90:306-386) and get_file_type /
    get_time / get_num_steps on a file from an independent writer: the
    requested record lands on planes 0..nz of the halo layout; absent
    datasets leave zeros; step -1 is the last record, 0 or past the last is
    the reference's error."""
    nx, ny, nz = GRID[:3]
    rng = np.random.default_rng(5)
    recs = {k: rng.normal(size=(3, nz + 1, ny, nx)) for k in ("x_vorticity", "z_vorticity", "buoyancy")}
    path = tmp_path / "f_fields.nc"
    _field_file(path, GRID, recs, [0.0, 0.5, 1.25])
    ft, t, n, grid = core.nc_file_info(path)
    assert (ft, t, n) == ("fields", 1.25, 3) and grid[:3] == (nx, ny, nz) and np.allclose(grid[4], GRID[4])
    for step, rec in ((-1, 2), (2, 1), (1, 0)):
        f, tt = core.nc_read_fields(path, GRID, step)
        assert tt == [0.0, 0.5, 1.25][rec]
        # halo layout [ix+1, iy+1, iz+1] <- file (z, y, x)
        want = lambda a: np.transpose(a[rec], (2, 1, 0))  # noqa: E731
        assert np.array_equal(f["vortg"][0][1:nx + 1, 1:ny + 1, 1:nz + 2], want(recs["x_vorticity"]))
        assert np.array_equal(f["vortg"][2][1:nx + 1, 1:ny + 1, 1:nz + 2], want(recs["z_vorticity"]))
        assert np.array_equal(f["tbuoyg"][1:nx + 1, 1:ny + 1, 1:nz + 2], want(recs["buoyancy"]))
        assert not f["vortg"][1].any() and not f["humg"].any()   # absent datasets
        assert f["vortg"][0][0].sum() == 0 and f["vortg"][0][:, :, 0].sum() == 0   # halo untouched (zero)
    for bad in (0, 4):
        with pytest.raises(core.EpicError, match="out of bounds"):
            core.nc_read_fields(path, GRID, bad)


def test_next_write_time():
    """get_next_write_time (utils.f90:48-54): freq (int(t / freq) + 1)."""
    assert core.next_write_time(0.5, 0.0) == 0.5
    assert core.next_write_time(0.5, 1.2) == 1.5
    assert core.next_write_time(0.25, 1.0) == 1.25   # a time on the grid moves on
    assert core.next_write_time(2.0, 7.99) == 8.0
