"""NetCDF files of a member on the GPU (epic_write_parcels_nc,
epic_read_parcels_nc, epic_write_fields_nc): a parcel file restarts a new
member with bitwise the same parcels (FP32 attributes widen exactly), and
the field file holds the member's fields (checked with scipy's independent
NetCDF reader)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from epic_amd import workloads as W
from epic_amd.core import EpicError, Member

pytestmark = pytest.mark.gpu


def test_parcel_restart_and_field_file(tmp_path):
    mem = W.rising_thermal(16, seed=97, humidity_plume=True, n_blobs=2)
    print(f"netcdf: thermal16 n={mem.n}", flush=True)
    a = Member.from_workload(mem)
    t = a.ls_rk_step(0.0, 0.01)
    pf = tmp_path / "run_0000000001_parcels.nc"
    a.write_parcels_nc(pf, t)
    b = Member(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, mem.glat, mem.lambda_c, mem.q0, mem.f_cor,
               mem.zeta_zero, max_parcels=mem.n)
    assert b.read_parcels_nc(pf) == t
    pa, pb = a.get_parcels(), b.get_parcels()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    # the restarted member steps like the original
    ta = a.ls_rk_step(t, 0.01, t_initial=0.0)
    tb = b.ls_rk_step(t, 0.01, t_initial=0.0)
    assert ta == tb
    # a grid mismatch is refused
    c = Member(2 * mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, max_parcels=mem.n)
    with pytest.raises(EpicError, match="box differs"):
        c.read_parcels_nc(pf)
    # field file
    a.par2grid()
    a.vor2vel()
    a.vorticity_tendency()
    ff = tmp_path / "run_fields.nc"
    a.write_fields_nc(ff, ta)
    nx, ny, nz = mem.nx, mem.ny, mem.nz
    with netcdf_file(ff, "r", mmap=False) as f:
        for name, field, comp in (("x_velocity", "velog", 0), ("z_vorticity", "vortg", 2), ("volume", "volg", None),
                                  ("y_vorticity_tendency", "vtend", 1), ("humidity", "humg", None)):
            g = a.get_field(field)
            g = g[comp] if comp is not None else g
            want = g[1:nx + 1, 1:ny + 1, 1:nz + 2].transpose(2, 1, 0)
            assert np.array_equal(f.variables[name][0], want), name
        assert np.array_equal(f.variables["nparg"][0], a.get_field("nparg")[1:nx + 1, 1:ny + 1, 1:nz + 2].transpose(2, 1, 0))
