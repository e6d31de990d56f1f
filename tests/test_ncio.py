"""The NetCDF classic reader/writer behind the output and restart files (pu_nc_*, the format of
Coupler::write_output_file / overwrite_with_restart, model/core/coupler.h:533-723), checked
against an independent implementation (scipy.io.netcdf_file reads and writes CDF-1 and CDF-2) and
against the published layout of CDF-5 headers.  Host transfers only: no device needed."""
import struct

import numpy as np
import pytest
from scipy.io import netcdf_file

import porturb_amd as pa


def write_sample(path, fmt=pa.PU_NC_AUTO):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((3, 4, 5))
    f = pa.NcFile(str(path), "w", fmt)
    f.def_dim("x", 5)
    f.def_dim("y", 4)
    f.def_dim("z", 3)
    f.def_dim("t", 1)
    f.def_var("temperature", pa.PU_NC_DOUBLE, ["z", "y", "x"])
    f.def_var("uvel", pa.PU_NC_FLOAT, ["z", "y", "x"])
    f.def_var("etime", pa.PU_NC_FLOAT, ["t"])
    f.def_var("file_counter", pa.PU_NC_INT, ["t"])
    f.enddef()
    f.put("temperature", a)
    f.put("uvel", a)
    f.put("etime", np.array([12.25]))
    f.put("file_counter", np.array([3], np.int32))
    # a module hook adds variables after the first enddef (dynamics_rk.h:1198-1210)
    f.redef()
    f.def_dim("z_halo", 7)
    f.def_var("hy_dens_cells", pa.PU_NC_FLOAT, ["z_halo"])
    f.enddef()
    f.put("hy_dens_cells", np.linspace(1.2, 0.9, 7))
    v = f.version()
    f.close()
    return a, v


def test_cdf2_files_read_by_scipy(tmp_path):
    p = tmp_path / "out_00000000.nc"
    a, v = write_sample(p)
    assert v == 2
    with netcdf_file(str(p), "r", mmap=False) as nc:
        assert nc.version_byte == 2
        assert nc.dimensions == {"x": 5, "y": 4, "z": 3, "t": 1, "z_halo": 7}
        t = nc.variables["temperature"]
        assert t.dimensions == ("z", "y", "x") and t.data.dtype == np.dtype(">f8")
        assert np.array_equal(t.data, a)
        u = nc.variables["uvel"]
        assert u.data.dtype == np.dtype(">f4") and np.array_equal(u.data, a.astype(np.float32))
        assert nc.variables["etime"].data[0] == np.float32(12.25)
        assert nc.variables["file_counter"].data.dtype == np.dtype(">i4") and nc.variables["file_counter"].data[0] == 3
        assert np.array_equal(nc.variables["hy_dens_cells"].data, np.linspace(1.2, 0.9, 7).astype(np.float32))


@pytest.mark.parametrize("version", [1, 2])
def test_reads_scipy_written_files(tmp_path, version):
    p = tmp_path / ("scipy_v%d.nc" % version)
    rng = np.random.default_rng(6)
    a = rng.standard_normal((2, 3, 4))
    with netcdf_file(str(p), "w", version=version) as nc:
        nc.createDimension("x", 4)
        nc.createDimension("y", 3)
        nc.createDimension("z", 2)
        nc.createVariable("density_dry", "d", ("z", "y", "x"))[:] = a
        nc.createVariable("wvel", "f", ("z", "y", "x"))[:] = a
        nc.createVariable("file_counter", "i", ("z",))[:] = [7, -2]
        nc.title = "attributes are skipped"
        nc.variables["wvel"].units = "m/s"
    f = pa.NcFile(str(p))
    assert f.version() == version
    assert f.inq("density_dry") == (pa.PU_NC_DOUBLE, (2, 3, 4))
    assert np.array_equal(f.get("density_dry"), a)
    assert np.array_equal(f.get("wvel"), a.astype(np.float32).astype(np.float64))
    assert np.array_equal(f.get("file_counter", np.int32), [7, -2])
    f.close()


def test_cdf5_header_layout_and_round_trip(tmp_path):
    p = tmp_path / "cdf5.nc"
    a, v = write_sample(p, pa.PU_NC_CDF5)
    assert v == 5
    raw = open(p, "rb").read(64)
    # magic, 64-bit numrecs, NC_DIMENSION tag, 64-bit count, first name (64-bit length, padded)
    assert raw[:4] == b"CDF\x05"
    assert struct.unpack(">Q", raw[4:12])[0] == 0
    assert struct.unpack(">I", raw[12:16])[0] == 0x0A
    assert struct.unpack(">Q", raw[16:24])[0] == 5
    assert struct.unpack(">Q", raw[24:32])[0] == 1 and raw[32:33] == b"x"
    assert struct.unpack(">Q", raw[36:44])[0] == 5          # dim length after the padded name
    f = pa.NcFile(str(p))
    assert f.version() == 5
    assert np.array_equal(f.get("temperature"), a)
    assert np.array_equal(f.get("uvel"), a.astype(np.float32).astype(np.float64))
    assert f.get("file_counter", np.int32)[0] == 3
    f.close()


def test_unsigned_bytes_select_cdf5_and_errors(tmp_path):
    p = tmp_path / "u8.nc"
    f = pa.NcFile(str(p), "w")
    f.def_dim("n", 3)
    f.def_var("mask", pa.PU_NC_UBYTE, ["n"])
    with pytest.raises(RuntimeError):
        f.def_var("bad", pa.PU_NC_FLOAT, ["missing"])
    with pytest.raises(RuntimeError):
        f.put("mask", np.zeros(3))                     # still in define mode
    f.enddef()
    assert f.version() == 5
    f.close()
    g = pa.NcFile(str(p), "w", pa.PU_NC_CDF2)
    g.def_dim("n", 3)
    g.def_var("mask", pa.PU_NC_UBYTE, ["n"])
    with pytest.raises(RuntimeError):
        g.enddef()                                      # CDF-2 has no unsigned bytes
    with pytest.raises(RuntimeError):
        g.close()                                       # closing retries the definition
    q = tmp_path / "junk.nc"
    q.write_bytes(b"HDF\x89 not netcdf classic")
    with pytest.raises(RuntimeError):
        pa.NcFile(str(q))
