"""Output and restart files on the device path: GPU encoding of device arrays (chunked, double
buffered) read back by scipy; Coupler::write_output_file through the C++ driver (variables,
names, hydrostatic profiles, the dycore's perturbation fields against the oracle); and
Coupler::overwrite_with_restart resuming a run bitwise from a double-precision file."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import porturb_amd as pa
from oracle import oracle_py
from porturb_amd import cases

pytestmark = pytest.mark.gpu


def test_device_encode_decode_multi_chunk(tmp_path):
    n = 5_000_000                       # 40 MB of doubles: two staging chunks of the double variable
    rng = np.random.default_rng(9)
    a = rng.standard_normal(n) * 1e3
    a[:4] = [0.0, -0.0, 1e-310, 3.4e38]
    ai = np.clip(a, -2e9, 2e9)           # in the range of int for the integer variable
    buf = pa.DeviceBuffer(a.nbytes)
    buf.upload(a)
    bufi = pa.DeviceBuffer(ai.nbytes)
    bufi.upload(ai)
    p = str(tmp_path / "dev.nc")
    f = pa.NcFile(p, "w")
    f.def_dim("n", n)
    f.def_var("d", pa.PU_NC_DOUBLE, ["n"])
    f.def_var("f", pa.PU_NC_FLOAT, ["n"])
    f.def_var("i", pa.PU_NC_INT, ["n"])
    f.enddef()
    f.put_device("d", buf.ptr, None)
    f.put_device("f", buf.ptr, None)
    f.put_device("i", bufi.ptr, None)
    f.close()
    with netcdf_file(p, "r", mmap=False) as nc:
        assert np.array_equal(nc.variables["d"].data, a)
        assert np.array_equal(nc.variables["f"].data, a.astype(np.float32))
        assert np.array_equal(nc.variables["i"].data, ai.astype(np.int32))
    back = pa.DeviceBuffer(a.nbytes)
    g = pa.NcFile(p)
    for name, ref in (("d", a), ("f", a.astype(np.float32).astype(np.float64)), ("i", ai.astype(np.int32).astype(np.float64))):
        g.get_device(name, back.ptr, None)
        assert np.array_equal(back.download(np.float64, (n,)), ref), name
    g.close()
    for b in (buf, bufi, back):
        b.free()


NX, NY, NZ, DX = 48, 32, 32, 2.0


def test_output_file_contents(tmp_path):
    drv = pa.Driver(pa.PU_CASE_ABL_URBAN, NX, NY, NZ, DX, ord=9, seed=81)
    init_state = drv.download()
    drv.step(2)
    st = drv.download()
    path = drv.write_output(str(tmp_path / "abl"))
    assert path.endswith("abl_00000000.nc")
    cfg = cases.abl_urban(NX, NY, NZ, DX, ord=9)
    names = ["density_dry", "uvel", "vvel", "wvel", "temp"] + drv.tracer_names()
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables
        assert nc.dimensions["z_halo"] == NZ + 8
        for n in names:
            fn = "temperature" if n == "temp" else n
            assert v[fn].dimensions == ("z", "y", "x") and np.array_equal(v[fn].data, st[n]), n
        assert np.allclose(v["x"].data, (np.arange(NX) + 0.5) * DX) and np.allclose(v["z"].data, (np.arange(NZ) + 0.5) * DX)
        assert v["file_counter"].data[0] == 0
        assert v["etime"].data[0] == pytest.approx(drv.elapsed_time(), rel=1e-15)
        imm = v["immersed_proportion"].data.copy()
        assert imm.max() == 1.0 and imm.min() == 0.0
        # the oracle's dycore built from the same initial state holds the same hydrostatics and
        # gives the same perturbation fields
        c0 = np.stack([init_state[n] for n in names])
        od = oracle_py.Dycore(cfg, c0, imm)
        hy = od.hydrostatics()
        for k, hn in enumerate(("hy_dens_cells", "hy_theta_cells", "hy_pressure_cells")):
            assert np.allclose(v[hn].data, hy[k], rtol=1e-13, atol=0), hn
        ref = od.output_fields(np.stack([st[n] for n in names]))
        for k, pn in enumerate(("density_pert", "theta_pert", "pressure_pert")):
            d = np.abs(v[pn].data - ref[k]).max() / np.abs(ref[k]).max()
            assert d < 1e-10, (pn, d)
    drv.destroy()


@pytest.mark.parametrize("precision", [pa.PU_FP64, pa.PU_FP32])
def test_restart_resumes_bitwise(tmp_path, precision):
    a = pa.Driver(pa.PU_CASE_ABL_URBAN, NX, NY, NZ, DX, ord=9, precision=precision, seed=82)
    a.step(3)
    path = a.write_output(str(tmp_path / "run"))
    a.step(2)
    sa = a.download()
    b = pa.Driver(pa.PU_CASE_ABL_URBAN, NX, NY, NZ, DX, ord=9, precision=precision, seed=5)   # other seed
    b.restart(path)
    assert b.elapsed_time() == pytest.approx(a.elapsed_time() - 2 * (a.elapsed_time() / 5), rel=1e-12)
    b.step(2)
    sb = b.download()
    for n in sa:
        assert np.array_equal(sa[n], sb[n]), n
    assert b.elapsed_time() == a.elapsed_time()
    # the restart advanced the file counter past the file it read
    assert b.write_output(str(tmp_path / "next")).endswith("next_00000001.nc")
    a.destroy()
    b.destroy()


def test_float_output_restart_rounds_the_state(tmp_path):
    a = pa.Driver(pa.PU_CASE_ABL_URBAN, NX, NY, NZ, DX, ord=9, seed=83)
    a.step(1)
    sa = a.download()
    path = a.write_output(str(tmp_path / "f32"), double_precision=False)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.variables["uvel"].data.dtype == np.dtype(">f4")
        assert nc.variables["hy_dens_cells"].data.dtype == np.dtype(">f4")
        assert nc.variables["density_pert"].data.dtype == np.dtype(">f8")      # real in the reference
    b = pa.Driver(pa.PU_CASE_ABL_URBAN, NX, NY, NZ, DX, ord=9, seed=83)
    b.restart(path)
    sb = b.download()
    for n in sa:
        assert np.array_equal(sb[n], sa[n].astype(np.float32).astype(np.float64)), n
    b.step(1)
    assert np.isfinite(b.download()["uvel"]).all()
    a.destroy()
    b.destroy()
