"""The driver's classic-netCDF reader/writer (qg-ybjp-model_amd/fortran/qg_ncio.c),
checked against SciPy's independent netCDF-3 implementation
(scipy.io.netcdf_file) in both directions, on the schema of the reference's
dumps and initial conditions (IO_ncf.f90: psi(x, y, z) / LAr, LAi over the
dimensions x, y, z, t with a one-element time; N2(z)).
"""
import ctypes
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import PKG

I32 = ctypes.c_int32


@pytest.fixture(scope="module")
def nc():
    so = os.path.join(PKG, "lib", "libqgncio.so")
    if not os.path.exists(so):
        import build
        build.build_ncio()
    lib = ctypes.CDLL(so)
    lib.qg_nc_error.restype = ctypes.c_char_p
    lib.qg_nc_write.restype = ctypes.c_int
    lib.qg_nc_read.restype = ctypes.c_int
    return lib


def _write(lib, path, dims, lens, variables):
    """variables: list of (name, dim indices slowest first, float32 array)."""
    names = ",".join(v[0] for v in variables).encode()
    nd = (I32 * len(variables))(*[len(v[1]) for v in variables])
    flat = [d for v in variables for d in v[1]]
    vd = (I32 * max(1, len(flat)))(*flat)
    arrs = [np.ascontiguousarray(v[2], dtype=np.float32) for v in variables]
    ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return lib.qg_nc_write(str(path).encode(), ",".join(dims).encode(), (I32 * len(lens))(*lens), len(lens), names,
                           nd, vd, ptrs, len(variables))


def _read(lib, path, var, dims, lens):
    out = np.zeros(int(np.prod(lens)) if lens else 1, np.float32)
    rc = lib.qg_nc_read(str(path).encode(), var.encode(), ",".join(dims).encode(), (I32 * max(1, len(lens)))(*lens),
                        len(lens), out.ctypes.data_as(ctypes.c_void_p))
    return rc, out.reshape(lens) if lens else out


def test_dump_schema_read_by_scipy(nc, tmp_path):
    n1, n3 = 16, 5
    rng = np.random.default_rng(0)
    lar = rng.standard_normal((n3, n1, n1)).astype(np.float32)  # C order (z, y, x) = Fortran (x, y, z)
    lai = rng.standard_normal((n3, n1, n1)).astype(np.float32)
    t = np.array([1234.5], np.float32)
    p = tmp_path / "la000.out.nc"
    assert _write(nc, p, ["x", "y", "z", "t"], [n1, n1, n3, 1],
                  [("LAr", [2, 1, 0], lar), ("LAi", [2, 1, 0], lai), ("time", [3], t)]) == 0, nc.qg_nc_error()
    assert open(p, "rb").read(4) == b"CDF\x02"
    with netcdf_file(p, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions == {"x": n1, "y": n1, "z": n3, "t": 1}
        assert list(f.variables) == ["LAr", "LAi", "time"]
        for name, a in (("LAr", lar), ("LAi", lai)):
            v = f.variables[name]
            assert v.dimensions == ("z", "y", "x") and v.data.dtype == np.dtype(">f4")
            assert np.array_equal(v.data, a)
        assert f.variables["time"].dimensions == ("t",)
        assert f.variables["time"].data[0] == t[0]
    # a profile over z alone (the N^2 dump)
    n2 = np.linspace(1.0, 100.0, n3).astype(np.float32)
    p2 = tmp_path / "N2.out.nc"
    assert _write(nc, p2, ["z"], [n3], [("N2", [0], n2)]) == 0
    with netcdf_file(p2, "r", mmap=False) as f:
        assert f.dimensions == {"z": n3} and np.array_equal(f.variables["N2"].data, n2)


@pytest.mark.parametrize("version", [1, 2])
def test_reads_scipy_files(nc, tmp_path, version):
    """Files written by another implementation, with attributes, a record
    (unlimited) time dimension and a FLOAT and a DOUBLE field."""
    n1, n3 = 12, 4
    rng = np.random.default_rng(version)
    psi64 = rng.standard_normal((n3, n1, n1)) * 1e3
    p = tmp_path / "psi000.in.nc"
    with netcdf_file(p, "w", version=version) as f:
        f.title = "initial condition"
        f.createDimension("t", None)
        f.createDimension("x", n1)
        f.createDimension("y", n1)
        f.createDimension("z", n3)
        tt = f.createVariable("time", "f4", ("t",))
        tt.units = "s"
        tt[0] = 42.0
        v = f.createVariable("psi", "f4", ("z", "y", "x"))
        v.units = "m2 s-1"
        v[:] = psi64.astype(np.float32)
        w = f.createVariable("psid", "f8", ("z", "y", "x"))
        w[:] = psi64
    rc, got = _read(nc, p, "psi", ["z", "y", "x"], [n3, n1, n1])
    assert rc == 0, nc.qg_nc_error()
    assert np.array_equal(got, psi64.astype(np.float32))
    rc, got = _read(nc, p, "psid", ["z", "y", "x"], [n3, n1, n1])
    assert rc == 0 and np.array_equal(got, psi64.astype(np.float32))  # DOUBLE rounded to the nearest float
    rc, t = _read(nc, p, "time", ["t"], [1])
    assert rc == 0 and t[0] == 42.0


def test_errors(nc, tmp_path):
    n1, n3 = 8, 3
    a = np.zeros((n3, n1, n1), np.float32)
    p = tmp_path / "a.nc"
    assert _write(nc, p, ["x", "y", "z", "t"], [n1, n1, n3, 1], [("psi", [2, 1, 0], a), ("time", [3], a[:1, 0, 0])]) == 0
    rc, _ = _read(nc, p, "psi", ["z", "y", "x"], [n3, n1, n1 + 2])  # another resolution
    assert rc != 0 and b"x(8), x(10) expected" in nc.qg_nc_error()
    rc, _ = _read(nc, p, "psi", ["y", "z", "x"], [n1, n3, n1])  # dimensions in another order
    assert rc != 0
    rc, _ = _read(nc, p, "LAr", ["z", "y", "x"], [n3, n1, n1])
    assert rc != 0 and b"no variable LAr" in nc.qg_nc_error()
    h = tmp_path / "h.nc"
    h.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    rc, _ = _read(nc, h, "psi", ["z", "y", "x"], [n3, n1, n1])
    assert rc != 0 and b"netCDF-4" in nc.qg_nc_error()
    short = tmp_path / "short.nc"
    short.write_bytes(p.read_bytes()[:-100])
    rc, _ = _read(nc, short, "psi", ["z", "y", "x"], [n3, n1, n1])
    assert rc != 0 and b"too short" in nc.qg_nc_error()
    rc, _ = _read(nc, tmp_path / "missing.nc", "psi", ["z", "y", "x"], [n3, n1, n1])
    assert rc != 0 and b"cannot open" in nc.qg_nc_error()
    assert _write(nc, p, ["x", "y"], [n1, 0], [("v", [1, 0], a)]) != 0  # zero-length dimension
