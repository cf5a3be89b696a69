"""Restart-file ingestion (suli_ps_restart_read; no GPU needed): the
record restart_salva writes (SuLi-IC/11_diversos.f90:581-599) parsed back
into the pressure step's arrays.

Pinned by (a) a Fortran program compiled here (amdflang) that writes a
SuLi-shaped record with the compiler's own list-directed output and the
same arrays unformatted, compared value for value; (b) records in the
gfortran style (17-digit fields, r*c repeats, D and bare exponents, commas,
a closing slash) written by tests/restart_helpers.py; (c) the error paths.
The reference ships no restart file, so no record written by SuLi itself
is available."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from suli_amd import _lib
from suli_amd.restart import SHAPES, fields_from_restart, read_restart
from suli_amd.cases import Grid
from tests.restart_helpers import write_record

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("SULI_FC", "/opt/rocm/lib/llvm/bin/amdflang")
NAMES = ("u", "v", "w", "prd0", "prd1", "prd", "rho", "ls_mu", "ls", "hs")

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libsuli_ps.so not built")


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    if not (os.path.exists(FLANG) or shutil.which(FLANG)):
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("rw")
    exe = str(d / "restart_writer")
    r = subprocess.run([FLANG, "-O1", "-o", exe, os.path.join(HERE, "golden", "restart_writer.F90")],
                       capture_output=True, text=True, cwd=str(d))
    assert r.returncode == 0, r.stderr
    return exe


def raw_arrays(path, nx, ny, nz):
    raw = np.fromfile(path, dtype=np.float64)
    out, o = {}, 0
    for name in NAMES:
        shape = SHAPES[name](nx, ny, nz)
        size = int(np.prod(shape))
        out[name] = raw[o:o + size].reshape(shape, order="F")
        o += size
    assert o == raw.size
    return out


@pytest.mark.parametrize("full", [0, 1])
@pytest.mark.parametrize("dims", [(6, 5, 4), (13, 3, 7)])
def test_compiler_written_record(writer, tmp_path, full, dims):
    nx, ny, nz = dims
    txt, raw = str(tmp_path / "arquivorestart"), str(tmp_path / "raw.bin")
    subprocess.run([writer, str(nx), str(ny), str(nz), str(full), txt, raw], check=True)
    got = read_restart(txt, nx, ny, nz, NAMES)
    exp = raw_arrays(raw, nx, ny, nz)
    for name in NAMES:
        np.testing.assert_array_equal(got[name], exp[name], err_msg=name)


def sample_arrays(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    out = {}
    for name in NAMES:
        a = rng.standard_normal(SHAPES[name](nx, ny, nz)) * 10.0 ** rng.integers(-5, 5)
        out[name] = np.asfortranarray(a)
    out["rho"][:, :, 0] = 998.0          # runs -> r*c repeats
    out["u"][0, 0, 0] = 1e-300           # three-digit exponents
    out["v"][0, 0, 0] = -2.5e150
    return out


@pytest.mark.parametrize("byxf_full", [True, False])
def test_gfortran_style_record(tmp_path, byxf_full):
    nx, ny, nz = 9, 4, 5
    arrays = sample_arrays(nx, ny, nz, 3)
    path = str(tmp_path / "arquivorestart")
    write_record(path, arrays, nx, ny, nz, byxf_full=byxf_full, slash=True)
    got = read_restart(path, nx, ny, nz, NAMES)
    for name in NAMES:
        np.testing.assert_array_equal(got[name], arrays[name], err_msg=name)


def test_null_values_keep_the_destination(tmp_path):
    nx, ny, nz = 5, 4, 3
    arrays = sample_arrays(nx, ny, nz, 4)
    path = str(tmp_path / "arquivorestart")
    write_record(path, arrays, nx, ny, nz, nulls=3)
    got = read_restart(path, nx, ny, nz, ("u", "hs"))
    u = got["u"].ravel(order="F")
    assert np.all(u[:3] == 0.0)          # the reader's buffers start at zero
    np.testing.assert_array_equal(u[3:], arrays["u"].ravel(order="F")[3:])
    np.testing.assert_array_equal(got["hs"], arrays["hs"])


def test_fields_for_the_solver(tmp_path):
    nx, ny, nz = 8, 6, 5
    arrays = sample_arrays(nx, ny, nz, 5)
    path = str(tmp_path / "arquivorestart")
    write_record(path, arrays, nx, ny, nz)
    f = fields_from_restart(path, Grid(nx, ny, nz, 0.01, 0.01, 0.01), with_hs=True)
    for name in ("u", "v", "w", "rho", "prd1", "hs"):
        a = getattr(f, name)
        assert a.flags.f_contiguous and a.dtype == np.float64
        np.testing.assert_array_equal(a, arrays[name])


def test_errors(tmp_path):
    nx, ny, nz = 6, 5, 4
    arrays = sample_arrays(nx, ny, nz, 6)
    path = str(tmp_path / "arquivorestart")
    write_record(path, arrays, nx, ny, nz)
    text = open(path).read()
    cut = str(tmp_path / "cut")
    open(cut, "w").write(text[: len(text) // 3])
    with pytest.raises(ValueError, match="ends after"):
        read_restart(cut, nx, ny, nz, ("rho",))
    with pytest.raises(ValueError, match="fewer than"):
        read_restart(cut, nx, ny, nz, ("hs",))
    bad = str(tmp_path / "bad")
    open(bad, "w").write(" 1.0 2.0 1.2.3x 4.0\n")
    with pytest.raises(ValueError, match="bad value"):
        read_restart(bad, nx, ny, nz, ("u",))
    with pytest.raises(ValueError, match="cannot open"):
        read_restart(str(tmp_path / "missing"), nx, ny, nz, ("u",))
    # a record of another grid size: its byxf gap comes out negative
    with pytest.raises(ValueError, match="fewer than"):
        read_restart(path, nx + 1, ny, nz, ("hs",))


def _plain_record(arrays, nx, ny, nz):
    """The record's values in order, as repr strings (a null-free record)."""
    from tests.restart_helpers import record_shapes
    vals = []
    for name, shape in record_shapes(nx, ny, nz):
        a = arrays.get(name)
        a = np.zeros(shape) if a is None else np.asarray(a, dtype=np.float64)
        vals.extend(repr(float(x)) for x in a.ravel(order="F"))
    return vals


def test_commas_mark_null_values(tmp_path):
    """List-directed input: nothing between two commas (blanks and line
    ends aside) is a null value, and so is a comma before the record's
    first value; a comma at a line end followed by a value on the next
    line is one separator."""
    nx, ny, nz = 5, 4, 3
    arrays = sample_arrays(nx, ny, nz, 7)
    vals = _plain_record(arrays, nx, ny, nz)
    # u(0..9): v0,,v2 , ,v4,<newline> v5 ,,, v8 v9 ...: nulls at 1, 3, 6, 7
    head = f"{vals[0]},,{vals[2]} , ,{vals[4]},\n {vals[5]} ,,, {vals[8]} "
    path = str(tmp_path / "commas")
    with open(path, "w") as f:
        f.write(" " + head + " ".join(vals[9:]) + "\n")
    u = read_restart(path, nx, ny, nz, ("u", "hs"))
    got = u["u"].ravel(order="F")
    exp = arrays["u"].ravel(order="F").copy()
    exp[[1, 3, 6, 7]] = 0.0               # the reader's buffers start at zero
    np.testing.assert_array_equal(got[:20], exp[:20])
    # the nulls count as values: everything after them is in place
    np.testing.assert_array_equal(got[9:], arrays["u"].ravel(order="F")[9:])
    np.testing.assert_array_equal(u["hs"], arrays["hs"])
    # a comma before the first value: u(0) null, the rest shifted by none
    lead = str(tmp_path / "lead")
    with open(lead, "w") as f:
        f.write(" , " + " ".join(vals[1:]) + "\n")
    got = read_restart(lead, nx, ny, nz, ("u",))["u"].ravel(order="F")
    assert got[0] == 0.0
    np.testing.assert_array_equal(got[1:], arrays["u"].ravel(order="F")[1:])


def test_overlong_value_is_an_error(tmp_path):
    nx, ny, nz = 5, 4, 3
    path = str(tmp_path / "long")
    with open(path, "w") as f:
        f.write(" 1.0 2.0 " + "1" * 200 + " 4.0\n")
    with pytest.raises(ValueError, match="longer than 127"):
        read_restart(path, nx, ny, nz, ("u",))
