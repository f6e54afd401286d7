"""Fluid dumps (restart files, modflu/RFLU_ModRestartData.F90) through the library's
host codec, on the CPU.

The fixtures in tests/golden/dumps were written by tests/golden/make_dump_fixtures.f90,
a Fortran program using the statement forms of the reference's writer with stream
access (amdflang), so they pin the byte layout a Fortran build produces.  The layout
restatement in oracle/dump.py checks the codec's bytes field by field."""
import os
import struct

import numpy as np
import pytest

from oracle import dump as odump
from rocflu_amd import _native as nat
from rocflu_amd import restart

HERE = os.path.dirname(os.path.abspath(__file__))
DUMPS = os.path.join(HERE, "golden", "dumps")


def fixture(name):
    return os.path.join(DUMPS, name + ".dat")


def fixture_cv():
    # cv(iCv, ic) = 10 ic + iCv + 0.125 (1-based), as the fixture program sets it
    return np.array([[10.0 * ic + iv + 0.125 for iv in range(1, 6)] for ic in range(1, 4)])


def test_fixture_layout_offsets():
    """The spec's offsets read straight from the Fortran-written bytes."""
    b = open(fixture("v5_gas"), "rb").read()
    assert len(b) == 48 + 8 * 15 + 4
    assert b[:8] == b"RFLUDUMP"
    assert struct.unpack_from("<i", b, 8)[0] == 5
    assert struct.unpack_from("<q", b, 12)[0] == 1234
    assert struct.unpack_from("<d", b, 20)[0] == 1.5e-4
    assert struct.unpack_from("<5i", b, 28) == (7, 3, 5, 1, 0)
    # one variable after the other: the first three doubles are variable 1 of cells 1..3
    assert struct.unpack_from("<3d", b, 48) == (11.125, 21.125, 31.125)
    assert struct.unpack_from("<d", b, 48 + 8 * 3)[0] == 12.125
    assert struct.unpack_from("<i", b, 48 + 8 * 15)[0] == 0


def test_oracle_restatement_matches_fixtures():
    assert odump.encode(fixture_cv(), 1234, 1.5e-4, 7) == open(fixture("v5_gas"), "rb").read()
    for name in ("v5_gas", "v5_picl", "v4_picl", "v5_nscbc", "v5_weighted"):
        b = open(fixture(name), "rb").read()
        d = odump.decode(b)
        assert d["size"] == len(b), name
        assert np.array_equal(d["cv"], fixture_cv()), name
    assert np.array_equal(odump.decode(open(fixture("v5_picl"), "rb").read())["picl_vf"],
                          [0.25, 0.0, 0.5])


def test_library_writes_the_fortran_bytes(tmp_path):
    p = str(tmp_path / "d.dat")
    restart.write_dump(p, fixture_cv(), 1234, 1.5e-4, 7)
    assert open(p, "rb").read() == open(fixture("v5_gas"), "rb").read()
    assert not os.path.exists(p + ".tmp")


def test_library_write_matches_restatement_on_random_state(tmp_path):
    rng = np.random.default_rng(5)
    cv = rng.standard_normal((1000, 5))
    cv[17, 2] = -0.0
    cv[18, 3] = np.finfo(float).tiny / 4     # subnormal: bit patterns travel unchanged
    p = str(tmp_path / "r.dat")
    restart.write_dump(p, cv, 99999999, 0.123456789, 99999)
    b = open(p, "rb").read()
    assert b == odump.encode(cv, 99999999, 0.123456789, 99999)
    back, step, time, region = restart.read_dump(p, 1000)
    assert back.tobytes() == cv.tobytes()
    assert (step, time, region) == (99999999, 0.123456789, 99999)


@pytest.mark.parametrize("name,header", [
    ("v5_gas", (1234, 1.5e-4, 7)),
    ("v5_picl", (20, 2e-5, 1)),       # PICL fields read past (version 5 layout)
    ("v4_picl", (5, 1e-6, 2)),        # PICL fields read past (version 4 layout)
])
def test_library_reads_fortran_dumps(name, header):
    cv, step, time, region = restart.read_dump(fixture(name), 3)
    assert np.array_equal(cv, fixture_cv())
    assert (step, time, region) == header


@pytest.mark.parametrize("name,code", [
    ("v5_nscbc", nat.RFLU_ERR_DUMP),      # face states for an NSCBC patch this region lacks
    ("v5_weighted", nat.RFLU_ERR_DUMP),   # a particle run's volume-weighted state
])
def test_library_refuses_what_the_reference_refuses(name, code):
    with pytest.raises(nat.RfluError) as e:
        restart.read_dump(fixture(name), 3)
    assert e.value.code == code


def test_library_read_errors(tmp_path):
    good = open(fixture("v5_gas"), "rb").read()
    cases = {
        "magic": (b"RFLUDUMQ" + good[8:], nat.RFLU_ERR_DUMP),
        "version3": (good[:8] + struct.pack("<i", 3) + good[12:], nat.RFLU_ERR_DUMP),
        "short_header": (good[:40], nat.RFLU_ERR_DUMP),
        "short_data": (good[:100], nat.RFLU_ERR_FILE),
        "no_nscbc_count": (good[:-4], nat.RFLU_ERR_FILE),
    }
    for key, (blob, code) in cases.items():
        p = str(tmp_path / (key + ".dat"))
        open(p, "wb").write(blob)
        with pytest.raises(nat.RfluError) as e:
            restart.read_dump(p, 3)
        assert e.value.code == code, key
    with pytest.raises(nat.RfluError) as e:      # cells differ from the region's
        restart.read_dump(fixture("v5_gas"), 4)
    assert e.value.code == nat.RFLU_ERR_DUMP
    with pytest.raises(nat.RfluError) as e:
        restart.read_dump(str(tmp_path / "missing.dat"), 3)
    assert e.value.code == nat.RFLU_ERR_FILE


def test_failed_write_leaves_no_file(tmp_path):
    p = str(tmp_path / "no_such_dir" / "d.dat")
    with pytest.raises(nat.RfluError) as e:
        restart.write_dump(p, fixture_cv(), 1, 0.0, 0)
    assert e.value.code == nat.RFLU_ERR_FILE
    assert not os.path.exists(p) and not os.path.exists(p + ".tmp")


def test_dump_names():
    assert restart.dump_name("", 0, 0) == "restart_data/fluid_00000000_00000.dat"
    assert restart.dump_name("out/", 1234, 7) == "out/restart_data/fluid_00001234_00007.dat"
    with pytest.raises(nat.RfluError):
        restart.dump_name("", 100000000, 0)
