"""The restart file (pyom2_amd/restart.py) against the reference's own
write_restart / read_restart (for_src/main/restart.f90, flang build with
-fconvert=big-endian as the reference's site_specific.mk files build it):
byte-identical files, and each side reads the other's."""
import os

import numpy as np
import pytest
import torch

from oracle import ref_run
from pyom2_amd import restart as R

needs_ref = pytest.mark.skipif(not ref_run.restart_available(),
                               reason="reference build not present")


def _state(nx, ny, nz, nisle, sf=True, seed=0, itt=1234):
    rng = np.random.default_rng(seed)
    sh3, sh2 = (nz, ny + 4, nx + 4), (ny + 4, nx + 4)
    fields = {n: (rng.standard_normal(sh3), rng.standard_normal(sh3)) for n in R.CORE_3D}
    psi = (rng.standard_normal(sh2), rng.standard_normal(sh2))
    dpsi = tuple(rng.standard_normal(sh2) for _ in range(3)) if sf else None
    dpsin = (rng.standard_normal(nisle), rng.standard_normal(nisle)) if sf else None
    return R.RestartState(nx, ny, nz, itt, fields, psi, dpsi, dpsin)


def _same(a, b):
    for n in R.CORE_3D:
        for x, y in zip(a.fields3[n], b.fields3[n]):
            np.testing.assert_array_equal(x, y, err_msg=n)
    for x, y in zip(a.psi, b.psi):
        np.testing.assert_array_equal(x, y)
    if a.dpsi is not None:
        for x, y in zip(a.dpsi, b.dpsi):
            np.testing.assert_array_equal(x, y)
        for x, y in zip(a.dpsin, b.dpsin):
            np.testing.assert_array_equal(x, y)


@needs_ref
@pytest.mark.parametrize("nx,ny,nz,nisle", [(12, 9, 3, 2), (7, 5, 1, 0), (20, 6, 4, 5)])
def test_file_identical_to_reference_writer(tmp_path, nx, ny, nz, nisle):
    st = _state(nx, ny, nz, nisle)
    ours = open(R.write_restart(str(tmp_path), st), "rb").read()
    ref = ref_run.run_restart_write(nx, ny, nz, st.itt, st.fields3, st.psi, st.dpsi, st.dpsin)
    assert ours == ref


@needs_ref
def test_reference_reads_our_file_and_we_read_its(tmp_path):
    nx, ny, nz, nisle = 10, 8, 3, 3
    st = _state(nx, ny, nz, nisle, seed=3, itt=77)
    path = R.write_restart(str(tmp_path), st)
    itt, fields, psi, dpsi, dpsin = ref_run.run_restart_read(open(path, "rb").read(), nx, ny,
                                                             nz, nisle)
    assert itt == 77
    _same(st, R.RestartState(nx, ny, nz, itt, fields, psi, dpsi, dpsin))
    ref = ref_run.run_restart_write(nx, ny, nz, 78, st.fields3, st.psi, st.dpsi, st.dpsin)
    open(path, "wb").write(ref)
    back = R.read_restart(str(tmp_path), nx, ny, nz, True, nisle)
    assert back.itt == 78
    _same(st, back)


@needs_ref
def test_without_streamfunction(tmp_path):
    """enable_streamfunction off: no dpsi/dpsin records."""
    nx, ny, nz = 9, 7, 2
    st = _state(nx, ny, nz, 0, sf=False, seed=4)
    ours = open(R.write_restart(str(tmp_path), st), "rb").read()
    z = tuple(np.zeros((ny + 4, nx + 4)) for _ in range(3))
    ref = ref_run.run_restart_write(nx, ny, nz, st.itt, st.fields3, st.psi, z, None, sf=False)
    assert ours == ref
    back = R.read_restart(str(tmp_path), nx, ny, nz, False)
    assert back.dpsi is None
    _same(st, back)


def test_checks_and_missing_file(tmp_path):
    st = _state(6, 5, 2, 1)
    R.write_restart(str(tmp_path), st)
    with pytest.raises(R.RestartError):
        R.read_restart(str(tmp_path), 6, 5, 3, True, 1)           # nz differs
    with pytest.raises(R.RestartError):
        R.read_restart(str(tmp_path), 6, 5, 2, True, 1, bounds=(0, 8, -1, 7))
    assert R.read_restart(str(tmp_path / "none"), 6, 5, 2) is None
    assert R.restart_filename(3) == "restart_PE_00003.dta"


def test_record_framing_and_subrecords(tmp_path, monkeypatch):
    """Big-endian 4-byte markers around every record; a record longer than a
    subrecord is split with the leading marker negated while more follow and
    the trailing one negated after the first (checked on a small subrecord
    size), and reads back; little-endian files round-trip too."""
    st = _state(8, 6, 2, 2, seed=5)
    path = R.write_restart(str(tmp_path), st)
    raw = open(path, "rb").read()
    assert int(np.frombuffer(raw[:4], ">i4")[0]) == 16
    assert tuple(np.frombuffer(raw[4:20], ">i4")) == (8, 6, 2, st.itt)
    # gfortran's default subrecord length (-fmax-subrecord-length): a 2048^2 x 40
    # field pair (2.7 GB) is split at 2**31 - 9 bytes
    assert R.MAX_SUB == 2 ** 31 - 9
    n = 2 * 40 * 2052 * 2052 * 8
    assert R.subrecord_sizes(n) == [2 ** 31 - 9, n - (2 ** 31 - 9)]
    assert R.subrecord_sizes(2 ** 31 - 9) == [2 ** 31 - 9]
    assert R.subrecord_sizes(2 ** 31 - 8) == [2 ** 31 - 9, 1]
    monkeypatch.setattr(R, "MAX_SUB", 1000)
    assert R.subrecord_sizes(3840) == [1000, 1000, 1000, 840]
    p2 = str(tmp_path / "split.dta")
    R.write_restart(p2, st)
    raw2 = open(p2, "rb").read()
    # the first field record: 2 x 2 x 10 x 12 doubles = 3840 bytes -> 1000+1000+1000+840
    off = 24 + 24
    marks = []
    for _ in range(4):
        lead = int(np.frombuffer(raw2[off:off + 4], ">i4")[0])
        k = abs(lead)
        trail = int(np.frombuffer(raw2[off + 4 + k:off + 8 + k], ">i4")[0])
        marks.append((lead, trail))
        off += 8 + k
    assert marks == [(-1000, 1000), (-1000, -1000), (-1000, -1000), (840, -840)]
    _same(st, R.read_restart(p2, 8, 6, 2, True, 2))
    p3 = str(tmp_path / "le.dta")
    R.write_restart(p3, st, endian="<")
    _same(st, R.read_restart(p3, 8, 6, 2, True, 2, endian="<"))


@pytest.mark.gpu
def test_barotropic_state_round_trip(tmp_path):
    """A restart into the solver's BarotropicState on the device and back."""
    from pyom2_amd.streamfunction import BarotropicState
    nx, ny, nz, nisle = 16, 12, 2, 3
    st = _state(nx, ny, nz, nisle, seed=6)
    R.write_restart(str(tmp_path), st)
    back = R.read_restart(str(tmp_path), nx, ny, nz, True, nisle)
    z = lambda: torch.zeros((3, ny + 4, nx + 4), dtype=torch.float64, device="cuda")
    bs = BarotropicState(z(), z(), np.zeros((nisle, 3)), None, np.zeros((nisle, nisle)), 3600.0)
    R.barotropic_from_restart(back, bs, 2, 0, 1)
    psi, dpsi, dpsin = R.barotropic_records(bs, nisle)
    for x, y in zip(psi, st.psi):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(dpsi, st.dpsi):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(dpsin, st.dpsin):
        np.testing.assert_array_equal(x, y)
