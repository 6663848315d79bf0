"""Reader of the reference's observation files and writer of its period
files (lsamd_read_table, lsamd_write_periods), on the CPU.

The reader is checked against numpy's parse of the reference's own data files
(the values are short decimals, so the float parse equals the rounded double
parse), against the token rules of the reference's stream reader (commas and
whitespace both separate values; a value that does not parse ends its
comma-separated field), and the writer against 6-significant-digit formatting
of the same values.
"""
import os

import numpy as np
import pytest

from conftest import DATA
from gpu_lomb_scargle_amd import _native


@pytest.mark.parametrize("name,with_dy", [("SDSS_stripe82_band_z.txt", True),
                                          ("8205_normalized_with_error.txt", True),
                                          ("8205_normalized.txt", False),
                                          ("243_normalized_ztf_filter2.txt", False)])
def test_read_table_matches_numpy(name, with_dy):
    path = os.path.join(DATA, name)
    ref = np.loadtxt(path, delimiter=",")
    for dtype in (np.float64, np.float32):
        ids, t, y, dy = _native.read_table(path, dtype, with_dy)
        assert ids.size == ref.shape[0]
        np.testing.assert_array_equal(ids, ref[:, 0].astype(dtype).astype(np.uint32))
        np.testing.assert_array_equal(t, ref[:, 1].astype(dtype))
        np.testing.assert_array_equal(y, ref[:, 2].astype(dtype))
        if with_dy:
            np.testing.assert_array_equal(dy, ref[:, 3].astype(dtype))


def test_read_table_token_rules(tmp_path):
    p = tmp_path / "obs.txt"
    # whitespace and commas both separate values; "x7" ends its field, so 9
    # after it in the same field is dropped; a trailing partial row is ignored
    p.write_text("1, 0.5, 2.5\n1,1.5,3.5\n2 ,2.5e1 ,-4\n3, 4, 5 x7 9, 6\n7")
    ids, t, y, _ = _native.read_table(str(p), np.float64, False)
    assert ids.tolist() == [1, 1, 2, 3]
    assert t.tolist() == [0.5, 1.5, 25.0, 4.0]
    assert y.tolist() == [2.5, 3.5, -4.0, 5.0]
    with pytest.raises(_native.LsamdError):
        _native.read_table(str(tmp_path / "missing.txt"))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_write_periods_formats(tmp_path, dtype):
    rng = np.random.default_rng(3)
    ids = np.array([5, 9, 12], np.uint32)
    per = rng.uniform(0.05, 10, (3, 3)).astype(dtype)
    pw = rng.uniform(0, 1, (3, 3)).astype(dtype)
    p3 = tmp_path / "bestperiods_top_three.txt"
    _native.write_periods(str(p3), ids, per, pw)
    want = "".join("%d, %s\n" % (i, ", ".join("%.6g" % float(v) for v in list(a) + list(b)))
                   for i, a, b in zip(ids, per, pw))
    assert p3.read_text() == want
    p1 = tmp_path / "bestperiods.txt"
    _native.write_periods(str(p1), ids, per[:, 0].copy(), pw[:, 0].copy())
    want = "".join("%d, %.6g, %.6g\n" % (i, float(a), float(b)) for i, a, b in zip(ids, per[:, 0], pw[:, 0]))
    assert p1.read_text() == want
