"""Python exception classes of every gt_status (CPU only).

The reference's bindings raise KeyboardInterrupt for Error::Interrupted,
ConnectionError for the adapter / device errors, FileNotFoundError for
FileNotFound / InvalidFileName, BaseException for ThreadError /
PolledAfterTermination / TiffWrite and ValueError for every other variant
(gpu_tracking_python/src/python_bindings.rs:159-192).  Each case below either
provokes the status through the C ABI or checks the class table directly.
"""
import numpy as np
import pytest

import gpu_tracking_amd as gt
from gpu_tracking_amd import _native as N

VALUE = [4, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15, 17, 18, 22, 100, 101]


def test_every_status_has_the_reference_builtin_class():
    assert issubclass(N.error(19, "").__class__, KeyboardInterrupt)
    assert not isinstance(N.error(19, ""), Exception)
    for s in (1, 2):
        assert isinstance(N.error(s, ""), ConnectionError)
    for s in (3, 16):
        assert isinstance(N.error(s, ""), FileNotFoundError)
    for s in (9, 20, 21):
        e = N.error(s, "")
        assert isinstance(e, BaseException) and not isinstance(e, Exception)
    for s in VALUE:
        assert isinstance(N.error(s, ""), ValueError), s
    # every code of the header is covered, and all are GtError with the status
    for s in N.STATUS_NAMES:
        if s == 0:
            continue
        e = N.error(s, "msg")
        assert isinstance(e, gt.GtError) and e.status == s and e.message == "msg"
        assert N.STATUS_NAMES[s] in str(e)


def test_missing_file_is_file_not_found(tmp_path):
    with pytest.raises(FileNotFoundError) as e:
        gt.VideoFile(str(tmp_path / "missing.tif"))
    assert e.value.status == 3


@pytest.mark.parametrize("name,status", [("noext", 11), ("x.png", 12)])
def test_bad_extension_is_value_error(tmp_path, name, status):
    p = tmp_path / name
    p.write_bytes(b"\0" * 16)
    with pytest.raises(ValueError) as e:
        gt.VideoFile(str(p))
    assert e.value.status == status


def test_unsorted_link_is_value_error_with_its_message():
    pts = np.array([[1, 0, 0], [0, 1, 1]], np.float32)
    with pytest.raises(ValueError) as e:
        gt.link(pts, 3)
    assert e.value.status == 6
    assert "sorted in ascending frame number" in e.value.message


def test_too_dense_to_link_is_value_error():
    # two frames of 80 points within one search range: one subnetwork whose
    # load exceeds the limit (linking.rs:677-679)
    rng = np.random.default_rng(0)
    xy = rng.uniform(0, 2, (2, 80, 2)).astype(np.float32)
    pts = np.concatenate([np.column_stack([np.full(80, f, np.float32), xy[f]]) for f in range(2)])
    with pytest.raises(ValueError) as e:
        gt.link(pts, 50)
    assert e.value.status == 18
    assert "abnormally high" in e.value.message


def test_no_device_is_connection_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(ConnectionError) as e:
        gt.Context(gt.make_params(9), 64, 64)
    assert e.value.status == 1
