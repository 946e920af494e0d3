"""Host-side argument checks of the Python layer (no device needed).

Every wrapper that hands raw device pointers to the C ABI must reject, before
any device call, buffers whose dtype, frame count, shape or strides would let
a kernel read or write outside them.  These tests build objects without a
device handle and pass CPU tensors: the checks fire before any pointer is
taken, so a ValueError naming the problem is raised and nothing else runs.
"""
import threading

import pytest

torch = pytest.importorskip("torch")

import dsmvc_hip as dh  # noqa: E402


def _executor():
    ex = object.__new__(dh.Executor)
    ex._handle = None
    ex.solve = "auto"
    return ex


@pytest.fixture(scope="module")
def plans():
    h = dh.Plan(dh.AxisRequest(96, 64, 64.0))
    v = dh.Plan(dh.AxisRequest(54, 36, 36.0))
    return h, v


def test_inverse_2d_rejects_wrong_output(plans):
    h, v = plans
    ex = _executor()
    x = torch.zeros((2, 54, 96))
    with pytest.raises(ValueError, match="frames"):
        ex.inverse_2d(h, v, x, out=torch.zeros((1, 36, 64)))
    with pytest.raises(ValueError, match="float32"):
        ex.inverse_2d(h, v, x, out=torch.zeros((2, 36, 64), dtype=torch.float64))
    with pytest.raises(ValueError, match="at least"):
        ex.inverse_2d(h, v, x, out=torch.zeros((2, 36, 63)))
    with pytest.raises(ValueError, match="contiguous"):
        ex.inverse_2d(h, v, x, out=torch.zeros((2, 64, 36)).transpose(1, 2))
    with pytest.raises(ValueError, match="does not match"):
        ex.inverse_2d(h, v, torch.zeros((2, 53, 96)))
    with pytest.raises(ValueError, match="conversion"):
        ex.inverse_2d(h, v, torch.zeros((54, 96), dtype=torch.uint8))
    with pytest.raises(ValueError, match="float32, uint8 or uint16"):
        ex.inverse_2d(h, v, torch.zeros((54, 96), dtype=torch.float64))


def test_single_axis_calls_check_sizes_and_dtype(plans):
    h, v = plans
    ex = _executor()
    with pytest.raises(ValueError, match="at least"):
        ex.inverse_rows(h, torch.zeros((5, 95)))
    with pytest.raises(ValueError, match="float32"):
        ex.inverse_rows(h, torch.zeros((5, 96), dtype=torch.float64))
    with pytest.raises(ValueError, match="at least"):
        ex.inverse_rows(h, torch.zeros((5, 96)), out=torch.zeros((4, 64)))
    with pytest.raises(ValueError, match="2-D"):
        ex.inverse_rows(h, torch.zeros((1, 5, 96)))
    with pytest.raises(ValueError, match="fewer rows"):
        ex.inverse_columns(v, torch.zeros((53, 7)))
    with pytest.raises(ValueError, match="at least"):
        ex.inverse_columns(v, torch.zeros((54, 7)), out=torch.zeros((36, 6)))
    with pytest.raises(ValueError, match="float32"):
        ex.inverse_columns(v, torch.zeros((54, 7), dtype=torch.uint16))


def test_batch_checks_each_job(plans):
    h, v = plans
    ex = _executor()
    x = torch.zeros((3, 54, 96))
    with pytest.raises(ValueError, match="frames"):
        dh.Batch(ex, [(h, v, x, torch.zeros((2, 36, 64)))])
    with pytest.raises(ValueError, match="float32"):
        dh.Batch(ex, [(h, v, x.double(), torch.zeros((3, 36, 64)))])
    with pytest.raises(ValueError, match="contiguous"):
        dh.Batch(ex, [(h, v, torch.zeros((3, 96, 54)).transpose(1, 2),
                       torch.zeros((3, 36, 64)))])


def test_server_checks_derived_chroma_sizes():
    server = object.__new__(dh.Server)
    server._handle = None
    server._keep = {}
    server._local = threading.local()
    server._rid_local = threading.local()
    luma_in, luma_out = torch.zeros((108, 192)), torch.zeros((72, 128))
    good_in, good_out = torch.zeros((54, 96)), torch.zeros((36, 64))
    with pytest.raises(ValueError, match="input plane 1"):
        server.submit([luma_in, torch.zeros((53, 96)), good_in],
                      [luma_out, good_out, good_out])
    with pytest.raises(ValueError, match="output plane 2"):
        server.submit([luma_in, good_in, good_in], [luma_out, good_out, torch.zeros((36, 63))])
    with pytest.raises(ValueError, match="float32"):
        server.submit([luma_in.double()], [luma_out])
