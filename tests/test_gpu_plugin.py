"""GPU tests of the plugin-shaped operators (dsmvc_hip.Descale and friends).

  * Descale with a custom kernel (the bicubic polynomial as a callable,
    taps = 2) equals Debicubic bit for bit;
  * argument errors keep the reference's messages (vs_plugin.cpp:362-509);
  * integer clips descaled on one axis (the reference converts them to
    float32, descales and converts back, vs_plugin.cpp:570-626) are within
    1 code of the oracle's F32 route on the normalised samples, rounded the
    same way.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import make_request  # noqa: E402
from test_custom_kernel import _bicubic  # noqa: E402
from workloads import integer_conversion  # noqa: E402

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def device():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return DEV


def test_custom_kernel_descale_equals_debicubic(device):
    x = torch.rand(2, 540, 960, device=device)
    expected = dh.Debicubic(x, 640, 360, b=1 / 3, c=1 / 3)
    got = dh.Descale(x, 640, 360, custom_kernel=_bicubic(1 / 3, 1 / 3), taps=2, b=1 / 3,
                     c=1 / 3)
    assert torch.equal(got, expected)
    legacy = dh.Descale(x, 640, 360, custom=_bicubic(1 / 3, 1 / 3), support=2)
    assert torch.equal(legacy, expected)


def test_argument_errors(device):
    x = torch.rand(108, 192, device=device)
    cases = [
        (dict(kernel="bicubic", custom_kernel=lambda v: 1.0), "not both"),
        (dict(), "kernel or custom kernel is required"),
        (dict(custom_kernel=lambda v: 1.0), "require taps or support"),
        (dict(kernel="lanczos", taps=0), "taps must be greater than zero"),
        (dict(kernel="nearest"), "kernel is invalid"),
        (dict(kernel="bicubic", blur=200.0), "blur exceeds"),
        (dict(kernel="bicubic", src_width=0.0), "src_width and src_height must be positive"),
    ]
    for kw, message in cases:
        with pytest.raises(ValueError, match=message):
            dh.Descale(x, 128, 72, **kw)
    with pytest.raises(ValueError, match="must not exceed"):
        dh.Descale(x, 200, 72, kernel="bicubic")


def test_single_axis_integer_descale(device, oracle):
    rng = np.random.default_rng(50)
    y = rng.integers(0, 65536, (1, 216, 384)).astype(np.uint16)
    to = lambda a: torch.from_numpy(a.astype(np.int32)).to(device).to(torch.uint16)  # noqa: E731
    conv = integer_conversion(16, chroma=False, full_range=False)
    for width, height, rows in ((256, 216, True), (384, 144, False)):
        out = dh.Debicubic(dh.Clip([to(y)], "gray", 0, 0, 16), width, height)
        got = out.planes[0][0].to(torch.int32).cpu().numpy()
        xf = (y[0].astype(np.float32) - np.float32(conv.input_offset)) \
            * np.float32(conv.input_scale)
        if rows:
            plan = oracle.build(make_request(384, 256))
            solved = oracle.rows(plan, xf)
        else:
            plan = oracle.build(make_request(216, 144))
            solved = oracle.columns(plan, xf)
        v = np.clip(solved * np.float32(conv.output_scale) + np.float32(conv.output_offset),
                    0, conv.output_maximum)
        expected = np.rint(v).astype(np.int32)
        diff = int(np.abs(got - expected).max())
        print(f"single-axis u16 ({'rows' if rows else 'columns'}): max code diff {diff}",
              flush=True)
        assert diff <= 1
