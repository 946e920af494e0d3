"""GPU tests of the continuous-batching request loop (dsmvc_hip_server_*).

Bars:
  * every request's result is bit-identical alone (one request in flight)
    and in a mix submitted concurrently from several threads, with F32 planes
    on a fixed route (fused "always", and "off": the windowed tables); in
    the default "warm" policy a plane moves to the fused route once its
    operator images exist, and both routes meet the oracle bars below;
  * every plane of every request, chroma included, against the oracle:
    F32-route planes within f32_bar(e_ref) of the FP64 oracle, Float64-route
    planes within 1 ULP of the oracle's Float64 route (requests derived by
    filter_requests, the reference's prepare_filter_requests);
  * a request larger than the whole memory budget is refused at submit;
    with a budget that holds only a few requests at once, every request
    still completes (admission waits for memory to free up).
"""
import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import (BICUBIC, BILINEAR, LANCZOS, SPLINE36, f32_bar, fp64_request, make_request,
                     ulp_distance)  # noqa: E402

DEV = torch.device("cuda:0")

KERNELS = [(BILINEAR, 3, 0.0, 0.5), (BICUBIC, 3, 0.0, 0.5), (LANCZOS, 3, 0.0, 0.5),
           (SPLINE36, 3, 0.0, 0.5)]


@pytest.fixture(scope="module")
def executor():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return dh.Executor(0, "auto")


def _requests(seed, count):
    """Mixed geometry: gray and 4:2:0 YUV, three source sizes, random target
    height, all kernels, all f64 modes."""
    rng = np.random.default_rng(seed)
    sources = [(1280, 720), (960, 540), (640, 360)]
    out = []
    for k in range(count):
        sw, sh = sources[rng.integers(len(sources))]
        dh_ = int(sh * rng.uniform(0.5, 0.95)) & ~1
        dw = int(np.floor(dh_ * 16 / 9 + 0.5))
        dw += dw & 1
        yuv = bool(rng.integers(2))
        kernel = KERNELS[rng.integers(len(KERNELS))]
        mode = int(rng.integers(3))
        planes = [(sw, sh, dw, dh_)]
        if yuv:
            planes += [(sw // 2, sh // 2, dw // 2, dh_ // 2)] * 2
        inputs = [torch.from_numpy(rng.uniform(0, 1, (h, w)).astype(np.float32)).to(DEV)
                  for (w, h, _, _) in planes]
        out.append({"planes": planes, "inputs": inputs, "kernel": kernel, "mode": mode})
    return out


def _submit(server, req):
    outputs = [torch.full((h, w), float("nan"), device=DEV) for (_, _, w, h) in req["planes"]]
    kernel, taps, b, c = req["kernel"]
    rid = server.submit(req["inputs"], outputs, kernel=kernel, taps=taps, b=b, c=c,
                        f64_mode=req["mode"])
    return rid, outputs


@pytest.mark.parametrize("fused", ["always", "off"])
def test_results_same_alone_and_in_the_mix(fused, executor):
    reqs = _requests(11, 24)
    server = dh.Server(executor, memory_budget=512 << 20, slots=3, fused=fused)
    alone = []
    for req in reqs:
        rid, outs = _submit(server, req)
        server.wait(rid)
        alone.append(outs)
    mixed = [None] * len(reqs)
    ids = [None] * len(reqs)

    def client(first):
        for k in range(first, len(reqs), 3):
            ids[k], mixed[k] = _submit(server, reqs[k])

    threads = [threading.Thread(target=client, args=(t,)) for t in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for rid in ids:
        server.wait(rid)
    stats = server.stats()
    print(stats, flush=True)
    assert stats["completed"] == 2 * len(reqs) and stats["failed"] == 0
    for k, (a, m) in enumerate(zip(alone, mixed)):
        for pa, pm in zip(a, m):
            assert torch.equal(pa, pm), k
    server.close()


@pytest.mark.parametrize("fused", ["warm", "always"])
def test_served_results_match_oracle_and_float64_route(fused, executor, oracle):
    reqs = _requests(12, 12)
    server = dh.Server(executor, memory_budget=512 << 20, slots=2, fused=fused)
    submitted = [_submit(server, r) for r in reqs]
    routes = set()
    for (rid, outs), req in zip(submitted, reqs):
        server.wait(rid)
        kernel, taps, b, c = req["kernel"]
        (sw, sh, dw, dh_) = req["planes"][0]
        hreq, vreq = dh.filter_requests(sw, sh, dw, dh_, len(req["planes"]), 1, 1,
                                        kernel=kernel, taps=taps, b=b, c=c, f64_mode=req["mode"])
        for index, (out, x) in enumerate(zip(outs, req["inputs"])):
            a, v = hreq[min(index, 1)], vreq[min(index, 1)]
            hr = make_request(a.source_size, a.destination_size, a.active_length, a.shift,
                              kernel, taps, b, c, f64_mode=req["mode"])
            vr = make_request(v.source_size, v.destination_size, v.active_length, v.shift,
                              kernel, taps, b, c, f64_mode=req["mode"])
            oh, ov = oracle.build(hr), oracle.build(vr)
            got = out.cpu().numpy()
            xs = x.cpu().numpy()
            if oh.c.requires_float64 or ov.c.requires_float64:
                routes.add("f64")
                assert int(ulp_distance(got, oracle.inverse_2d(oh, ov, xs)).max()) <= 1
            else:
                routes.add("f32")
                expected = oracle.inverse_2d(oh, ov, xs, ordered=True)
                ref64 = oracle.inverse_2d_f64_exact(oracle.build(fp64_request(hr)),
                                                    oracle.build(fp64_request(vr)),
                                                    xs.astype(np.float64))
                err = float(np.abs(got.astype(np.float64) - ref64).max())
                e_ref = float(np.abs(expected.astype(np.float64) - ref64).max())
                assert err <= f32_bar(e_ref), (index, err, e_ref)
    assert routes == {"f32", "f64"}
    server.close()


def test_admission_waits_for_memory(executor):
    reqs = [r for r in _requests(13, 16) if r["mode"] != 2]
    big = max(4 * ((w + 63) // 64 * 64) * h for r in reqs for (w, h, dw, _) in r["planes"])
    # Room for about two requests at a time (windowed tables: every plane
    # needs its intermediate).
    server = dh.Server(executor, memory_budget=3 * big, slots=4, fused="off")
    submitted = [_submit(server, r) for r in reqs]
    for rid, outs in submitted:
        server.wait(rid)
        assert all(bool(torch.isfinite(o).all()) for o in outs)
    assert server.stats()["memory_high_water"] <= 3 * big
    tiny = dh.Server(executor, memory_budget=4096, slots=1)
    with pytest.raises(ValueError, match="budget"):
        _submit(tiny, reqs[0])
    tiny.close()
    server.close()
