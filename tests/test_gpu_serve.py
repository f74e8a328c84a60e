"""The serving loop (``serve.Server``): every result the same alone and in the mix (bitwise),
served results equal to the plain per-request path, admission within the HBM budget, and the
latency statistics.

Requests: a Poisson stream (``workloads.request_stream``) of evaluations and 60-step NVT segments
on 256- to 4096-atom systems at 21 nm^-3 (edges stretched to at least 2.45 nm, above twice the
table's cutoff), LJ and 200-knot tables, fp32. Tolerance against the plain path: evaluations
max|dF| / max|F| <= 2e-5 (the batch and the drop-in list sum in different orders); MD segments
bitwise (the same simulation code and schedule).
"""

import pytest
import torch

from nabladyn_amd import serve, workloads

pytestmark = pytest.mark.gpu


def _specs(count=12, seed=11, qps=2000.0):
    return workloads.request_stream(count, qps, seed=seed, low=256, high=4096, md_fraction=0.4,
                                    md_steps=60)


def _server(device, **kw):
    kw.setdefault("md_chunk", 25)
    kw.setdefault("md_streams", 4)
    return serve.Server(device, **kw)


def _equal(a: dict, b: dict) -> bool:
    return a.keys() == b.keys() and all(torch.equal(a[k], b[k]) for k in a)


def test_same_alone_and_in_the_mix(device):
    specs = _specs()
    assert {s["kind"] for s in specs} == {"efw", "md"}
    mix = serve.make_requests(specs, device)
    stats = _server(device).run(mix)
    assert stats["requests"] == len(specs)
    for spec, served in zip(specs, mix):
        alone = serve.make_requests([dict(spec, arrival=0.0)], device)
        _server(device).run(alone)
        assert _equal(served.result, alone[0].result), (spec["kind"], spec["atoms"])


def test_served_results_equal_the_plain_path(device):
    specs = _specs(count=8, seed=12)
    served = serve.make_requests(specs, device)
    _server(device).run(served)
    plain = serve.make_requests(specs, device)
    serve.serve_plain(plain)
    for s, p in zip(served, plain):
        if s.kind == "md":
            assert _equal(s.result, p.result)
        else:
            scale = float(p.result["forces"].abs().max())
            assert float((s.result["forces"] - p.result["forces"]).abs().max()) <= 2e-5 * scale
            wscale = float(p.result["virial"].abs().max())
            assert float((s.result["virial"] - p.result["virial"]).abs().max()) <= 2e-5 * wscale
            e = float(p.result["energy"])
            assert abs(float(s.result["energy"]) - e) <= 2e-5 * max(1.0, abs(e)) * 10


def test_admission_keeps_within_the_budget(device):
    specs = _specs(count=10, seed=13, qps=1e6)  # all arrive at once
    requests = serve.make_requests(specs, device)
    largest = max(serve.REQUEST_BYTES + serve.MD_BYTES_PER_ATOM * r.atoms for r in requests)
    budget = 2 * largest
    server = _server(device, hbm_budget=budget)
    stats = server.run(requests)
    assert stats["requests"] == len(requests)
    assert server.peak_resident_bytes <= budget
    assert server.resident_bytes == 0
    assert stats["rounds"] > 1
    assert stats["p99_latency_ms"] >= stats["p50_latency_ms"] > 0


def test_oversized_and_malformed_requests_are_refused(device):
    request = serve.make_requests(_specs(count=1, seed=14), device)[0]
    with pytest.raises(ValueError):
        _server(device, hbm_budget=1024).submit(request)
    request.kind = "bogus"
    with pytest.raises(ValueError):
        _server(device).submit(request)


def test_full_stream_alone_equals_in_the_mix(device):
    """The serving workload at its stated size: the 64-request Poisson stream the bench serves
    (seed 4, 40 QPS, 256 .. 262,144 atoms, 30% 500-step NVT segments, LJ or 200-knot tables,
    fp32). Every request's result is bitwise the same served in the mix and served alone; the
    largest LJ and table evaluations are within 1e-5 of the reference's fp32 route on the
    oracle's pairs."""
    import numpy as np

    from oracle import neighbors as on
    from oracle import potentials as op
    from _parity import check_against_route, lj_route, table_route

    specs = workloads.request_stream(64, 40.0, seed=4)
    assert max(s["atoms"] for s in specs) > 200_000
    assert sum(s["kind"] == "md" for s in specs) >= 10
    mix = serve.make_requests(specs, device)
    stats = serve.Server(device).run(mix)
    assert stats["requests"] == len(specs)
    print(f"served 64: p50 {stats['p50_latency_ms']:.1f} ms, p99 {stats['p99_latency_ms']:.1f} ms",
          flush=True)
    for spec, served in zip(specs, mix):
        alone = serve.make_requests([dict(spec, arrival=0.0)], device)
        serve.Server(device).run(alone)
        assert _equal(served.result, alone[0].result), (spec["kind"], spec["atoms"])
    for kind in ("lj", "table"):
        k = max((i for i, s in enumerate(specs) if s["kind"] == "efw" and s["potential"] == kind),
                key=lambda i: specs[i]["atoms"])
        request = mix[k]
        x = request.positions.cpu().numpy()
        cell = request.cell.cpu().numpy()
        xw = on.wrap(x, cell)[0]
        b = on.build(xw, cell, request.model.cutoff)
        pairs = b.pairs[:, :b.pair_total]
        types = np.zeros(len(x), np.int64)
        m = request.model
        if kind == "lj":
            params = (m.sigma.detach().cpu().numpy(), m.epsilon.detach().cpu().numpy())
            fp64 = op.lj_efw(xw.astype(np.float64), cell.astype(np.float64), pairs, types,
                             *params, m.cutoff)
            route = lj_route(xw, cell, pairs, types, *params, m.cutoff)
        else:
            params = (m.table.detach().cpu().numpy(), m.start.cpu().numpy())
            spacing = m.spacing.cpu().numpy()
            fp64 = op.table_efw(xw.astype(np.float64), cell.astype(np.float64), pairs, types,
                                *params, m.cutoff, spacing=spacing)
            route = table_route(xw, cell, pairs, types, *params, spacing, m.cutoff)
        got = (float(request.result["energy"]), request.result["forces"].cpu().numpy(),
               request.result["virial"].cpu().numpy())
        check_against_route(f"served {kind} evaluation, {len(x)} atoms", got, route, fp64)
