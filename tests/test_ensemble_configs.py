"""Ensemble / serving workload definitions (host only)."""
import numpy as np

from mcac_amd import ensemble


def test_ensemble_box_list():
    boxes = ensemble.ensemble_configs(512)
    assert len(boxes) == 512
    n = np.array([b["n"] for b in boxes])
    assert n.min() >= 500 and n.max() <= 50000
    assert 7000 < n.mean() < 15000  # log-uniform [500, 5e4]: E[N] ~ 10.7k
    assert [b["seed"] for b in boxes] == list(range(1000, 1512))
    fv = np.array([b["fv"] for b in boxes])
    t = np.array([b["temperature"] for b in boxes])
    assert fv.min() >= 1e-4 and fv.max() <= 5e-3 and t.min() >= 1000 and t.max() <= 2000
    assert "number_of_aggregates = %d" % (boxes[0]["n"] // 10) in boxes[0]["ini"]
    assert ensemble.ensemble_configs(8) == boxes[:8]


def test_serving_arrivals():
    reqs = ensemble.serving_requests(400, rate_per_s=2.0)
    a = np.array([r["arrival"] for r in reqs])
    assert a[0] == 0.0 and np.all(np.diff(a) > 0)
    assert 0.15 < np.diff(a).mean() < 0.85  # mean gap 1/rate
    n = np.array([r["n"] for r in reqs])
    assert n.min() >= 1000 and n.max() <= 200000
