"""Autotuner bookkeeping (CPU): decisions are the fastest candidate, cached
per key and persisted to the JSON cache file."""

import json

import pytest


def test_decide_picks_fastest_and_persists(tmp_path, monkeypatch):
    from addft import tuning
    monkeypatch.setenv("ADDFT_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(tuning, "_LOADED", False)
    monkeypatch.setattr(tuning, "_DECISIONS", {})
    times = {"a": 3.0, "b": 1.0, "c": 2.0}
    monkeypatch.setattr(tuning, "time_ms", lambda fn, reps=3: fn())
    choice = tuning.decide("k1", {k: (lambda v=v: v) for k, v in times.items()})
    assert choice == "b"
    # cached: candidates are not timed again
    assert tuning.decide("k1", {"a": lambda: pytest.fail("re-timed"), "b": lambda: pytest.fail("re-timed")}) == "b"
    data = json.loads((tmp_path / "tune.json").read_text())
    assert data["k1"]["choice"] == "b" and data["k1"]["ms"]["c"] == 2.0


def test_sliced_inner_matches_plain_product():
    import torch
    from addft.solver.linalg import inner_sliced
    g = torch.Generator().manual_seed(0)
    A = torch.randn(7, 3001, generator=g, dtype=torch.float64).to(torch.complex128)
    B = torch.randn(5, 3001, generator=g, dtype=torch.float64).to(torch.complex128)
    ref = A @ B.conj().T
    for nc in (1, 2, 7, 16):
        assert torch.allclose(inner_sliced(A, B, nc), ref, atol=1e-12)
