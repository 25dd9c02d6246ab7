"""Synthetic problem definitions: text round trip and cross-section rules."""
import numpy as np
import pytest

from openmoc_amd import workloads as w


@pytest.mark.parametrize("name", list(w.WORKLOADS))
def test_text_round_trip(name):
    p = w.get(name)
    q = w.Problem.from_text(p.to_text())
    assert q.to_text() == p.to_text()


def test_committed_inputs_match_generator():
    from pathlib import Path
    gdir = Path(__file__).resolve().parent / "golden"
    for name in ("pin_cell_7g", "lattice_17x17_7g", "c5g7_pin_cell", "hom_inf_medium"):
        assert (gdir / f"{name}.txt").read_text() == w.get(name).to_text(), name


@pytest.mark.parametrize("G", [7, 70])
def test_synthetic_xs_rules(G):
    rng = np.random.default_rng(11)
    for fissile in (True, False):
        m = w.synthetic_material(rng, G, "m", fissile)
        s = m.sigma_s.reshape(G, G)            # [orig, dest]
        assert np.all(s.sum(axis=1) <= 0.95 * m.sigma_t + 1e-15)
        up = np.argwhere(np.tril(s, -1) > 0)   # dest < orig
        thermal = 5 if G == 7 else int(round(0.65 * G))
        assert np.all(up[:, 0] >= thermal) and np.all(up[:, 1] >= thermal)
        if fissile:
            assert abs(m.chi.sum() - 1) < 1e-12 and np.all(m.chi[3:] == 0)
            assert np.all((m.nu_sigma_f >= 0.005) & (m.nu_sigma_f <= 0.3))
        else:
            assert not m.fissionable


def test_perturbation_keeps_scatter_ratio():
    rng = np.random.default_rng(5)
    m = w.synthetic_material(rng, 7, "f", True)
    p = w.perturb_material(m, np.random.default_rng(9))
    s = p.sigma_s.reshape(7, 7)
    assert np.all(s.sum(axis=1) <= 0.95 * p.sigma_t + 1e-15)
    assert np.all(np.abs(p.sigma_t / m.sigma_t - 1) <= 0.1 + 1e-12)
