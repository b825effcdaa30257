"""Touchstone export (fem/physics/edm/touchstone.py:6-49, emdata.py:306-333): file layout and round trip."""
import numpy as np
import pytest

from emerge_amd.emdata import EMSimData
from emerge_amd.touchstone import generate_touchstone, read_touchstone


@pytest.mark.parametrize("fmt", ["RI", "MA", "DB"])
@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_round_trip(tmp_path, fmt, n):
    rng = np.random.default_rng(n)
    f = np.linspace(1e9, 3e9, 7)
    S = (rng.standard_normal((7, n, n)) + 1j * rng.standard_normal((7, n, n))) * 0.3
    out = generate_touchstone(str(tmp_path / "net"), f, S, fmt)
    assert out.endswith(f".s{n}p")
    f2, S2 = read_touchstone(out)
    assert np.allclose(f2, f, rtol=1e-12)
    assert np.abs(S2 - S).max() < 1e-10


def test_two_port_column_order(tmp_path):
    """v1 2-port rows are S11 S21 S12 S22."""
    S = np.array([[[1 + 0j, 2 + 0j], [3 + 0j, 4 + 0j]]])
    out = generate_touchstone(str(tmp_path / "x"), np.array([1e9]), S, "RI")
    row = [float(t) for t in open(out).read().splitlines()[2].split()]
    assert row[1::2] == [1.0, 3.0, 2.0, 4.0]
    assert open(out).read().splitlines()[1] == "# Hz S RI R 50"


def test_emsimdata_export(tmp_path):
    data = EMSimData()
    for k, f in enumerate([1e9, 2e9]):
        ds = data.new(freq=f)
        ds.init_sp([1, 2])
        ds.Sp.arry[:] = [[0.1 * k, 0.9j], [0.9j, 0.2]]
    out = data.export_touchstone(str(tmp_path / "d"), "MA")
    f, S = read_touchstone(out)
    assert np.allclose(S, data.S_array()) and np.allclose(f, [1e9, 2e9])


def test_validation():
    with pytest.raises(ValueError):
        generate_touchstone("x", np.zeros(2), np.zeros((2, 2, 3)), "RI")
    with pytest.raises(ValueError):
        generate_touchstone("x", np.zeros(3), np.zeros((2, 2, 2)), "RI")
