"""The CPU restatement of FractionalStepMultiphase (oracle/multiphase_oracle.py,
the checker of the device multiphase driver) against what the projection
method fixes: after each step the corrected face velocities are discretely
divergence-free (src::div(u_f) = 0 up to the direct solves: a face correction
with the wrong density, sign or time step leaves a divergence); a closed box
keeps sum V gamma; a single fluid (rho1 = rho2) starting at rest stays at rest
under gravity (grad rho = 0, so no gravity source).  With several
pressure-corrector passes per step the same holds, and the passes of one step
contract to a fixed point."""
import numpy as np
import pytest

from phase_amd import meshgen as G


def _box(oracle, nx, ny):
    xy, cptr, cind = G.rectilinear(nx, ny, width=1.0, height=2.0)
    m = oracle.Mesh(xy, cptr, cind)
    f = m.faces()
    a, b = xy[f["lnode"]], xy[f["rnode"]]
    top = (f["rcell"] < 0) & (a[:, 1] == 2.0) & (b[:, 1] == 2.0)
    pbc = np.where(top, oracle.BC_FIXED, oracle.BC_NORMAL_GRADIENT).astype(np.int32)
    ubc = np.full(m.n_faces, oracle.BC_FIXED, np.int32)
    return m, ubc, pbc


@pytest.fixture(scope="module")
def mp():
    from oracle.multiphase_oracle import MultiphaseOracle
    return MultiphaseOracle


@pytest.mark.parametrize("correctors", [1, 2])
def test_projection_and_conservation(oracle, mp, correctors):
    m, ubc, pbc = _box(oracle, 12, 24)
    vol, cc = m.cells()
    F = m.n_faces
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 1.2)
    gam = np.clip(0.5 - (r - 0.25) * 12, 0.0, 1.0)  # a heavy drop, smeared over about a cell
    s = mp(m, 1.0, 1000.0, 0.01, 1.0, (0.0, -9.81), ubc, np.zeros(2 * F), pbc, np.zeros(F), gam)
    mass0 = np.dot(vol, s.gC)
    for _ in range(3):
        div = s.step(2e-3, correctors)
        assert div <= 1e-10 * np.abs(s.uF).max()
    assert np.abs(s.uC).max() > 1e-4  # the drop falls
    assert abs(np.dot(vol, s.gC) - mass0) <= 1e-10 * mass0


@pytest.mark.parametrize("correctors", [1, 2])
def test_single_fluid_stays_at_rest(oracle, mp, correctors):
    m, ubc, pbc = _box(oracle, 6, 12)
    F = m.n_faces
    s = mp(m, 3.0, 3.0, 0.5, 0.5, (0.0, -9.81), ubc, np.zeros(2 * F), pbc, np.zeros(F),
           np.linspace(0.0, 1.0, m.n_cells))
    for _ in range(2):
        s.step(1e-2, correctors)
    assert np.all(s.uC == 0.0) and np.all(s.uF == 0.0) and np.all(s.pC == 0.0)


def test_static_drop_with_surface_tension(oracle, mp):
    """A drop at rest without gravity: the CELESTE force is balanced by the
    pressure jump; the projection still leaves divergence-free faces and the
    spurious currents stay small against the capillary velocity sigma / mu."""
    from oracle.celeste_oracle import CelesteOracle
    n = 24
    xy, cptr, cind = G.rectilinear(n, n)
    m = oracle.Mesh(xy, cptr, cind)
    F = m.n_faces
    _, cc = m.cells()
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 0.5)
    gam = 0.5 * (1.0 - np.tanh((r - 0.25) / (1.5 / n)))
    ubc = np.full(F, oracle.BC_FIXED, np.int32)
    pbc = np.full(F, oracle.BC_NORMAL_GRADIENT, np.int32)
    pbc[0] = oracle.BC_FIXED  # one wall face fixes the pressure level
    model = CelesteOracle(m, cptr, cind, kernel_width=2.5 / n)
    s = mp(m, 1.0, 1.0, 0.1, 0.1, (0.0, 0.0), ubc, np.zeros(2 * F), pbc, np.zeros(F), gam,
           surface_tension=(model, 0.05, None))
    assert np.abs(s.fstF).max() > 0
    for _ in range(3):
        div = s.step(1e-3)
        assert div <= 1e-10 * max(np.abs(s.uF).max(), 1e-30)
    assert np.abs(s.uC).max() < 1e-3
    # the pressure inside the drop exceeds the outside by about sigma / R
    dp = s.pC[r < 0.1].mean() - s.pC[r > 0.4].mean()
    assert 0.85 * 0.05 / 0.25 < dp < 1.15 * 0.05 / 0.25


def test_corrector_passes_contract(oracle, mp):
    """The falling heavy drop: within one step, the change of u from one
    pressure-corrector pass to the next at least halves (measured: 8.8x,
    then 4.0x at density ratio 1000)."""
    m, ubc, pbc = _box(oracle, 12, 24)
    _, cc = m.cells()
    F = m.n_faces
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 1.2)
    gam = np.clip(0.5 - (r - 0.25) * 12, 0.0, 1.0)
    out = []
    for k in (1, 2, 3, 4):
        s = mp(m, 1.0, 1000.0, 0.01, 1.0, (0.0, -9.81), ubc, np.zeros(2 * F), pbc, np.zeros(F), gam)
        for _ in range(3):
            s.step(2e-3)
        s.step(2e-3, k)
        out.append(s.uC)
    d = [np.abs(b - a).max() for a, b in zip(out, out[1:])]
    assert d[0] > 0.0 and d[1] <= 0.5 * d[0] and d[2] <= 0.5 * d[1], d
