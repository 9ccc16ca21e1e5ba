"""Pins of the CPU driver restatements (oracle/fstep_oracle.py,
oracle/simple_oracle.py) by what the mathematics fixes, independent of the
device drivers they check:

* FractionalStep (FractionalStep.cpp:79-117), pressure-driven Poiseuille
  channel: p = p_in on x-, 0 on x+, no-slip walls.  The steady state is
  u = G y (H - y) / (2 mu), v = 0 with G = p_in / L, and a linear pressure.
  The restatement reaches it with a second-order error (1/63 of the peak at
  8 cells across, 1/255 at 16) and an exactly linear pressure.
* FractionalStep, viscous shear mode u = sin(pi y) between no-slip walls
  (an exact Navier-Stokes solution with zero convection): its amplitude decays
  as exp(-nu pi^2 t); the theta = 0.5 viscous term makes the error second order
  in (h, dt) together.
* FractionalStep, a closed box at rest stays exactly at rest.
* SIMPLE and PISO marched to steady state on the Re = 100 cavity land on the
  centreline data of Ghia, Ghia & Shin (1982), Table I/II.
* PISO transient of the impulsively started cavity: converges (first order in
  dt) to the FractionalStep transient of the same flow.
* FractionalStep with several pressure-corrector passes per step: the
  Poiseuille and shear-mode pins hold unchanged (each pass starts from the
  same old-time fields), and on the impulsively started cavity the passes of
  one step contract to a fixed point (the change from pass to pass shrinks at
  least fourfold), as a fixed-point iteration of one implicit step must.

Each restatement takes a `plant` argument that plants one mistake (dropping
u += dt grad p, the sign of src(grad p), the viscous theta, the ddt
coefficient, old-time fields taken again before every corrector pass); the
tests show the pins above catch every one of them.
"""
import numpy as np
import pytest
from scipy.interpolate import LinearNDInterpolator

from oracle.fstep_oracle import FractionalStepOracle
from oracle.simple_oracle import SimpleOracle
from phase_amd import meshgen as G

# Ghia, Ghia & Shin (1982), Re = 100: u on x = 0.5, v on y = 0.5
GHIA_Y = [0.0547, 0.0625, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5, 0.6172, 0.7344, 0.8516, 0.9531, 0.9609,
          0.9688, 0.9766]
GHIA_U = [-0.03717, -0.04192, -0.04775, -0.06434, -0.10150, -0.15662, -0.21090, -0.20581, -0.13641, 0.00332,
          0.23151, 0.68717, 0.73722, 0.78871, 0.84123]
GHIA_X = [0.0625, 0.0703, 0.0781, 0.0938, 0.1563, 0.2266, 0.2344, 0.5, 0.8047, 0.8594, 0.9063, 0.9453, 0.9531,
          0.9609, 0.9688]
GHIA_V = [0.09233, 0.10091, 0.10890, 0.12317, 0.16077, 0.17507, 0.17527, 0.05454, -0.24533, -0.22445, -0.16914,
          -0.10313, -0.08864, -0.07391, -0.05906]


def _channel(O, nx, ny, L, H):
    xy, cptr, cind = G.rectilinear(nx, ny, width=L, height=H)
    m = O.Mesh(xy, cptr, cind)
    f = m.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy, x1=L, y1=H)
    return m, pid


def poiseuille(O, ny, plant=None, L=2.0, H=1.0, p_in=1.0, mu=1.0, dt=0.02, steps=200, correctors=1):
    """Relative velocity error, max |v| and max pressure error at steady state."""
    m, pid = _channel(O, 2 * ny, ny, L, H)
    F = m.n_faces
    ends = (pid == 0) | (pid == 1)
    ubc = np.where(ends, O.BC_NORMAL_GRADIENT, O.BC_FIXED).astype(np.int32)
    pbc = np.where(ends, O.BC_FIXED, O.BC_NORMAL_GRADIENT).astype(np.int32)
    pF = np.where(pid == 0, p_in, 0.0)
    ref = FractionalStepOracle(m, 1.0, mu, ubc, np.zeros((F, 2)), p_bc=pbc, p_face0=pF, pin=-1, plant=plant)
    with np.errstate(all="ignore"):
        for _ in range(steps):
            ref.step(dt, correctors)
    _, cc = m.cells()
    u_ex = p_in / L / (2.0 * mu) * cc[:, 1] * (H - cc[:, 1])
    p_ex = p_in * (1.0 - cc[:, 0] / L)
    return (np.abs(ref.uC[:, 0] - u_ex).max() / u_ex.max(), np.abs(ref.uC[:, 1]).max(),
            np.abs(ref.pC - p_ex).max())


def shear_decay(O, ny, dt, plant=None, mu=0.1, T=1.0, L=2.0, correctors=1):
    """Error of the decaying mode sin(pi y) exp(-nu pi^2 T), relative to its amplitude."""
    m, pid = _channel(O, 8, ny, L, 1.0)
    F = m.n_faces
    ends = (pid == 0) | (pid == 1)
    ubc = np.where(ends, O.BC_NORMAL_GRADIENT, O.BC_FIXED).astype(np.int32)
    ref = FractionalStepOracle(m, 1.0, mu, ubc, np.zeros((F, 2)), pin=0, plant=plant)
    _, cc = m.cells()
    ref.uC[:, 0] = np.sin(np.pi * cc[:, 1])
    ref.uF = O.interpolate_faces(m, ref.uC.reshape(-1), ref.uF.reshape(-1), bc_type=ref.ubc,
                                 ncomp=2).reshape(F, 2)
    with np.errstate(all="ignore"):
        for _ in range(int(round(T / dt))):
            ref.step(dt, correctors)
    amp = np.exp(-mu * np.pi ** 2 * T)
    return np.abs(ref.uC[:, 0] - amp * np.sin(np.pi * cc[:, 1])).max() / amp


def _poiseuille_ok(e8, e16):
    return e8[0] <= 0.02 and e16[0] <= 0.005 and e8[0] >= 3.5 * e16[0] and max(e8[1], e16[1]) <= 1e-12 \
        and max(e8[2], e16[2]) <= 1e-12


def _shear_ok(a, b):
    return a <= 3e-3 and b <= 8e-4 and a >= 3.5 * b


def test_fstep_poiseuille_channel(oracle):
    e8, e16 = poiseuille(oracle, 8), poiseuille(oracle, 16)
    # second order: the peak error is 1/(n^2 - 1) of the peak velocity
    assert abs(e8[0] - 1.0 / 63.0) <= 1e-9 and abs(e16[0] - 1.0 / 255.0) <= 1e-9, (e8, e16)
    assert _poiseuille_ok(e8, e16)


def test_fstep_shear_mode_decays_at_crank_nicolson_rate(oracle):
    a, b = shear_decay(oracle, 16, 0.1), shear_decay(oracle, 32, 0.05)
    assert _shear_ok(a, b), (a, b)


def test_fstep_closed_box_stays_at_rest(oracle):
    m, pid = _channel(oracle, 8, 8, 1.0, 1.0)
    F = m.n_faces
    ubc = np.full(F, oracle.BC_FIXED, np.int32)
    ref = FractionalStepOracle(m, 1.0, 0.01, ubc, np.zeros((F, 2)), pin=0)
    for _ in range(5):
        div = ref.step(0.05)
    assert div == 0.0 and np.all(ref.uC == 0.0) and np.all(ref.pC == 0.0) and np.all(ref.uF == 0.0)


@pytest.mark.parametrize("plant", ["drop_restore", "flip_src", "theta1"])
def test_fstep_pins_catch_planted_mistakes(oracle, plant):
    e8, e16 = poiseuille(oracle, 8, plant), poiseuille(oracle, 16, plant)
    a, b = shear_decay(oracle, 16, 0.1, plant), shear_decay(oracle, 32, 0.05, plant)
    assert not (_poiseuille_ok(e8, e16) and _shear_ok(a, b)), (plant, e8, e16, a, b)


def _cavity(O, n):
    xy, cptr, cind = G.rectilinear(n, n)
    om = O.Mesh(xy, cptr, cind)
    f = om.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    ubc = np.full(om.n_faces, O.BC_FIXED, np.int32)
    ub = np.zeros((om.n_faces, 2))
    ub[pid == 3, 0] = 1.0  # lid y+
    return om, ubc, ub


def _ghia_error(om, u):
    _, cc = om.cells()
    ip = LinearNDInterpolator(cc, u)
    uu = ip(np.c_[np.full(len(GHIA_Y), 0.5), GHIA_Y])[:, 0]
    vv = ip(np.c_[GHIA_X, np.full(len(GHIA_X), 0.5)])[:, 1]
    return max(np.abs(uu - GHIA_U).max(), np.abs(vv - GHIA_V).max())


def test_piso_cavity_steady_state_matches_ghia(oracle):
    om, ubc, ub = _cavity(oracle, 32)
    ref = SimpleOracle(om, 1.0, 0.01, 1.0, 1.0, ubc, ub, pin=0)
    hist = ref.piso(0.02, 300, correctors=2)
    assert hist[-1, 1] < 1e-6  # pressure corrections vanishing: steady
    # first-order upwind convection on 32x32 cells (0.025 / 0.027 from the table)
    assert _ghia_error(om, ref.uC) <= 0.03


def test_simple_cavity_converges_to_ghia(oracle):
    om, ubc, ub = _cavity(oracle, 32)
    ref = SimpleOracle(om, 1.0, 0.01, 0.7, 0.3, ubc, ub, pin=0)
    hist = ref.iterate(400)
    assert hist[-1, 1] < 1e-6
    assert _ghia_error(om, ref.uC) <= 0.03


def piso_vs_fstep(O, plant=None, n=16, T=0.5, mu=0.01):
    om, ubc, ub = _cavity(O, n)
    out = []
    for dt in (0.01, 0.005):
        k = int(round(T / dt))
        a = SimpleOracle(om, 1.0, mu, 1.0, 1.0, ubc, ub, pin=0, plant=plant)
        b = FractionalStepOracle(om, 1.0, mu, ubc, ub, pin=0)
        with np.errstate(all="ignore"):
            a.piso(dt, k, correctors=2)
            for _ in range(k):
                b.step(dt)
        out.append(np.abs(a.uC - b.uC).max())
    return out


def _piso_transient_ok(d):
    return d[0] <= 6e-3 and d[1] <= 2.5e-3 and d[0] >= 2.0 * d[1]


def test_piso_transient_converges_to_fractional_step(oracle):
    d = piso_vs_fstep(oracle)
    assert _piso_transient_ok(d), d


@pytest.mark.parametrize("plant", ["flip_src", "ddt_half"])
def test_piso_pins_catch_planted_mistakes(oracle, plant):
    d = piso_vs_fstep(oracle, plant)
    assert not _piso_transient_ok(d), (plant, d)


def corrector_contraction(O, plant=None, n=16, dt=0.02, passes=4):
    """Changes of u from pass to pass within one step of the impulsively
    started cavity (after 10 ordinary steps)."""
    om, ubc, ub = _cavity(O, n)
    ref = FractionalStepOracle(om, 1.0, 0.01, ubc, ub, pin=0, plant=plant)
    for _ in range(10):
        ref.step(dt)
    state = (ref.uC.copy(), ref.uF.copy(), ref.pC.copy(), ref.pF.copy(), ref.gpC.copy(), ref.gpF.copy())
    out = []
    for k in range(1, passes + 1):
        ref.uC, ref.uF, ref.pC, ref.pF, ref.gpC, ref.gpF = (a.copy() for a in state)
        ref.step(dt, correctors=k)
        out.append(ref.uC.copy())
    return [np.abs(b - a).max() for a, b in zip(out, out[1:])]


def _contraction_ok(d):
    return d[0] > 0.0 and d[1] <= 0.25 * d[0] and d[2] <= 0.25 * d[1]


def test_fstep_corrector_passes_keep_the_pins(oracle):
    e8, e16 = poiseuille(oracle, 8, correctors=2), poiseuille(oracle, 16, correctors=2)
    assert _poiseuille_ok(e8, e16), (e8, e16)
    a, b = shear_decay(oracle, 16, 0.1, correctors=2), shear_decay(oracle, 32, 0.05, correctors=2)
    assert _shear_ok(a, b), (a, b)


def test_fstep_corrector_passes_contract(oracle):
    d = corrector_contraction(oracle)
    assert _contraction_ok(d), d


def test_fstep_corrector_pins_catch_planted_mistake(oracle):
    a, b = shear_decay(oracle, 16, 0.1, "save_per_corrector", correctors=2), \
        shear_decay(oracle, 32, 0.05, "save_per_corrector", correctors=2)
    d = corrector_contraction(oracle, "save_per_corrector")
    assert not _shear_ok(a, b) and not _contraction_ok(d), (a, b, d)
