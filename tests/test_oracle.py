"""Pins of the CPU oracle (oracle/pbf_oracle.cpp) against what the reference
and the mathematics fix, so that the oracle can then check the HIP path.

The reference ships no golden vectors or tests (its only fixtures are the
input scenarios, copied under tests/golden/).  What pins the oracle here:
  * the kernel constants against the -D strings src/main.cpp:149-168
    produces for h = 0.1 (6-significant-digit formatting of the closed
    forms);
  * perfect-lattice closed forms of rho and lambda in both physics modes;
  * the two-particle scene of assets/scenarios/two.{par,in}, reduced to a
    scalar recurrence on the separation and worked in FP64;
  * brute-force O(N^2) neighbour counts, numpy cell ids and a numpy stable
    sort against the oracle's linked-cell search;
  * antisymmetry of the pressure term (sum of displacements = 0 away from
    walls), mirror symmetry of a lattice, wall / wave-generator clamps,
    the orientation of the vorticity of a rigid rotation, FP32 vs FP64;
  * the PAPER-only terms against closed forms of Macklin & Mueller's
    equations: s_corr through a two-particle scalar recurrence, the
    vorticity magnitude of a rigid rotation over the lattice shells, the
    confinement force on three particles with a known |omega| gradient, and
    the obstacle projection with its order relative to the wall clamp.
"""
import math
import os

import numpy as np
import pytest

import pbf
import pbf_oracle as orc
import scenarios


def fmt6(v: float) -> np.float32:
    """The float a '-DNAME=<%#g>f' flag of src/main.cpp:149-168 denotes."""
    return np.float32(float(f"{v:#g}"))


def params(box=(1.0, 1.0, 1.0), n=(10, 10, 10), physics=pbf.PHYSICS_REF, **kw):
    p = pbf.Params()
    p.physics = physics
    p.iterations = 4
    p.gravity_y = -9.81
    p.xsph_c = 0.01
    if physics == pbf.PHYSICS_REF:
        p.relaxation = 1e4
    else:
        p.relaxation = 600.0
        p.scorr_k, p.scorr_n, p.scorr_dq = 0.001, 4.0, 0.2
        p.vorticity_eps = 1e-4
    p.time_step_length = np.float32(1.0 / 120.0)
    p.x_min = p.y_min = p.z_min = 0.0
    p.x_max, p.y_max, p.z_max = box
    p.x_n, p.y_n, p.z_n = (float(v) for v in n)
    p.rest_density = 6378.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ----------------------------------------------------------------------------
# constants
# ----------------------------------------------------------------------------

def test_ref_constants_match_reference_flags():
    # box 2 x 1 x 1, 20 x 10 x 10 cells: the flags src/main.cpp:149-168 emits
    p = params(box=(2.0, 1.0, 1.0), n=(20, 10, 10))
    c = orc.constants(p)
    expect = {"h": "0.100000", "h2": "0.0100000", "poly6": "1.56668e+09",
              "spiky": "1.43239e+07", "dt": "0.00833333", "rho0": "6378.00",
              "cl_x": "0.100000", "max_x": "2.00000", "n_x": "20.0000"}
    for k, s in expect.items():
        assert c[k] == np.float32(float(s)), (k, c[k], s)


def test_paper_constants_exact():
    p = params(box=(2.0, 1.0, 1.0), n=(20, 10, 10), physics=pbf.PHYSICS_PAPER)
    c = orc.constants(p)
    h = 2.0 / 20.0
    assert c["poly6"] == np.float32(315.0 / (64.0 * math.pi * h ** 9))
    assert c["spiky"] == np.float32(45.0 / (math.pi * h ** 6))
    assert c["dt"] == np.float32(1.0 / 120.0)


# ----------------------------------------------------------------------------
# perfect lattice closed forms
# ----------------------------------------------------------------------------

# interior shells of a cubic lattice at spacing h/2 inside the support:
# (count, q^2 / h^2); the 6 sites at q = h are excluded by r^2 < h^2
SHELLS = [(6, 0.25), (12, 0.5), (8, 0.75)]


def _lattice_state(h=0.125, n=9, origin=0.5):
    s = h / 2
    idx = np.arange(n)
    g = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1).reshape(-1, 3)
    pos = np.zeros((g.shape[0], 4), np.float32)
    pos[:, :3] = origin + g * s   # exact binary fractions
    vel = np.zeros_like(pos)
    vel[:, 3] = 1.0
    centre = int(np.flatnonzero((g == n // 2).all(1))[0])
    return pos, vel, centre


@pytest.mark.parametrize("physics", [pbf.PHYSICS_REF, pbf.PHYSICS_PAPER])
def test_lattice_density_and_lambda_closed_form(physics):
    h = 0.125
    p = params(box=(2.0, 2.0, 2.0), n=(16, 16, 16), physics=physics, gravity_y=0.0)
    pos, vel, ci = _lattice_state(h)
    out = orc.probe_f32(p, pos, vel)
    rho0 = 6378.0
    if physics == pbf.PHYSICS_REF:
        poly6 = float(fmt6(315.0 / (64.0 * math.pi * h ** 9)))
        spiky = float(fmt6(45.0 / (math.pi * h ** 6)))
        hh, h2 = float(fmt6(h)), float(fmt6(h * h))
    else:
        poly6 = 315.0 / (64.0 * math.pi * h ** 9)
        spiky = 45.0 / (math.pi * h ** 6)
        hh, h2 = h, h * h
    rho = sum(c * poly6 * (h2 - f * h * h) ** 3 for c, f in SHELLS)
    grad = [(c, spiky * (hh - math.sqrt(f) * h) ** 2) for c, f in SHELLS]
    C = rho / rho0 - 1.0
    if physics == pbf.PHYSICS_REF:
        sum_abs = sum(c * g for c, g in grad)          # |sum grad W| = 0
        lam = -C / (sum_abs ** 2 / rho0 ** 2 + 1e4)
    else:
        sum_sq = sum(c * g * g for c, g in grad)
        lam = -C / (sum_sq / rho0 ** 2 + 600.0)
    assert out["nbr_count"][ci] == 26
    assert out["rho"][ci] == pytest.approx(rho, rel=2e-6)
    assert out["lambda"][ci] == pytest.approx(lam, rel=2e-5)


def test_lattice_numbers_h01():
    """h = 0.1: rho (self excluded) = 6511.52 and lambda_paper = -2.507166e-5
    at the centre of the lattice.  The spacing 0.05 is not a binary fraction,
    so three of the six sites at q = h (those on the + side of each axis)
    land just inside the support in FP32 (q = 0.0999999): the centre has 29
    neighbours, counted here by brute force on the FP32 coordinates, and the
    three extra pairs add (h^2 - q^2)^3 ~ 1e-21 to rho and (h - q)^4 ~ 1e-28
    to the lambda denominator, so the closed forms of the 26-neighbour
    interior still hold to FP32 rounding."""
    h = 0.1
    poly6 = 315.0 / (64.0 * math.pi * h ** 9)
    rho = sum(c * poly6 * (h * h - f * h * h) ** 3 for c, f in SHELLS)
    assert rho == pytest.approx(6511.5199, rel=1e-6)
    p = params(box=(4.0, 4.0, 4.0), n=(40, 40, 40), physics=pbf.PHYSICS_PAPER,
               gravity_y=0.0)
    pos, vel, ci = _lattice_state(h=0.1, n=9, origin=1.0)
    out = orc.probe_f32(p, pos, vel)
    # brute force on the FP32 coordinates: r^2 left to right, 0 < r^2 < h^2
    r = (pos[ci, :3] - pos[:, :3]).astype(np.float32)
    r2 = (r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]
    inside = (r2 > 0) & (r2 < np.float32(orc.constants(p)["h2"]))
    q = np.sqrt(r2[inside].astype(np.float64))
    assert inside.sum() == 29 and (np.abs(q - h) < 1e-6).sum() == 3
    assert out["nbr_count"][ci] == 29
    assert out["rho"][ci] == pytest.approx(6511.5199, rel=1e-5)
    assert out["lambda"][ci] == pytest.approx(-2.507166e-5, rel=1e-4)


def test_lattice_centre_stays_put_fp64():
    # rest density = the lattice's own density, so the block only relaxes
    h = 0.125
    poly6 = 315.0 / (64.0 * math.pi * h ** 9)
    rho_lattice = sum(c * poly6 * (h * h - f * h * h) ** 3 for c, f in SHELLS)
    p = params(box=(2.0, 2.0, 2.0), n=(16, 16, 16), physics=pbf.PHYSICS_PAPER,
               gravity_y=0.0, rest_density=rho_lattice)
    # a quarter spacing off the cell faces: the particles never change cell,
    # so every candidate set is mirror-symmetric about the centre
    pos, vel, ci = _lattice_state(h, origin=0.53125)
    x, v = orc.run_f64(p, pos, vel, 3)
    assert np.abs(x[ci, :3] - pos[ci, :3]).max() < 1e-12
    assert np.abs(v[ci, :3]).max() < 1e-9
    # the surface particles did move: the test is not vacuous
    assert np.abs(x[:, :3] - pos[:, :3]).max() > 1e-6


# ----------------------------------------------------------------------------
# two particles (assets/scenarios/two.{par,in})
# ----------------------------------------------------------------------------

def _two_params():
    p = params(box=(1.0, 1.0, 1.0), n=(100, 100, 100))
    p.time_step_length = 0.005
    p.rest_density = 998200.0
    return p


def test_two_particle_scene_scalar_recurrence(golden_dir):
    parts = np.loadtxt(os.path.join(golden_dir, "two.in"), skiprows=1)
    pos, vel = orc.state_from_particles(np.pad(parts, ((0, 0), (0, 3))), np.float64)
    p = _two_params()
    x, v = orc.run_f64(p, pos, vel, 1)

    # the same step reduced to the separation d along x (ref physics), in
    # the constants the kernels see
    h = float(fmt6(0.01))
    h2 = float(fmt6(0.01 ** 2))
    poly6 = float(fmt6(315.0 / (64.0 * math.pi * float(np.float32(0.01)) ** 9)))
    spiky = float(fmt6(45.0 / (math.pi * float(np.float32(0.01)) ** 6)))
    dt = float(fmt6(np.float32(0.005)))
    rho0 = float(fmt6(998200.0))
    d = float(pos[0, 0] - pos[1, 0])
    vy = float(np.float32(-9.81)) * dt   # the kernel's -9.81f literal
    rho = 0.0
    for _ in range(4):
        rho = poly6 * (h2 - d * d) ** 3
        g = spiky * (h - d) ** 2
        lam = -(rho / rho0 - 1.0) / ((2 * g) ** 2 / rho0 ** 2 + 1e4)
        step = 2 * lam * g / rho0      # |dp| of each particle, toward the other
        d -= 2 * step
    # both particles: equal and opposite x displacement, common fall in y
    mid = 0.5 * (pos[0, 0] + pos[1, 0])
    assert x[0, 0] == pytest.approx(mid + d / 2, abs=1e-13)
    assert x[1, 0] == pytest.approx(mid - d / 2, abs=1e-13)
    assert x[0, 1] == pytest.approx(0.5 + dt * vy, abs=1e-13)
    # velocity (x* - x)/dt, then XSPH with the last lambda pass's density
    vx0 = (x[0, 0] - pos[0, 0]) / dt
    w = poly6 * (h2 - d * d) ** 3
    dvx = float(np.float32(0.01)) * (1.0 / rho) * (-vx0 - vx0) * w
    assert v[0, 0] == pytest.approx(vx0 + dvx, rel=1e-11)
    assert v[1, 0] == pytest.approx(-(vx0 + dvx), rel=1e-11)
    assert v[0, 1] == pytest.approx(vy, rel=1e-12)
    assert x[0, 3] == pytest.approx(math.hypot(v[0, 0], v[0, 1]), rel=1e-12)
    # rho < rho0: the bilateral constraint pulls the pair together
    assert d < pos[0, 0] - pos[1, 0]


# ----------------------------------------------------------------------------
# neighbour search: cell ids, stable sort, stencil counts
# ----------------------------------------------------------------------------

@pytest.fixture(scope="module")
def jittered():
    sc = scenarios.small_dam(dims=(16, 16, 16), box=(2.0, 2.0, 2.0),
                             physics=pbf.PHYSICS_REF, seed=11)
    pos, vel = orc.state_from_particles(sc.particles[0])
    return sc.params[0], pos, vel


def test_cell_ids_sort_and_table(jittered):
    p, pos, vel = jittered
    out = orc.probe_f32(p, pos, vel, 20 ** 3)
    pred = out["pred"]
    # predict in float32, as predict_positions.cl:8-9
    dt = fmt6(p.time_step_length)
    vy = np.float32(vel[:, 1] + dt * np.float32(-9.81))
    assert np.array_equal(pred[:, 1], np.float32(pos[:, 1] + dt * vy))
    cl = fmt6(0.1)
    cell = np.zeros(len(pos), np.int64)
    stride = 1
    for a in range(3):
        c = np.trunc((pred[:, a] - np.float32(0.0)) / cl).astype(np.int64)
        cell += np.clip(c, 0, 19) * stride
        stride *= 20
    assert np.array_equal(out["cell_of"], cell)
    order = np.argsort(cell, kind="stable")
    assert np.array_equal(out["sorted_ids"], order)
    starts = np.searchsorted(cell[order], np.arange(20 ** 3 + 1), side="left")
    assert np.array_equal(out["cell_start"], starts)


def test_stencil_counts_equal_brute_force(jittered):
    p, pos, vel = jittered
    out = orc.probe_f32(p, pos, vel, 20 ** 3)
    brute = orc.brute_neighbors(out["pred"], float(fmt6(0.01)))
    assert np.array_equal(out["nbr_count"], brute)
    assert brute.mean() > 20  # a dense block, not a vacuous comparison


# ----------------------------------------------------------------------------
# invariants of the step
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("physics", [pbf.PHYSICS_REF, pbf.PHYSICS_PAPER])
def test_pressure_term_is_antisymmetric(physics):
    sc = scenarios.small_dam(dims=(10, 10, 10), box=(2.0, 2.0, 2.0),
                             physics=physics, seed=5)
    p = sc.params[0]
    p.gravity_y = 0.0
    pos, vel = orc.state_from_particles(sc.particles[0], np.float64)
    pos[:, :3] += 0.5  # away from every wall
    x, _ = orc.run_f64(p, pos, vel, 1)
    d = x[:, :3] - pos[:, :3]
    assert np.abs(d).sum() > 1e-4
    assert np.abs(d.sum(0)).max() < 1e-12 * max(1.0, np.abs(d).sum())


def test_wall_and_wave_clamp():
    p = params(box=(1.0, 1.0, 1.0), n=(10, 10, 10), gravity_y=0.0)
    parts = np.zeros((3, 10), np.float32)
    parts[:, 0] = 1.0
    parts[0, 1:4] = [0.5, 0.02, 0.5]   # below the floor band
    parts[1, 1:4] = [0.97, 0.5, 0.5]   # beyond x_max - h
    parts[2, 1:4] = [0.25, 0.5, 0.5]   # inside x_min + wave + h for wave 0.2
    pos, vel = orc.state_from_particles(parts)
    x, _ = orc.run_f32(p, pos, vel, 1, wave=0.2)
    h = fmt6(0.1)
    assert x[0, 1] == np.float32(np.float32(0.0) + h)
    assert x[1, 0] == np.float32(np.float32(1.0) - h)
    assert x[2, 0] == np.float32(np.float32(np.float32(0.0) + np.float32(0.2)) + h)


def test_vorticity_of_rigid_rotation_points_along_axis():
    p = params(box=(2.0, 2.0, 2.0), n=(16, 16, 16), physics=pbf.PHYSICS_PAPER)
    pos, _, ci = _lattice_state()
    c = pos[ci, :3].astype(np.float64)
    omega = np.array([0.3, -0.2, 1.0])
    vel = np.zeros((pos.shape[0], 4))
    vel[:, :3] = np.cross(omega, pos[:, :3] - c)
    w = orc.vorticity_f64(p, pos.astype(np.float64), vel)[ci, :3]
    cosang = w @ omega / (np.linalg.norm(w) * np.linalg.norm(omega))
    assert cosang > 1 - 1e-9


@pytest.mark.parametrize("physics", [pbf.PHYSICS_REF, pbf.PHYSICS_PAPER])
def test_fp32_tracks_fp64(physics):
    sc = scenarios.small_dam(dims=(10, 10, 10), box=(1.0, 1.0, 1.0),
                             physics=physics, seed=3)
    p = sc.params[0]
    pos, vel = orc.state_from_particles(sc.particles[0])
    x32, v32 = orc.run_f32(p, pos, vel, 10)
    x64, v64 = orc.run_f64(p, pos, vel, 10)
    assert np.abs(x32[:, :3] - x64[:, :3]).max() < 2e-5
    assert np.abs(v32[:, :3] - v64[:, :3]).max() < 5e-3 * max(1.0, np.abs(v64).max())


def test_literal_scorr_k01_diverges():
    """s_corr k = 0.1 (with eps 600, rho0 6378, h 0.1, m 1) pushes the
    surface layer ~0.03 m per iteration and diverges; k = 0.001 does not.
    This is why PAPER physics defaults to k = 0.001."""
    sc = scenarios.small_dam(dims=(12, 12, 12), box=(2.0, 2.0, 2.0),
                             physics=pbf.PHYSICS_PAPER, seed=3)
    p = sc.params[0]
    pos, vel = orc.state_from_particles(sc.particles[0], np.float64)
    p.scorr_k = 0.1
    _, v = orc.run_f64(p, pos, vel, 5)
    assert not np.all(np.abs(v[:, :3]) < 100.0)
    p.scorr_k = 0.001
    _, v = orc.run_f64(p, pos, vel, 5)
    assert np.all(np.abs(v[:, :3]) < 5.0)


def _zero_density_pair():
    """Two particles 0.16 m apart (beyond h = 0.1), one 0.02 m from the x-min
    wall; one iteration, no gravity.  In the only lambda pass both have no
    neighbour (density 0); the delta pass clamps the wall particle to
    x_min + h, within h of the other, before the XSPH pass."""
    parts = np.zeros((2, 10), np.float32)
    parts[:, 0] = 1.0
    parts[0, 1:4] = [0.02, 0.5, 0.5]
    parts[1, 1:4] = [0.18, 0.5, 0.5]
    return parts


def test_ref_xsph_divides_by_a_zero_density():
    """REF keeps the reference's unguarded 1/rho_j of the XSPH sum
    (apply_vorticity_and_viscosity.cl:61): a neighbour that had no pair in
    the last lambda pass but is within h afterwards gives an infinite
    velocity.  This is what ends a long REF run of the 1M-particle dam break
    (first at step 82).  PAPER skips such a neighbour and stays finite."""
    parts = _zero_density_pair()
    p = params(gravity_y=0.0, iterations=1)
    pos, vel = orc.state_from_particles(parts)
    x, v = orc.run_f32(p, pos, vel, 1)
    # clamped to x_min + h, through the update x* + (future - x*)
    x0, fut = np.float32(0.02), np.float32(np.float32(0.0) + fmt6(0.1))
    assert x[0, 0] == np.float32(x0 + np.float32(fut - x0))
    assert np.isfinite(x[:, :3]).all()
    assert not np.isfinite(v[:, 0]).any()
    q = params(physics=pbf.PHYSICS_PAPER, gravity_y=0.0, iterations=1)
    x, v = orc.run_f64(q, pos.astype(np.float64), vel.astype(np.float64), 1)
    assert np.isfinite(x).all() and np.isfinite(v).all()


# ----------------------------------------------------------------------------
# PAPER-only physics: s_corr, vorticity, vorticity confinement, obstacle.
# The reference has none of these (s_corr is a commented sketch at
# src/kernels/compute_delta.cl:65-76, apply_vorticity_and_viscosity.cl only
# does XSPH, the clamp of compute_delta.cl:128-150 knows only walls), so they
# are pinned against closed forms of Macklin & Mueller's equations.
# ----------------------------------------------------------------------------

def _paper_pair(d, k, n, iterations):
    """Two particles a distance d apart along x at the centre of a 2 m box,
    at rest, no gravity: PAPER physics with s_corr (k, n, dq = 0.2h)."""
    p = params(box=(2.0, 2.0, 2.0), n=(20, 20, 20), physics=pbf.PHYSICS_PAPER,
               gravity_y=0.0, iterations=iterations, scorr_k=k,
               scorr_n=float(n))
    pos = np.zeros((2, 4))
    vel = np.zeros((2, 4))
    pos[0, :3] = [1.0 + d / 2, 1.03, 0.97]
    pos[1, :3] = [1.0 - d / 2, 1.03, 0.97]
    vel[:, 3] = 1.0
    return p, pos, vel


@pytest.mark.parametrize("iterations", [1, 4])
@pytest.mark.parametrize("n", [2, 4])
@pytest.mark.parametrize("k", [1e-3, 0.1])
def test_scorr_two_particle_recurrence(k, n, iterations):
    """s_corr = -k (W(d) / W(dq))^n added to lambda_i + lambda_j (eqs. 13,
    14), reduced to a scalar recurrence on the separation d: with
    g = spiky (h - d)^2 the gradient magnitude, each iteration moves each
    particle by (2 lambda + s_corr) g / rho0 toward the other, lambda from the
    paper's denominator (|sum grad W|^2 + sum |grad W|^2)/rho0^2 + eps =
    2 g^2 / rho0^2 + eps.  Then v = (x* - x)/dt, XSPH with the density of the
    last lambda pass, and no vorticity (v parallel to r)."""
    h = 0.1
    poly6 = 315.0 / (64.0 * math.pi * h ** 9)
    spiky = 45.0 / (math.pi * h ** 6)
    rho0, eps, dt = 6378.0, 600.0, float(np.float32(1.0 / 120.0))
    # the parameters travel as float32 (pbf_params)
    dq = float(np.float32(0.2)) * h
    w_dq = poly6 * (h * h - dq * dq) ** 3
    p, pos, vel = _paper_pair(0.5 * h, k, n, iterations)
    strong = k == 0.1
    k = float(np.float32(k))
    x, v = orc.run_f64(p, pos, vel, 1)

    d0 = pos[0, 0] - pos[1, 0]
    d = d0
    rho = 0.0
    for _ in range(iterations):
        if d >= h:          # the pair has left the support: no terms
            rho = 0.0
            continue
        w = poly6 * (h * h - d * d) ** 3
        g = spiky * (h - d) ** 2
        rho = w
        lam = -(rho / rho0 - 1.0) / (2.0 * g * g / rho0 ** 2 + eps)
        scorr = -k * (w / w_dq) ** n
        d -= 2.0 * (2.0 * lam + scorr) * g / rho0
    mid = 0.5 * (pos[0, 0] + pos[1, 0])
    assert x[0, 0] == pytest.approx(mid + d / 2, abs=1e-13)
    assert x[1, 0] == pytest.approx(mid - d / 2, abs=1e-13)
    assert np.array_equal(x[:, 1:3], pos[:, 1:3])
    vx = (x[0, 0] - pos[0, 0]) / dt
    dvx = 0.0
    if d < h and rho > 0.0:   # XSPH, weighted by the last pass's density
        wf = poly6 * (h * h - d * d) ** 3
        dvx = float(np.float32(0.01)) * (1.0 / rho) * (-2.0 * vx) * wf
    assert v[0, 0] == pytest.approx(vx + dvx, rel=1e-11)
    assert v[1, 0] == pytest.approx(-(vx + dvx), rel=1e-11)
    # what the equations fix whatever the algebra: at d = h/2 the pair is
    # under-dense (lambda > 0 pulls it together); the artificial pressure
    # s_corr is repulsive, and at k = 0.1 it outweighs 2 lambda in the first
    # iteration (later ones start from the wider, more under-dense pair)
    if iterations == 1:
        assert (d > d0) if strong else (d < d0)


def test_scorr_is_what_separates_the_pair():
    """The same pair with k = 0 and k = 0.1 (n = 4): the difference of the
    separations after one iteration is 2 k (W/W(dq))^4 g / rho0 exactly."""
    h = 0.1
    poly6 = 315.0 / (64.0 * math.pi * h ** 9)
    spiky = 45.0 / (math.pi * h ** 6)
    p0, pos, vel = _paper_pair(0.4 * h, 0.0, 4, 1)
    p1, _, _ = _paper_pair(0.4 * h, 0.1, 4, 1)
    x0, _ = orc.run_f64(p0, pos, vel, 1)
    x1, _ = orc.run_f64(p1, pos, vel, 1)
    d = 0.4 * h
    dq = float(np.float32(0.2)) * h
    q = (h * h - d * d) ** 3 / (h * h - dq * dq) ** 3
    expect = 2.0 * float(np.float32(0.1)) * q ** 4 * spiky * (h - d) ** 2 / 6378.0
    sep0 = x0[0, 0] - x0[1, 0]
    sep1 = x1[0, 0] - x1[1, 0]
    assert sep1 - sep0 == pytest.approx(expect, rel=1e-9)
    assert poly6 > 0  # W(d)/W(dq) needs no poly6: the constant cancels


def test_vorticity_magnitude_of_rigid_rotation():
    """omega_i = sum_j (v_j - v_i) x grad_j W_ij (eq. 15) for the rigid
    rotation v = Omega x (x - c) at the centre of a perfect lattice (h = 1/8,
    binary-exact sites).  With r = x_i - x_j, v_j - v_i = -Omega x r and
    grad_j W_ij = spiky (h - q)^2 r/q, so each pair adds
    spiky (h - q)^2 q (Omega - rhat (rhat . Omega)); over a shell of c_k
    sites sum rhat rhat^T = (c_k / 3) I, hence
    omega = (2/3) Omega sum_k c_k spiky (h - q_k)^2 q_k over the shells
    (6, h/2), (12, h/sqrt 2), (8, sqrt(3) h/2)."""
    h = 0.125
    p = params(box=(2.0, 2.0, 2.0), n=(16, 16, 16), physics=pbf.PHYSICS_PAPER)
    pos, _, ci = _lattice_state(h)
    c = pos[ci, :3].astype(np.float64)
    omega = np.array([0.3, -0.2, 1.0])
    vel = np.zeros((pos.shape[0], 4))
    vel[:, :3] = np.cross(omega, pos[:, :3] - c)
    w = orc.vorticity_f64(p, pos.astype(np.float64), vel)[ci]
    spiky = 45.0 / (math.pi * h ** 6)
    mag = sum(cnt * spiky * (h - math.sqrt(f) * h) ** 2 * math.sqrt(f) * h
              for cnt, f in SHELLS)
    expect = (2.0 / 3.0) * mag * omega
    assert np.allclose(w[:3], expect, rtol=1e-12, atol=0.0)
    assert w[3] == pytest.approx(np.linalg.norm(expect), rel=1e-12)


def _three_in_a_row(axis, d1, d2, omegas):
    """Particles A, B, C at B - d1 u, B, B + d2 u along unit vector `axis`
    (d1 + d2 > h: A and C are not neighbours), vorticities `omegas`."""
    u = np.asarray(axis, np.float64)
    u = u / np.linalg.norm(u)
    b = np.array([1.0, 1.0, 1.0])
    pos = np.zeros((3, 4))
    pos[0, :3] = b - d1 * u
    pos[1, :3] = b
    pos[2, :3] = b + d2 * u
    om = np.zeros((3, 4))
    om[:, :3] = omegas
    return pos, om, u


@pytest.mark.parametrize("axis", [(1.0, 0.0, 0.0), (1.0, 2.0, 2.0)])
def test_vorticity_force_points_up_the_gradient(axis):
    """eta = grad |omega| (eq. 16), N = eta/|eta|, dv = dt eps (N x omega_i).
    On a line of three particles with |omega| growing along u, every
    particle's gradient estimate points along +u whatever the kernel's
    magnitudes, so N = u exactly and dv = dt eps (u x omega_i); reversing the
    growth reverses N, and equal magnitudes give no force."""
    h = 0.1
    p = params(box=(2.0, 2.0, 2.0), n=(20, 20, 20), physics=pbf.PHYSICS_PAPER)
    dt = float(np.float32(1.0 / 120.0))
    eps = float(np.float32(1e-4))
    oms = np.array([[0.0, 0.0, 1.0], [0.0, 0.5, 2.0], [0.0, 0.0, 4.0]])
    pos, om, u = _three_in_a_row(axis, 0.4 * h, 0.7 * h, oms)
    dv = orc.vorticity_force_f64(p, pos, om)
    for i in range(3):
        expect = dt * eps * np.cross(u, oms[i])
        assert np.allclose(dv[i, :3], expect, rtol=0, atol=1e-12 * np.abs(expect).max())
    # reversed growth: N = -u
    rev = oms[::-1].copy()
    dv = orc.vorticity_force_f64(p, pos, np.pad(rev, ((0, 0), (0, 1))))
    for i in range(3):
        expect = dt * eps * np.cross(-u, rev[i])
        assert np.allclose(dv[i, :3], expect, rtol=0, atol=1e-12 * np.abs(expect).max())
    # equal magnitudes: eta = 0, no force
    same = np.array([[0.0, 0.0, 2.0], [0.0, 2.0, 0.0], [2.0, 0.0, 0.0]])
    dv = orc.vorticity_force_f64(p, pos, np.pad(same, ((0, 0), (0, 1))))
    assert np.abs(dv[:, :3]).max() < 1e-18


def _one_particle_step(at, sphere, box=(2.0, 2.0, 2.0)):
    """One PAPER step of a single particle at rest without gravity: no
    neighbours, so dp = 0 and the step reduces to the collision handling."""
    p = params(box=box, n=(20, 20, 20), physics=pbf.PHYSICS_PAPER, gravity_y=0.0,
               iterations=1)
    for a in range(4):
        p.sphere[a] = sphere[a]
    pos = np.zeros((1, 4))
    pos[0, :3] = at
    vel = np.zeros((1, 4))
    vel[0, 3] = 1.0
    x, v = orc.run_f64(p, pos, vel, 1)
    return x[0, :3], v[0, :3]


def test_sphere_obstacle_projection():
    """A point within r + h of the obstacle's centre moves to distance r + h
    along its radius; the centre itself moves to +y; a point outside is
    untouched.  The velocity is the displacement over dt."""
    dt = float(np.float32(1.0 / 120.0))
    c = np.array([1.0, 1.0, 1.0])
    r, h = float(np.float32(0.2)), 0.1   # the radius travels as float32
    sph = (1.0, 1.0, 1.0, r)
    inside = np.array([1.1, 1.05, 0.95])
    x, v = _one_particle_step(inside, sph)
    expect = c + (inside - c) * (r + h) / np.linalg.norm(inside - c)
    assert np.allclose(x, expect, rtol=0, atol=1e-15)
    assert np.linalg.norm(x - c) == pytest.approx(r + h, rel=1e-14)
    assert np.allclose(v, (x - inside) / dt, rtol=1e-12, atol=0)
    x, _ = _one_particle_step(c, sph)
    assert np.array_equal(x, c + np.array([0.0, r + h, 0.0]))
    outside = np.array([1.5, 1.0, 1.0])
    x, v = _one_particle_step(outside, sph)
    assert np.array_equal(x, outside) and np.array_equal(v, np.zeros(3))


def test_sphere_projection_runs_before_the_wall_clamp():
    """Obstacle of radius 0.05 centred 0.2 m above the floor (r + h = 0.15):
    a particle at y = 0.12 is projected down to y = 0.05, inside the floor
    band, and the wall clamp then lifts it to min + h = 0.1.  In the other
    order it would end at 0.05 (the clamp leaves y = 0.12 alone)."""
    x, _ = _one_particle_step([1.0, 0.12, 1.0], (1.0, 0.2, 1.0, 0.05))
    assert x[1] == 0.1
    assert x[0] == 1.0 and x[2] == 1.0


def test_grid_density_equals_brute_force():
    """The cell-list FP64 density (used at sizes where the brute-force sum is
    out of reach) against the O(N^2) sum on a settled dam."""
    sc = scenarios.small_dam(dims=(16, 16, 16), box=(2.0, 2.0, 2.0),
                             physics=pbf.PHYSICS_PAPER, seed=4)
    p = sc.params[0]
    pos, vel = orc.state_from_particles(sc.particles[0], np.float64)
    x, _ = orc.run_f64(p, pos, vel, 20)
    a = orc.density_f64(p, x)
    b = orc.density_grid_f64(p, x)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-9)
    assert a.max() > 6000.0
