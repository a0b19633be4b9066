// pbf_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's per-step PBF path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// It is never linked into, loaded by or called from the product library.
//
// Two instantiations of the same step:
//   * Real = float, physics REF: follows the active OpenCL kernels operation
//     by operation (strict IEEE, no contraction: build with -ffp-contract=off
//     and no FMA target), with the kernel constants rounded to 6 significant
//     digits as src/main.cpp:149-168 prints them, linked cell lists built by
//     serial atomic_xchg insertion (update_cells.cl:18, so each list runs in
//     descending particle id) and the x-outer / z-inner stencil order of
//     compute_scaling.cl:33-88.
//   * Real = double: the FP64 oracle of both physics modes (REF and PAPER)
//     that gives per-result density error and position RMS.
// REF passes walk the cell stencil of the current positions in every pass,
// as the reference kernels do; PAPER passes sum over neighbour lists found
// once per step at the predicted positions (Macklin & Mueller, Algorithm 1).
//
// Builtins the reference takes from its OpenCL implementation (sqrt,
// length, division under -cl-fast-relaxed-math) are fixed here to IEEE
// sqrt / division and length(v) = sqrt(x*x + y*y + z*z).  No reference test
// or fixture pins what an OpenCL implementation returns for them: parity
// unpinned at that boundary.
//
// Pins (tests/test_oracle.py): lattice closed forms for rho and lambda in
// both modes, the two-particle scene of assets/scenarios/two.in worked by
// hand, brute-force O(N^2) neighbour counts, antisymmetry of the pressure
// term (sum of dp = 0 away from walls), stable-sort permutation, clamp
// cases, FP32-vs-FP64 agreement.  PAPER-only terms, which the reference does
// not have: s_corr by a two-particle scalar recurrence (n = 2, 4; k = 1e-3,
// 0.1), the vorticity magnitude of a rigid rotation on the lattice shells,
// the confinement force on a line of three particles with a known |omega|
// gradient (orc_vorticity_force_f64), and the sphere projection (inside, at
// the centre, outside, before the wall clamp).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../include/pbf.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Kernel constants.  src/main.cpp:149-168 streams each value with
// std::showpoint and the default precision 6 (printf "%#g") into -D flags,
// so the kernels see the float value of that 6-significant-digit text.
// h is computed in float from the box (main.cpp:164), h^2 / POLY6 /
// GRAD_SPIKY in double from that float h (main.cpp:166-168).
// ---------------------------------------------------------------------------
float fmt6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%#g", v);
  return std::strtof(buf, nullptr);
}

template <typename R>
struct Consts {
  R min[3], max[3];
  R ncell[3];     // NUMBER_OF_CELLS_* (floats in the reference)
  R cell_len[3];  // CELL_LENGTH_*
  R dt, rho0, h, h2, poly6, spiky;
  R eps, c_xsph, g_y;
  R scorr_k, scorr_w_dq;  // k and W(dq)
  int scorr_n;
  R vort_eps;
  R sphere[4];
  int nc[3];
  long C;
  int iterations;
  int physics;
};

template <typename R>
Consts<R> make_consts(const pbf_params &p) {
  Consts<R> k{};
  const float bmin[3] = {p.x_min, p.y_min, p.z_min};
  const float bmax[3] = {p.x_max, p.y_max, p.z_max};
  const float n[3] = {p.x_n, p.y_n, p.z_n};
  const bool ref = p.physics == PBF_PHYSICS_REF;
  const float hf = (p.x_max - p.x_min) / p.x_n;  // main.cpp:164 (float)
  for (int a = 0; a < 3; ++a) {
    const float cl = (bmax[a] - bmin[a]) / n[a];  // main.cpp:159-161
    if (ref) {
      k.min[a] = (R)fmt6(bmin[a]);
      k.max[a] = (R)fmt6(bmax[a]);
      k.ncell[a] = (R)fmt6(n[a]);
      k.cell_len[a] = (R)fmt6(cl);
    } else {
      k.min[a] = (R)bmin[a];
      k.max[a] = (R)bmax[a];
      k.ncell[a] = (R)n[a];
      k.cell_len[a] = (R)((double)bmax[a] - (double)bmin[a]) / (R)n[a];
    }
    k.nc[a] = (int)n[a];
  }
  k.C = (long)k.nc[0] * k.nc[1] * k.nc[2];
  const double pi = 3.14159265358979323846;
  if (ref) {
    k.dt = (R)fmt6(p.time_step_length);                           // :162
    k.rho0 = (R)fmt6(p.rest_density);                             // :163
    k.h = (R)fmt6(hf);                                            // :165
    k.h2 = (R)fmt6(std::pow((double)hf, 2));                      // :166
    k.poly6 = (R)fmt6(315.0 / (64.0 * pi * std::pow((double)hf, 9)));  // :167
    k.spiky = (R)fmt6(45.0 / (pi * std::pow((double)hf, 6)));     // :168
  } else {
    const double h = ((double)p.x_max - (double)p.x_min) / (double)p.x_n;
    k.dt = (R)p.time_step_length;
    k.rho0 = (R)p.rest_density;
    k.h = (R)h;
    k.h2 = (R)(h * h);
    k.poly6 = (R)(315.0 / (64.0 * pi * std::pow(h, 9)));
    k.spiky = (R)(45.0 / (pi * std::pow(h, 6)));
  }
  k.eps = (R)p.relaxation;
  k.c_xsph = (R)p.xsph_c;
  k.g_y = (R)p.gravity_y;
  k.scorr_k = (R)p.scorr_k;
  k.scorr_n = (int)p.scorr_n;
  {
    const double h = (double)k.h;
    const double dq = (double)p.scorr_dq * h;
    const double t = h * h - dq * dq;
    k.scorr_w_dq = (R)((double)k.poly6 * t * t * t);
  }
  k.vort_eps = (R)p.vorticity_eps;
  for (int a = 0; a < 4; ++a) k.sphere[a] = (R)p.sphere[a];
  k.iterations = p.iterations;
  k.physics = p.physics;
  return k;
}

// Truncating float -> int conversion of the kernels' "(int)(...)", made
// total: NaN maps to 0 and the value is bounded before conversion.
template <typename R>
int trunc_int(R v) {
  if (!(v == v)) return 0;
  if (v > (R)1e9) v = (R)1e9;
  if (v < (R)-1e9) v = (R)-1e9;
  return (int)v;
}

// Cell coordinate of one axis: (int)((x - SYSTEM_MIN) / CELL_LENGTH)
// (update_cells.cl:9-15, compute_scaling.cl:21-26).
template <typename R>
int cell_coord(const Consts<R> &k, R x, int a) {
  return trunc_int<R>((x - k.min[a]) / k.cell_len[a]);
}

// Cell id used for the list/sort (update_cells.cl:9-15).  The reference
// indexes out of bounds when a particle leaves the grid; here each
// coordinate is clamped into the grid.
template <typename R>
long cell_id(const Consts<R> &k, const R *x4) {
  long id = 0, stride = 1;
  for (int a = 0; a < 3; ++a) {
    int c = cell_coord(k, x4[a], a);
    c = std::min(std::max(c, 0), k.nc[a] - 1);
    id += (long)c * stride;
    stride *= k.nc[a];
  }
  return id;
}

template <typename R>
struct State {
  int n = 0;
  Consts<R> k;
  std::vector<R> pos, vel, pred, delta, dvel, omega;  // 4 per particle
  std::vector<R> scaling;
  std::vector<int> heads, next;
  std::vector<std::vector<int>> nbrs;  // PAPER: N_i of the step
  R wave = 0;
};

// predictPositions (predict_positions.cl:5-9).
template <typename R>
void predict(State<R> &s) {
  const auto &k = s.k;
  for (int i = 0; i < s.n; ++i) {
    R *v = &s.vel[4 * i];
    const R *x = &s.pos[4 * i];
    R *p = &s.pred[4 * i];
    const R gx = (R)0, gy = k.g_y, gz = (R)0;
    v[0] = v[0] + k.dt * gx;
    v[1] = v[1] + k.dt * gy;
    v[2] = v[2] + k.dt * gz;
    p[0] = x[0] + k.dt * v[0];
    p[1] = x[1] + k.dt * v[1];
    p[2] = x[2] + k.dt * v[2];
  }
}

// initCellsOld + updateCells (init_cells_old.cl:6-12, update_cells.cl:9-18,
// driver Simulation.cpp:330-349), executed serially in work-item order so
// that next[i] = atomic_xchg(&heads[cell], i) yields descending-id lists.
template <typename R>
void build_lists(State<R> &s) {
  std::fill(s.heads.begin(), s.heads.end(), -1);
  std::fill(s.next.begin(), s.next.end(), -1);
  for (int i = 0; i < s.n; ++i) {
    const long c = cell_id(s.k, &s.pred[4 * i]);
    s.next[i] = s.heads[c];
    s.heads[c] = i;
  }
}

// Visit the 27-cell stencil of particle i in the reference order
// (x outer, z inner; compute_scaling.cl:33-53) and every list entry
// (compute_scaling.cl:56-87): f(j) for each j != i in the lists.
template <typename R, typename F>
void for_candidates(const State<R> &s, int i, F &&f) {
  const auto &k = s.k;
  const R *xi = &s.pred[4 * i];
  int cc[3];
  for (int a = 0; a < 3; ++a) cc[a] = cell_coord(k, xi[a], a);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int nb[3] = {cc[0] + dx, cc[1] + dy, cc[2] + dz};
        bool skip = false;
        for (int a = 0; a < 3; ++a)
          if (nb[a] < 0 || (R)nb[a] >= k.ncell[a]) skip = true;
        if (skip) continue;
        const long ci = nb[0] + (long)nb[1] * k.nc[0] +
                        (long)nb[2] * k.nc[0] * k.nc[1];
        for (int j = s.heads[ci]; j != -1; j = s.next[j])
          if (j != i) f(j);
      }
}

// PAPER neighbour lists, Algorithm 1 of Macklin & Mueller (lines 5-7):
// N_i = {j != i : 0 < |x*_i - x*_j|^2 < h^2} at the predicted positions,
// fixed for the step.
template <typename R>
void build_neighbours(State<R> &s) {
  const auto &k = s.k;
  s.nbrs.resize(s.n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    std::vector<int> &l = s.nbrs[i];
    l.clear();
    const R *xi = &s.pred[4 * i];
    for_candidates(s, i, [&](int j) {
      const R *xj = &s.pred[4 * j];
      const R d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      const R r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 > (R)0 && r2 < k.h2) l.push_back(j);
    });
  }
}

// The pairs a pass sums over: REF walks the 27-cell stencil of the current
// position (compute_scaling.cl:21-88); PAPER the step's N_i.  Each pass
// tests 0 < r^2 < h^2 at its own positions.
template <typename R, typename F>
void for_pairs(const State<R> &s, int i, F &&f) {
  if (s.k.physics == PBF_PHYSICS_PAPER) {
    for (int j : s.nbrs[i]) f(j);
  } else {
    for_candidates(s, i, f);
  }
}

// r = x_i - x_j and r^2 = r.x*r.x + r.y*r.y + r.z*r.z, left to right
// (compute_scaling.cl:58-59).
template <typename R>
R dist2(const R *xi, const R *xj, R *r) {
  r[0] = xi[0] - xj[0];
  r[1] = xi[1] - xj[1];
  r[2] = xi[2] - xj[2];
  return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

template <typename R>
R len3(const R *v) {
  return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// computeScaling (compute_scaling.cl:12-140): density (self excluded) into
// pred.w and lambda.  REF: denominator (sum|gradW| + |sum gradW|)^2/rho0^2
// + eps (:80, :131, :139-140).  PAPER: (|sum gradW|^2 + sum |gradW|^2)
// / rho0^2 + eps (Macklin & Mueller eqs. 8, 9, 11).
template <typename R>
void compute_scaling(State<R> &s) {
  const auto &k = s.k;
  const bool ref = k.physics == PBF_PHYSICS_REF;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    const R *xi = &s.pred[4 * i];
    R density = 0, gsum = 0, gsum2 = 0;
    R gi[3] = {0, 0, 0};
    for_pairs(s, i, [&](int j) {
      R r[3];
      const R r2 = dist2(xi, &s.pred[4 * j], r);
      if (r2 > (R)0 && r2 < k.h2) {
        const R rl = std::sqrt(r2);
        R g[3];
        for (int a = 0; a < 3; ++a)
          g[a] = r[a] / rl * k.spiky * (k.h - rl) * (k.h - rl);
        const R w = k.poly6 * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
        density += w;
        if (ref) {
          gsum += len3(g);
        } else {
          gsum2 += g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        }
        gi[0] += g[0];
        gi[1] += g[1];
        gi[2] += g[2];
      }
    });
    R denom;
    if (ref) {
      gsum += len3(gi);
      denom = gsum * gsum / (k.rho0 * k.rho0) + k.eps;
    } else {
      denom = (gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2] + gsum2) /
                  (k.rho0 * k.rho0) +
              k.eps;
    }
    s.pred[4 * i + 3] = density;
    const R constraint = (density / k.rho0) - (R)1;
    s.scaling[i] = (R)-1 * constraint / denom;
  }
}

// Wall clamp of compute_delta.cl:128-150 (x-min wall offset by the wave
// generator), after the PAPER sphere obstacle projection.
template <typename R>
void collide(const Consts<R> &k, R wave, R *f) {
  if (k.physics == PBF_PHYSICS_PAPER && k.sphere[3] > (R)0) {
    R d[3] = {f[0] - k.sphere[0], f[1] - k.sphere[1], f[2] - k.sphere[2]};
    const R rr = k.sphere[3] + k.h;
    const R d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (d2 < rr * rr) {
      if (d2 > (R)0) {
        const R sc = rr / std::sqrt(d2);
        for (int a = 0; a < 3; ++a) f[a] = k.sphere[a] + d[a] * sc;
      } else {
        f[1] = k.sphere[1] + rr;
      }
    }
  }
  if ((f[0] - k.h) < (k.min[0] + wave)) {
    f[0] = k.min[0] + wave + k.h;
  } else if ((f[0] + k.h) > k.max[0]) {
    f[0] = k.max[0] - k.h;
  }
  for (int a = 1; a < 3; ++a) {
    if ((f[a] - k.h) < k.min[a]) {
      f[a] = k.min[a] + k.h;
    } else if ((f[a] + k.h) > k.max[a]) {
      f[a] = k.max[a] - k.h;
    }
  }
}

// computeDelta (compute_delta.cl:13-152) then updatePredicted
// (update_predicted.cl:4-7).  PAPER adds s_corr = -k (W/W(dq))^n inside the
// sum (eqs. 13, 14).
template <typename R>
void compute_delta_and_update(State<R> &s) {
  const auto &k = s.k;
  const bool ref = k.physics == PBF_PHYSICS_REF;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    const R *xi = &s.pred[4 * i];
    R sum[3] = {0, 0, 0};
    const R li = s.scaling[i];
    for_pairs(s, i, [&](int j) {
      R r[3];
      const R r2 = dist2(xi, &s.pred[4 * j], r);
      if (r2 > (R)0 && r2 < k.h2) {
        const R rl = std::sqrt(r2);
        R g[3];
        for (int a = 0; a < 3; ++a)
          g[a] = (R)-1 * r[a] / rl * k.spiky * (k.h - rl) * (k.h - rl);
        R coef = li + s.scaling[j];
        if (!ref) {
          const R w = k.poly6 * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
          const R q = w / k.scorr_w_dq;
          R pw = 1;
          for (int e = 0; e < k.scorr_n; ++e) pw *= q;
          coef = coef + (-k.scorr_k * pw);
        }
        for (int a = 0; a < 3; ++a) sum[a] += coef * g[a];
      }
    });
    R fut[3];
    for (int a = 0; a < 3; ++a) fut[a] = xi[a] + sum[a] / k.rho0;
    collide(k, s.wave, fut);
    for (int a = 0; a < 3; ++a) s.delta[4 * i + a] = fut[a] - xi[a];
  }
  for (int i = 0; i < s.n; ++i)
    for (int a = 0; a < 3; ++a)
      s.pred[4 * i + a] = s.pred[4 * i + a] + s.delta[4 * i + a];
}

// updateVelocities (update_velocities.cl:5-8).
template <typename R>
void update_velocities(State<R> &s) {
  for (int i = 0; i < s.n; ++i)
    for (int a = 0; a < 3; ++a)
      s.vel[4 * i + a] = (s.pred[4 * i + a] - s.pos[4 * i + a]) / s.k.dt;
}

// applyVorticityAndViscosity (apply_vorticity_and_viscosity.cl:12-104):
// XSPH dv = c sum (1/rho_j)(v_j - v_i) W_ij with the density of the last
// lambda pass.  PAPER also accumulates the vorticity
// omega_i = sum (v_j - v_i) x grad_j W_ij (eq. 15).
template <typename R>
void xsph_vorticity(State<R> &s) {
  const auto &k = s.k;
  const bool ref = k.physics == PBF_PHYSICS_REF;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    const R *xi = &s.pred[4 * i];
    const R *vi = &s.vel[4 * i];
    R vs[3] = {0, 0, 0};
    R om[3] = {0, 0, 0};
    for_pairs(s, i, [&](int j) {
      R r[3];
      const R *xj = &s.pred[4 * j];
      const R r2 = dist2(xi, xj, r);
      if (r2 > (R)0 && r2 < k.h2) {
        const R *vj = &s.vel[4 * j];
        const R v[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
        const R w = k.poly6 * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
        // PAPER skips a neighbour whose last density is 0 (it entered the
        // support after the last lambda pass); REF divides as the kernel does
        const R inv = (ref || xj[3] > (R)0) ? (R)1 / xj[3] : (R)0;
        for (int a = 0; a < 3; ++a) vs[a] += inv * v[a] * w;
        if (!ref) {
          const R rl = std::sqrt(r2);
          const R gs = k.spiky * (k.h - rl) * (k.h - rl) / rl;
          const R g[3] = {r[0] * gs, r[1] * gs, r[2] * gs};  // grad_j W_ij
          om[0] += v[1] * g[2] - v[2] * g[1];
          om[1] += v[2] * g[0] - v[0] * g[2];
          om[2] += v[0] * g[1] - v[1] * g[0];
        }
      }
    });
    for (int a = 0; a < 3; ++a) s.dvel[4 * i + a] = k.c_xsph * vs[a];
    for (int a = 0; a < 3; ++a) s.omega[4 * i + a] = om[a];
    s.omega[4 * i + 3] = len3(om);
  }
}

// PAPER vorticity force (eq. 16): eta = sum (|w_j| - |w_i|) grad_i W_ij,
// N = eta/|eta|, dv += dt * eps (N x w_i).
template <typename R>
void vorticity_force(State<R> &s) {
  const auto &k = s.k;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    const R *xi = &s.pred[4 * i];
    const R wi = s.omega[4 * i + 3];
    R eta[3] = {0, 0, 0};
    for_pairs(s, i, [&](int j) {
      R r[3];
      const R r2 = dist2(xi, &s.pred[4 * j], r);
      if (r2 > (R)0 && r2 < k.h2) {
        const R rl = std::sqrt(r2);
        const R gs = -k.spiky * (k.h - rl) * (k.h - rl) / rl;  // grad_i W
        const R dw = s.omega[4 * j + 3] - wi;
        for (int a = 0; a < 3; ++a) eta[a] += dw * r[a] * gs;
      }
    });
    const R e2 = eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2];
    if (e2 > (R)0) {
      const R inv = (R)1 / std::sqrt(e2);
      const R N[3] = {eta[0] * inv, eta[1] * inv, eta[2] * inv};
      const R *w = &s.omega[4 * i];
      const R f[3] = {N[1] * w[2] - N[2] * w[1], N[2] * w[0] - N[0] * w[2],
                      N[0] * w[1] - N[1] * w[0]};
      for (int a = 0; a < 3; ++a)
        s.dvel[4 * i + a] += k.dt * k.vort_eps * f[a];
    }
  }
}

// updatePositions (update_positions.cl:6-12).
template <typename R>
void update_positions(State<R> &s) {
  for (int i = 0; i < s.n; ++i) {
    for (int a = 0; a < 3; ++a) s.pos[4 * i + a] = s.pred[4 * i + a];
    for (int a = 0; a < 3; ++a)
      s.vel[4 * i + a] = s.vel[4 * i + a] + s.dvel[4 * i + a];
    s.pos[4 * i + 3] = len3(&s.vel[4 * i]);
  }
}

// Simulation::step (Simulation.cpp:455-576): predict, cell lists,
// iterations x (lambda, delta, update), velocities, XSPH (+ vorticity),
// positions.
template <typename R>
void step(State<R> &s) {
  predict(s);
  build_lists(s);
  if (s.k.physics == PBF_PHYSICS_PAPER) build_neighbours(s);
  for (int it = 0; it < s.k.iterations; ++it) {
    compute_scaling(s);
    compute_delta_and_update(s);
  }
  update_velocities(s);
  xsph_vorticity(s);
  if (s.k.physics == PBF_PHYSICS_PAPER && s.k.vort_eps != (R)0)
    vorticity_force(s);
  update_positions(s);
}

template <typename R>
void init_state(State<R> &s, const pbf_params &p, int n) {
  s.n = n;
  s.k = make_consts<R>(p);
  s.pos.assign(4 * (size_t)n, 0);
  s.vel.assign(4 * (size_t)n, 0);
  s.pred.assign(4 * (size_t)n, 0);
  s.delta.assign(4 * (size_t)n, 0);
  s.dvel.assign(4 * (size_t)n, 0);
  s.omega.assign(4 * (size_t)n, 0);
  s.scaling.assign(n, 0);
  s.heads.assign(s.k.C, -1);
  s.next.assign(n, -1);
}

bool params_ok(const pbf_params *p, int n) {
  if (!p || n < 0) return false;
  if (!(p->x_n >= 1 && p->y_n >= 1 && p->z_n >= 1)) return false;
  if (p->iterations < 0 || p->scorr_n < 0 || p->scorr_n > 16) return false;
  return true;
}

template <typename R, typename T>
int run(const pbf_params *p, int n, T *pos4, T *vel4, int nsteps, float wave) {
  if (!params_ok(p, n) || nsteps < 0) return 1;
  State<R> s;
  init_state(s, *p, n);
  s.wave = (R)wave;
  for (size_t i = 0; i < 4 * (size_t)n; ++i) {
    s.pos[i] = (R)pos4[i];
    s.vel[i] = (R)vel4[i];
  }
  for (int t = 0; t < nsteps; ++t) step(s);
  for (size_t i = 0; i < 4 * (size_t)n; ++i) {
    pos4[i] = (T)s.pos[i];
    vel4[i] = (T)s.vel[i];
  }
  return 0;
}

}  // namespace

extern "C" {

int orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

// Number of cells of the grid the oracle bins on; cell_start of
// orc_probe_f32 has this many entries + 1.
long orc_num_cells(const pbf_params *p) {
  if (!params_ok(p, 0)) return -1;
  return make_consts<double>(*p).C;
}

// Constants the kernels see: [h, h2, poly6, spiky, dt, rho0, cell_len x3,
// min x3, max x3, ncell x3, W(dq)] (main.cpp:149-168 rounding in REF).
int orc_constants(const pbf_params *p, float *out) {
  if (!params_ok(p, 0) || !out) return 1;
  const Consts<float> k = make_consts<float>(*p);
  const float v[] = {k.h,           k.h2,          k.poly6,      k.spiky,
                     k.dt,          k.rho0,        k.cell_len[0], k.cell_len[1],
                     k.cell_len[2], k.min[0],      k.min[1],     k.min[2],
                     k.max[0],      k.max[1],      k.max[2],     k.ncell[0],
                     k.ncell[1],    k.ncell[2],    k.scorr_w_dq};
  std::memcpy(out, v, sizeof v);
  return 0;
}

// nsteps of Simulation::step in FP32, state in place (float4 records).
int orc_run_f32(const pbf_params *p, int n, float *pos4, float *vel4,
                int nsteps, float wave) {
  return run<float>(p, n, pos4, vel4, nsteps, wave);
}

// The same in FP64 (double4 records in and out).
int orc_run_f64(const pbf_params *p, int n, double *pos4, double *vel4,
                int nsteps, float wave) {
  return run<double>(p, n, pos4, vel4, nsteps, wave);
}

// Neighbour-search stage on the current state without advancing it (FP32):
// predict, cell ids (update_cells.cl:9-15), the permutation of a stable
// sort of particle ids by cell, the cell start table, stencil neighbour
// counts and rho / lambda of the first constraint pass.  Any output may be
// NULL.
int orc_probe_f32(const pbf_params *p, int n, const float *pos4,
                  const float *vel4, int *cell_of, int *sorted_ids,
                  int *cell_start, int *nbr, float *rho, float *lambda,
                  float *pred4) {
  if (!params_ok(p, n)) return 1;
  State<float> s;
  init_state(s, *p, n);
  std::memcpy(s.pos.data(), pos4, sizeof(float) * 4 * n);
  std::memcpy(s.vel.data(), vel4, sizeof(float) * 4 * n);
  predict(s);
  build_lists(s);
  if (s.k.physics == PBF_PHYSICS_PAPER) build_neighbours(s);
  std::vector<long> cell(n);
  for (int i = 0; i < n; ++i) cell[i] = cell_id(s.k, &s.pred[4 * i]);
  if (cell_of)
    for (int i = 0; i < n; ++i) cell_of[i] = (int)cell[i];
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int a, int b) { return cell[a] < cell[b]; });
  if (sorted_ids) std::copy(perm.begin(), perm.end(), sorted_ids);
  if (cell_start) {
    std::vector<long> cnt(s.k.C + 1, 0);
    for (int i = 0; i < n; ++i) cnt[cell[i]]++;
    long acc = 0;
    for (long c = 0; c <= s.k.C; ++c) {
      cell_start[c] = (int)acc;
      acc += cnt[c];
    }
  }
  if (nbr) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      int cnt = 0;
      for_candidates(s, i, [&](int j) {
        float r[3];
        const float r2 = dist2(&s.pred[4 * i], &s.pred[4 * j], r);
        if (r2 > 0.0f && r2 < s.k.h2) ++cnt;
      });
      nbr[i] = cnt;
    }
  }
  if (pred4) std::memcpy(pred4, s.pred.data(), sizeof(float) * 4 * n);
  if (rho || lambda) {
    compute_scaling(s);
    for (int i = 0; i < n; ++i) {
      if (rho) rho[i] = s.pred[4 * i + 3];
      if (lambda) lambda[i] = s.scaling[i];
    }
  }
  return 0;
}

// Brute-force O(N^2) neighbour counts over all pairs with the reference
// predicate i != j && 0 < r^2 < h2, r^2 summed left to right in FP32.
int orc_brute_neighbors_f32(int n, const float *pred4, float h2, int *counts) {
  if (n < 0 || !counts) return 1;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    int c = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      float r[3];
      const float r2 = dist2(&pred4[4 * i], &pred4[4 * j], r);
      if (r2 > 0.0f && r2 < h2) ++c;
    }
    counts[i] = c;
  }
  return 0;
}

// FP64 density rho_i = sum_{j != i, 0 < r^2 < h^2} W_ij at the given
// positions (brute force; small n).
int orc_density_f64(const pbf_params *p, int n, const double *pos4,
                    double *rho) {
  if (!params_ok(p, n) || !rho) return 1;
  const Consts<double> k = make_consts<double>(*p);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    double acc = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double r[3];
      const double r2 = dist2(&pos4[4 * i], &pos4[4 * j], r);
      if (r2 > 0 && r2 < k.h2) {
        const double t = k.h2 - r2;
        acc += k.poly6 * t * t * t;
      }
    }
    rho[i] = acc;
  }
  return 0;
}

// The same density through the cell lists of the parameters' grid (any n):
// the 27-cell stencil holds every j within h when the cells are at least h
// long (REF grids: cell length = h; PAPER needs y, z cells >= h and x cells
// = h).
int orc_density_grid_f64(const pbf_params *p, int n, const double *pos4,
                         double *rho) {
  if (!params_ok(p, n) || !rho) return 1;
  State<double> s;
  init_state(s, *p, n);
  std::memcpy(s.pred.data(), pos4, sizeof(double) * 4 * (size_t)n);
  build_lists(s);
  const Consts<double> &k = s.k;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double acc = 0;
    for_candidates(s, i, [&](int j) {
      double r[3];
      const double r2 = dist2(&s.pred[4 * (size_t)i], &s.pred[4 * (size_t)j], r);
      if (r2 > 0 && r2 < k.h2) {
        const double t = k.h2 - r2;
        acc += k.poly6 * t * t * t;
      }
    });
    rho[i] = acc;
  }
  return 0;
}

// PAPER vorticity omega_i = sum (v_j - v_i) x grad_j W_ij (eq. 15) at the
// given positions and velocities (FP64; lists built from the positions).
int orc_vorticity_f64(const pbf_params *p, int n, const double *pos4,
                      const double *vel4, double *omega4) {
  if (!params_ok(p, n) || !omega4) return 1;
  pbf_params q = *p;
  q.physics = PBF_PHYSICS_PAPER;
  State<double> s;
  init_state(s, q, n);
  for (size_t i = 0; i < 4 * (size_t)n; ++i) {
    s.pos[i] = pos4[i];
    s.pred[i] = pos4[i];
    s.vel[i] = vel4[i];
  }
  for (int i = 0; i < n; ++i) s.pred[4 * i + 3] = 1.0;  // rho_j for XSPH
  build_lists(s);
  build_neighbours(s);
  xsph_vorticity(s);
  std::memcpy(omega4, s.omega.data(), sizeof(double) * 4 * n);
  return 0;
}

// PAPER vorticity confinement (eq. 16) alone: the velocity change
// dt eps (N x omega_i) of vorticity_force for given positions and
// vorticities omega4 (x, y, z; .w is recomputed as |omega|), FP64, lists
// built from the positions.
int orc_vorticity_force_f64(const pbf_params *p, int n, const double *pos4,
                            const double *omega4, double *dv4) {
  if (!params_ok(p, n) || !dv4) return 1;
  pbf_params q = *p;
  q.physics = PBF_PHYSICS_PAPER;
  State<double> s;
  init_state(s, q, n);
  for (size_t i = 0; i < 4 * (size_t)n; ++i) {
    s.pos[i] = pos4[i];
    s.pred[i] = pos4[i];
    s.omega[i] = omega4[i];
  }
  for (int i = 0; i < n; ++i) s.omega[4 * i + 3] = len3(&s.omega[4 * i]);
  build_lists(s);
  build_neighbours(s);
  vorticity_force(s);
  std::memcpy(dv4, s.dvel.data(), sizeof(double) * 4 * n);
  return 0;
}

}  // extern "C"
