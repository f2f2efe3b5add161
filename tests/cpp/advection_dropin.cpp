// A caller of OpenCurrent's scalar-advection equation written against the
// ocu headers, running the checks of the reference's advection tests
// (advection3ddoubletest.cpp:90-267) on the GPU library:
//
//  symmetry: a periodic 32^3 box, velocity +-1.5 along one axis (spacing .6
//    along it, 1 across), phi = (c/32)^2 with c the cell index along the
//    flow (mirrored for negative flow); ten advance(1.0) calls of the six
//    runs (+x, +y, +z, -x, -y, -z) must agree cell for cell after each call
//    (axes permuted, mirrored) to 1e-12, and each run must keep sum(phi)
//    to tol * sum (1e-12 upwind, 1e-11 centred and second-order upwind; the
//    sum of |phi| stands in for the sum once values grow);
//  band: a periodic 128^3 box, velocity 1.5 along x, phi = 1 on the open
//    central cube (n/4, 3n/4)^3: a hundred advance(1.0) calls stay finite
//    and exactly zero outside the band j, k in (n/4, 3n/4).
//
// Usage: advection_dropin <scheme: 1 upwind, 2 centred, 4 second-order upwind>
// Prints one line per check and exits 0 when all pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "ocuequation/eqn_scalaradvection3d.h"
#include "ocustorage/grid3d.h"

using namespace ocu;

namespace {

// the cell (i, j, k) of run `dir` (0..2: +x, +y, +z; 3..5: the same axes
// negative) that corresponds to cell (a, b, c) of the +x run
void map_cell(int dir, int n, int a, int b, int c, int &i, int &j, int &k) {
  const int axis = dir % 3;
  const int s = dir < 3 ? a : n - 1 - a;  // coordinate along the flow
  i = a, j = b, k = c;
  if (axis == 0) i = s;
  if (axis == 1) i = b, j = s;  // (i, j) swapped
  if (axis == 2) i = c, k = s;  // (i, k) swapped
}

bool make(int n, int dir, InterpolationType scheme, bool square, Eqn_ScalarAdvection3DD &eqn, double &sum0) {
  Eqn_ScalarAdvection3DParamsD p;
  p.init_grids(n, n, n);
  p.hx = p.hy = p.hz = 1;
  p.advection_scheme = scheme;
  const int axis = dir % 3;
  const double vel = dir < 3 ? 1.5 : -1.5;
  p.u.clear(axis == 0 ? vel : 0);
  p.v.clear(axis == 1 ? vel : 0);
  p.w.clear(axis == 2 ? vel : 0);
  BoundaryCondition per;
  per.type = BC_PERIODIC;
  p.bc = BoundaryConditionSet(per);
  const int q = n / 4;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        double v;
        if (square) {
          v = (i > q && i < 3 * q && j > q && j < 3 * q && k > q && k < 3 * q) ? 1 : 0;
        } else {
          const int c = axis == 0 ? i : axis == 1 ? j : k;
          const float x = (float)(dir < 3 ? c : n - 1 - c) / n;  // in float, as the reference builds it
          v = x * x;
        }
        p.initial_values.at(i, j, k) = v;
      }
  if (!square) {
    if (axis == 0) p.hx = .6f;
    if (axis == 1) p.hy = .6f;
    if (axis == 2) p.hz = .6f;
  }
  if (!p.initial_values.reduce_sum(sum0)) return false;
  return eqn.set_parameters(p);
}

int symmetry(InterpolationType scheme, double tol) {
  const int n = 32;
  std::unique_ptr<Eqn_ScalarAdvection3DD> eq[6];
  std::unique_ptr<Grid3DHostD> host[6];
  double sum0[6];
  for (int d = 0; d < 6; d++) {
    eq[d].reset(new Eqn_ScalarAdvection3DD());
    host[d].reset(new Grid3DHostD());
    if (!make(n, d, scheme, false, *eq[d], sum0[d]) || !host[d]->init_congruent(eq[d]->phi)) {
      printf("symmetry scheme=%d: set-up failed\n", (int)scheme);
      return 1;
    }
  }
  double worst = 0;
  for (int t = 0; t < 10; t++) {
    for (int d = 0; d < 6; d++)
      if (!eq[d]->advance(1.0) || !host[d]->copy_all_data(eq[d]->phi)) {
        printf("symmetry scheme=%d: advance failed\n", (int)scheme);
        return 1;
      }
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int c = 0; c < n; c++) {
          const double ref = host[0]->at(a, b, c);
          for (int d = 1; d < 6; d++) {
            int i, j, k;
            map_cell(d, n, a, b, c, i, j, k);
            worst = fmax(worst, fabs(host[d]->at(i, j, k) - ref));
          }
        }
  }
  // the sum's change relative to the larger of the initial sum and the
  // final sum of |phi| (second-order upwind under forward Euler at the
  // equation's CFL-1 step amplifies long waves, so its values grow; the
  // flux form still conserves to rounding of those values)
  double cons = 0;
  for (int d = 0; d < 6; d++) {
    double s = 0, sa = 0;
    eq[d]->phi.reduce_sum(s);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) sa += fabs(host[d]->at(i, j, k));
    cons = fmax(cons, fabs(s - sum0[d]) / fmax(fabs(sum0[d]), sa));
  }
  const bool ok = worst <= 1e-12 && cons <= tol;
  printf("symmetry scheme=%d max_diff=%.3e rel_sum_change=%.3e ok=%d\n", (int)scheme, worst, cons, (int)ok);
  return ok ? 0 : 1;
}

int band(InterpolationType scheme) {
  const int n = 128, q = n / 4;
  Eqn_ScalarAdvection3DD eqn;
  Grid3DHostD h;
  double sum0 = 0;
  if (!make(n, 0, scheme, true, eqn, sum0) || !h.init_congruent(eqn.phi)) {
    printf("band scheme=%d: set-up failed\n", (int)scheme);
    return 1;
  }
  long bad = 0, nonfinite = 0;
  for (int t = 0; t < 100; t++) {
    if (!eqn.advance(1.0) || !h.copy_all_data(eqn.phi)) {
      printf("band scheme=%d: advance failed\n", (int)scheme);
      return 1;
    }
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
          const double v = h.at(i, j, k);
          if (!std::isfinite(v)) nonfinite++;
          if (!(j > q && j < 3 * q && k > q && k < 3 * q) && v != 0) bad++;
        }
  }
  const bool ok = bad == 0 && nonfinite == 0;
  printf("band scheme=%d nonzero_outside=%ld nonfinite=%ld ok=%d\n", (int)scheme, bad, nonfinite, (int)ok);
  return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
  const int s = argc > 1 ? atoi(argv[1]) : 1;
  const InterpolationType scheme = (InterpolationType)s;
  const double tol = scheme == IT_FIRST_ORDER_UPWIND ? 1e-12 : 1e-11;
  int rc = symmetry(scheme, tol);
  // the reference runs the band check with upwind only; centred forward
  // Euler is unconditionally unstable and is not run for 100 steps
  if (scheme != IT_SECOND_ORDER_CENTERED) rc |= band(scheme);
  return rc;
}
