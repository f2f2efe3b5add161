// A caller written against OpenCurrent's C++ API, unchanged: it includes the
// ocu headers, configures Sol_MultigridPressure3DDeviceD through its public
// fields and calls initialize_storage / solve exactly as the reference's own
// tests do (mgdoubletest.cpp:60-120).  Built against libocu.so +
// libocu_hip.so it runs on the GPU.
//
// Problem: all-Neumann Poisson on n^3 with the RHS a discrete eigenmode
// lambda * cos(pi (i+1/2)/n) cos(2 pi (j+1/2)/n), whose solution is the mode
// itself up to a constant (removed before comparing).  Prints one line and
// exits 0 when the solve converges and the error is below 1e-8.
#include <cmath>
#include <cstdio>

#include "ocuequation/sol_mgpressure3d.h"
#include "ocustorage/grid3d.h"

using namespace ocu;

int main(int argc, char **argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 64;
  const double h = 1.0 / n, pi = 3.14159265358979323846;
  const double lam = -(4 / (h * h)) * (pow(sin(pi / (2.0 * n)), 2) + pow(sin(2 * pi / (2.0 * n)), 2));

  Grid3DHostD hb, hp;
  if (!hb.init(n, n, n, 1, 1, 1) || !hp.init(n, n, n, 1, 1, 1)) {
    printf("mg_dropin: could not allocate pinned host grids (no GPU?)\n");
    return 5;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        hb.at(i, j, k) = lam * cos(pi * (i + .5) / n) * cos(2 * pi * (j + .5) / n);

  Grid3DDeviceD rhs;
  if (!rhs.init(n, n, n, 1, 1, 1) || !rhs.copy_all_data(hb)) return 2;

  Sol_MultigridPressure3DDeviceD solver;
  BoundaryCondition neumann;
  neumann.type = BC_NEUMANN;
  solver.bc.xpos = solver.bc.xneg = solver.bc.ypos = solver.bc.yneg = solver.bc.zpos = solver.bc.zneg = neumann;
  solver.convergence = CONVERGENCE_L2;
  if (!solver.initialize_storage(n, n, n, h, h, h, &rhs)) return 3;

  double residual = 0;
  const bool ok = solver.solve(residual, 1e-9, 15);
  if (!hp.copy_all_data(solver.pressure())) return 4;

  double mean = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) mean += hp.at(i, j, k);
  mean /= (double)n * n * n;
  double err = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        err = fmax(err, fabs(hp.at(i, j, k) - mean - cos(pi * (i + .5) / n) * cos(2 * pi * (j + .5) / n)));
  printf("mg_dropin n=%d ok=%d residual=%.3e vcycles=%d levels=%d max_err=%.3e\n", n, (int)ok, residual,
         solver.num_vcycles, solver.num_levels(), err);
  return ok && err < 1e-8 ? 0 : 1;
}
