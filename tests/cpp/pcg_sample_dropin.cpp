// A caller written against OpenCurrent's C++ API: the PCG pressure solver,
// the two-layer ghost fill and point sampling, through the reference's
// classes and free functions.  Built against libocu.so + libocu_hip.so it
// runs on the GPU.
//
// 1. Sol_PCGPressure3DDeviceD with the multigrid preconditioner on an n^3
//    all-Neumann problem with coefficient 1: the operator is -Laplacian, so
//    a rhs lam * mode (lam the positive eigenvalue) gives back the mode.
// 2. apply_3d_boundary_conditions_level2_nocorners on a linear field with
//    matching Neumann gradients, then sample_points_3d across the whole box
//    in x (near the x faces the lookup reads the face ghosts) and inside the
//    cell-centre hull in y and z (the fill leaves edges): the linear field
//    comes back.
// Prints one line per part ("... ok=1") and exits 0 when both pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "ocuequation/sol_pcgpressure3d.h"
#include "ocustorage/grid1d.h"
#include "ocustorage/grid3d.h"
#include "ocustorage/grid3dboundary.h"
#include "ocustorage/grid3dsample.h"

using namespace ocu;

static int pcg_part(int n) {
  const double h = 1.0 / n, pi = 3.14159265358979323846;
  const double lam = (4 / (h * h)) * (pow(sin(pi / (2.0 * n)), 2) + pow(sin(2 * pi / (2.0 * n)), 2));
  Grid3DHostD hb, hp;
  if (!hb.init(n, n, n, 1, 1, 1) || !hp.init(n, n, n, 1, 1, 1)) return 5;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) hb.at(i, j, k) = lam * cos(pi * (i + .5) / n) * cos(2 * pi * (j + .5) / n);
  Grid3DDeviceD rhs, coeff;
  if (!rhs.init(n, n, n, 1, 1, 1) || !rhs.copy_all_data(hb) || !coeff.init_congruent(rhs)) return 2;
  coeff.clear(1.0);
  Sol_PCGPressure3DDeviceD solver;
  BoundaryCondition neumann;
  neumann.type = BC_NEUMANN;
  solver.bc.xpos = solver.bc.xneg = solver.bc.ypos = solver.bc.yneg = solver.bc.zpos = solver.bc.zneg = neumann;
  solver.preconditioner = PRECOND_MULTIGRID;
  if (!solver.initialize_storage(n, n, n, (float)h, (float)h, (float)h, &rhs, &coeff)) return 3;
  double residual = 0;
  const bool ok = solver.solve(residual, 1e-8, 100);
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
  printf("pcg n=%d ok=%d residual=%.3e iterations=%d max_err=%.3e\n", n, (int)(ok && err < 1e-6), residual,
         solver.iterations(), err);
  return ok && err < 1e-6 ? 0 : 1;
}

static int sample_part() {
  const int nx = 12, ny = 9, nz = 10, m = 500;
  const float hx = .25f, hy = .5f, hz = .125f;
  const double a = .3, bx = .375, by = -.75, bz = 1.125;  // gradients exact in float
  Grid3DHostD hphi;
  if (!hphi.init(nx, ny, nz, 2, 2, 2)) return 5;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) hphi.at(i, j, k) = a + bx * (i + .5) * hx + by * (j + .5) * hy + bz * (k + .5) * hz;
  Grid3DDeviceD phi;
  if (!phi.init(nx, ny, nz, 2, 2, 2) || !phi.copy_all_data(hphi)) return 2;
  BoundaryConditionSet bc;
  bc.xpos.type = bc.xneg.type = bc.ypos.type = bc.yneg.type = bc.zpos.type = bc.zneg.type = BC_NEUMANN;
  bc.xpos.value = bc.xneg.value = (float)bx;
  bc.ypos.value = bc.yneg.value = (float)by;
  bc.zpos.value = bc.zneg.value = (float)bz;
  if (!apply_3d_boundary_conditions_level2_nocorners(phi, bc, hx, hy, hz)) return 3;
  Grid1DHost<float> px, py, pz;
  Grid1DHostD out;
  if (!px.init(m, 0) || !py.init(m, 0) || !pz.init(m, 0) || !out.init(m, 0)) return 5;
  srand(11);
  for (int q = 0; q < m; q++) {
    px.at(q) = (float)(nx * (rand() / (double)RAND_MAX) * hx);
    py.at(q) = (float)((.5 + (ny - 1) * (rand() / (double)RAND_MAX)) * hy);
    pz.at(q) = (float)((.5 + (nz - 1) * (rand() / (double)RAND_MAX)) * hz);
  }
  Grid1DDevice<float> dx, dy, dz;
  Grid1DDeviceD dout;
  if (!dx.init(m, 0) || !dy.init(m, 0) || !dz.init(m, 0) || !dout.init(m, 0)) return 2;
  if (!dx.copy_all_data(px) || !dy.copy_all_data(py) || !dz.copy_all_data(pz)) return 2;
  if (!sample_points_3d(dout, dx, dy, dz, phi, bc, (float)nx, (float)ny, (float)nz, hx, hy, hz, .5f, .5f, .5f))
    return 3;
  if (!out.copy_all_data(dout)) return 4;
  double err = 0;
  for (int q = 0; q < m; q++) {
    const double want = a + bx * px.at(q) + by * py.at(q) + bz * pz.at(q);
    err = fmax(err, fabs(out.at(q) - want));
  }
  printf("sample m=%d ok=%d max_err=%.3e\n", m, (int)(err < 1e-5), err);
  return err < 1e-5 ? 0 : 1;
}

int main(int argc, char **argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 32;
  const int a = pcg_part(n), b = sample_part();
  return a ? a : b;
}
