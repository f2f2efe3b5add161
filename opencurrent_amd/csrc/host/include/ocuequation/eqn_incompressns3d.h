// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/eqn_incompressns3d.h: Boussinesq incompressible Navier-Stokes
// on a MAC grid (eqn_incompressns3d.cu)
#ifndef __OCU_EQUATION_EQN_INCOMPRESS_NS_3D_H__
#define __OCU_EQUATION_EQN_INCOMPRESS_NS_3D_H__

#include "ocuequation/equation.h"
#include "ocuequation/parameters.h"
#include "ocuequation/sol_laplaciancent3d.h"
#include "ocuequation/sol_passiveadvection3d.h"
#include "ocuequation/sol_project3d.h"
#include "ocuequation/sol_selfadvection3d.h"
#include "ocuutil/direction.h"
#include "ocuutil/timestep.h"

#include <type_traits>

namespace ocu {

template <typename T>
class Eqn_IncompressibleNS3DParams : public Parameters {
 public:
  int nx, ny, nz;
  double hx, hy, hz;
  double max_divergence;
  BoundaryConditionSet temp_bc;
  BoundaryConditionSet flow_bc;
  DirectionType vertical_direction;
  T gravity;
  T bouyancy;
  Grid3DHost<T> init_u, init_v, init_w;
  Grid3DHost<T> init_temp;
  T thermal_diffusion;
  T viscosity;
  InterpolationType advection_scheme;
  TimeStepType time_step;
  double cfl_factor;
  // not in the reference: project with FP32 V-cycles inside FP64 iterative
  // refinement (Sol_ProjectDivergenceMixed3DDeviceD); T = double only
  bool mixed_precision_projection;
  // not in the reference: advance with the fused step kernel
  // (ocu_ns_step_fused) where it applies; false runs the reference's
  // operator-by-operator sequence.  Both give the same values.
  bool fused_step;

  // defaults of eqn_incompressns3d.h:59-74
  Eqn_IncompressibleNS3DParams() {
    advection_scheme = IT_FIRST_ORDER_UPWIND;
    viscosity = 0;
    thermal_diffusion = 0;
    max_divergence = 1e-5;
    nx = ny = nz = 0;
    hx = hy = hz = 0;
    cfl_factor = 1;
    gravity = -9.8;
    bouyancy = 1;
    time_step = TS_FORWARD_EULER;
    vertical_direction = DIR_YPOS;
    mixed_precision_projection = false;
    fused_step = true;
  }

  // eqn_incompressns3d.h:78-86: paddings (0,1,1), (1,0,1), (1,1,0), (1,1,1)
  // give the four grids one layout
  void init_grids(int nx_val, int ny_val, int nz_val, bool pinned = true) {
    nx = nx_val;
    ny = ny_val;
    nz = nz_val;
    init_u.init(nx + 1, ny, nz, 1, 1, 1, pinned, 0, 1, 1);
    init_v.init(nx, ny + 1, nz, 1, 1, 1, pinned, 1, 0, 1);
    init_w.init(nx, ny, nz + 1, 1, 1, 1, pinned, 1, 1, 0);
    init_temp.init(nx, ny, nz, 1, 1, 1, pinned, 1, 1, 1);
  }
};

template <typename T>
class Eqn_IncompressibleNS3DBase : public Equation {
 protected:
  BoundaryConditionSet _thermalbc;
  double _lastdt;
  DirectionType _vertical_direction;
  int _nx, _ny, _nz;
  double _hx, _hy, _hz;
  double _max_divergence;
  double _cfl_factor;
  TimeStepType _time_step;
  T _gravity, _bouyancy;

 public:
  int num_steps;
  // print get_max_stable_timestep() results like the reference (:325)
  bool verbose;

  Eqn_IncompressibleNS3DBase();
  bool set_base_parameters(const Eqn_IncompressibleNS3DParams<T> &params);

  int nx() const { return _nx; }
  int ny() const { return _ny; }
  int nz() const { return _nz; }
  double hx() const { return _hx; }
  double hy() const { return _hy; }
  double hz() const { return _hz; }
};

template <typename T>
class Eqn_IncompressibleNS3D : public Eqn_IncompressibleNS3DBase<T> {
 protected:
  Sol_ProjectDivergence3DDevice<T> _projection_solver;
  Sol_ProjectDivergenceMixed3DDeviceD _mixed_projection_solver;  // used when _mixed_projection
  bool _mixed_projection;
  Sol_SelfAdvection3DDevice<T> _advection_solver;
  Sol_PassiveAdvection3DDevice<T> _thermal_solver;
  Sol_LaplacianCentered3DDevice<T> _thermal_diffusion;
  Sol_LaplacianCentered3DDevice<T> _u_diffusion;
  Sol_LaplacianCentered3DDevice<T> _v_diffusion;
  Sol_LaplacianCentered3DDevice<T> _w_diffusion;

  Grid3DDevice<T> _u, _v, _w;
  Grid3DDevice<T> _deriv_udt, _deriv_vdt, _deriv_wdt;
  Grid3DDevice<T> _temp;
  Grid3DDevice<T> _deriv_tempdt;
  Grid3DDevice<T> _last_deriv_tempdt;
  Grid3DDevice<T> _last_deriv_udt, _last_deriv_vdt, _last_deriv_wdt;
  // fused step: the buffers the new u, v, w, T are written to (swapped in)
  bool _fused;
  Grid3DDevice<T> _u_next, _v_next, _w_next, _temp_next;

  void add_thermal_force();
  bool advance_fused(double dt);

 public:
  Eqn_IncompressibleNS3D();

  bool set_parameters(const Eqn_IncompressibleNS3DParams<T> &params);
  double get_max_stable_timestep() const;
  bool advance_one_step(double dt);

  T viscosity_coefficient() const { return _u_diffusion.coefficient; }
  T thermal_diffusion_coefficient() const { return _thermal_diffusion.coefficient; }

  const Grid3DDevice<T> &get_u() const { return _u; }
  const Grid3DDevice<T> &get_v() const { return _v; }
  const Grid3DDevice<T> &get_w() const { return _w; }
  const Grid3DDevice<T> &get_temperature() const { return _temp; }
  const Grid3DDevice<T> &get_pressure() const {
    if constexpr (std::is_same<T, double>::value)
      if (_mixed_projection) return _mixed_projection_solver.pressure_solver.pressure();
    return _projection_solver.pressure_solver.pressure();
  }
  bool mixed_projection() const { return _mixed_projection; }
  bool fused_step() const { return _fused; }
  Sol_ProjectDivergence3DDevice<T> &projection_solver() { return _projection_solver; }
};

typedef Eqn_IncompressibleNS3DParams<float> Eqn_IncompressibleNS3DParamsF;
typedef Eqn_IncompressibleNS3DParams<double> Eqn_IncompressibleNS3DParamsD;
typedef Eqn_IncompressibleNS3D<float> Eqn_IncompressibleNS3DF;
typedef Eqn_IncompressibleNS3D<double> Eqn_IncompressibleNS3DD;

}  // namespace ocu

#endif
