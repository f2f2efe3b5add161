/*
 * ocu_types.h - plain-C types shared by the HIP kernel library (ocu_hip.h),
 * the host solver library (ocu.h) and their callers.
 *
 * No HIP, CUDA or torch types appear here: a grid is a base pointer plus the
 * padded layout of OpenCurrent's Grid3D (ocustorage/grid3d.h:29-147,
 * ocustorage/grid3ddev.cu:358-400).  Cell (i,j,k) of a grid lives at element
 *     shift + i*pny*pnz + j*pnz + k
 * of the allocation, i.e. z (k) has unit stride and x (i) is the slowest axis.
 */
#ifndef OCU_TYPES_H
#define OCU_TYPES_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types of a grid */
enum ocu_dtype { OCU_F32 = 0, OCU_F64 = 1, OCU_I32 = 2 };

/* Boundary-condition kinds.  Same numeric values as ocu::BoundaryConditionType
 * (ocuutil/boundary_condition.h:26-36) so a reference BoundaryConditionSet can
 * be passed by value. */
enum ocu_bc_type {
  OCU_BC_INVALID = 0,
  OCU_BC_NONE = 1,
  OCU_BC_PERIODIC = 2,
  OCU_BC_DIRICHELET = 3,
  OCU_BC_NEUMANN = 4,
  OCU_BC_SECOND_DERIV = 5,
  OCU_BC_DIRICHELET_AND_NEUMANN = 6,
  OCU_BC_FORCED_INFLOW_VARIABLE_SLIP = 7,
  OCU_BC_SCALAR_SLIP = 8
};

/* ocu::BoundaryCondition (ocuutil/boundary_condition.h:38-62): same field
 * order and widths. */
typedef struct ocu_bc {
  int32_t type;
  float value;
  float aux_value;
} ocu_bc;

/* ocu::BoundaryConditionSet (ocuutil/boundary_condition.h:66-112). */
typedef struct ocu_bc_set {
  ocu_bc xpos, xneg, ypos, yneg, zpos, zneg;
} ocu_bc_set;

/* Convergence criteria, same bit values as ocu::ConvergenceType
 * (ocuutil/convergence.h:27-40). */
enum ocu_convergence {
  OCU_CONVERGENCE_CALC_L2 = 0x01,
  OCU_CONVERGENCE_CALC_LINF = 0x02,
  OCU_CONVERGENCE_CRITERIA_L2 = 0x04,
  OCU_CONVERGENCE_CRITERIA_LINF = 0x08,
  OCU_CONVERGENCE_CRITERIA_NONE = 0x10,
  OCU_CONVERGENCE_L2 = 0x05,
  OCU_CONVERGENCE_LINF = 0x0a,
  OCU_CONVERGENCE_NONE = 0x10,
  OCU_CONVERGENCE_NONE_CALC_L2LINF = 0x13
};

/* Interpolation schemes (ocuutil/interpolation.h:26-31).  The value 4 has
 * no reference counterpart: second-order (linear) upwind, the face value
 * extrapolated from the two upstream cells, x_up + (x_up - x_upup)/2; across
 * a periodic axis the stencil wraps, and at a face whose second upstream
 * cell lies outside the one-cell ghost layer of a non-periodic axis it is
 * first-order upwind. */
enum ocu_interp {
  OCU_IT_ERROR = 0,
  OCU_IT_FIRST_ORDER_UPWIND = 1,
  OCU_IT_SECOND_ORDER_CENTERED = 2,
  OCU_IT_THIRD_ORDER_VAN_LEER_FLUX_LIMITED = 3,
  OCU_IT_SECOND_ORDER_UPWIND = 4
};

/* Direction flags (ocuutil/direction.h:24-38). */
enum ocu_direction {
  OCU_DIR_POSITIVE_FLAG = 0x01,
  OCU_DIR_NEGATIVE_FLAG = 0x02,
  OCU_DIR_XAXIS_FLAG = 0x04,
  OCU_DIR_YAXIS_FLAG = 0x08,
  OCU_DIR_ZAXIS_FLAG = 0x10,
  OCU_DIR_XPOS = 0x05,
  OCU_DIR_XNEG = 0x06,
  OCU_DIR_YPOS = 0x09,
  OCU_DIR_YNEG = 0x0a,
  OCU_DIR_ZPOS = 0x11,
  OCU_DIR_ZNEG = 0x12
};

/* A 3-D grid: allocation start, interior size, ghost widths, padded extents.
 * `shift` is the element offset of cell (0,0,0) from `base`, `allocated` the
 * number of elements in the allocation (grid3ddev.cu:387-397). */
typedef struct ocu_grid3d {
  void *base;
  int32_t nx, ny, nz;
  int32_t gx, gy, gz;
  int32_t pnx, pny, pnz;
  int32_t dtype;
  int64_t shift;
  int64_t allocated;
} ocu_grid3d;

/* Red-black Gauss-Seidel arithmetic variants.  The reference has three
 * different operation orders for the same update; each is kept so results
 * can be compared bit-for-bit:
 *   OCU_RELAX_DEVICE: T=double follows Sol_MultigridPressure3DDeviceD_relax_DIAG
 *     (ocuequation/sol_mgpressure3ddevd.cu:542-599), T=float follows
 *     Sol_MultigridPressure3DDeviceF_relax (sol_mgpressure3ddevf.cu:200-252).
 *     Both launch nz/2 threads along k, so for odd nz the last k plane of
 *     one parity is never updated - kept as-is.
 *   OCU_RELAX_HOST: Sol_MultigridPressure3DDevice<T>::do_cpu_solve
 *     (sol_mgpressure3ddev.cu:82-125), the coarsest-level relaxation, whose
 *     coefficient products are formed in double even for T=float. */
enum ocu_relax_variant { OCU_RELAX_DEVICE = 0, OCU_RELAX_HOST = 1 };
/* options of ocu_mg_relax_fused_bc */
enum ocu_fused_flags { OCU_FUSED_WRITE_GHOSTS = 1, OCU_FUSED_ZERO_INPUT = 2 };

/* Per-level constants of the multigrid operator
 *   f_x (u_{i+1}+u_{i-1}) + f_y (..) + f_z (..) - 2(f_x+f_y+f_z) u = h^2 b
 * with h = 2^level * min(hx,hy,hz) and f_a = h0^2 / ha^2
 * (sol_mgpressure3d.cpp:401-437).  mods[] are bc_diag_mod() values in the
 * order xpos, xneg, ypos, yneg, zpos, zneg (sol_mgpressure3d.cpp:49-64). */
typedef struct ocu_mg_level_params {
  double h;
  double fx, fy, fz;
  double omega;
  double mods[6];
} ocu_mg_level_params;

/* Public configuration of a multigrid pressure solver: the public fields of
 * Sol_MultigridPressure3DBase / Sol_MultigridPressure3DDevice
 * (sol_mgpressure3d.h:71-77, 133) plus the prolongation variant.
 * prolong_quirk = 1 reproduces the reference prolongation exactly
 * (sol_mgpressure3ddevd.cu:181-185); 0 uses the symmetric corner. */
typedef struct ocu_mg_config {
  ocu_bc_set bc;
  int32_t nu1, nu2;
  int32_t convergence;
  int32_t make_symmetric_operator;
  int32_t disable_relax_on_host;
  int32_t prolong_quirk;
} ocu_mg_config;

/* One explicit NS time step done by ocu_ns_step_fused: the tendencies of
 * Eqn_IncompressibleNS3D<T>::advance_one_step (eqn_incompressns3d.cu:331-393)
 * and their forward-Euler / Adams-Bashforth combination. */
typedef struct ocu_ns_step_params {
  int32_t nx, ny, nz;            /* cells */
  int32_t advection_scheme;      /* OCU_IT_* */
  int32_t vertical_direction;    /* OCU_DIR_* (thermal force) */
  int32_t flow_periodic;         /* periodic-axes mask (bit 0 x, 1 y, 2 z) of the flow BCs */
  int32_t temp_periodic;         /* the same for the temperature BCs */
  int32_t keep_tendencies;       /* 1: store this step's tendencies (Adams-Bashforth) */
  int32_t use_last;              /* 1: add a_ab * the last step's tendencies */
  double hx, hy, hz;
  double viscosity;              /* <= 0: no momentum diffusion */
  double thermal_diffusion;      /* <= 0: no temperature diffusion */
  double thermal_coef;           /* dirmult * gravity * bouyancy, in T */
  double a_dt;                   /* dt (Euler) or dt - ab (Adams-Bashforth), in T */
  double a_ab;                   /* ab = -dt^2 / (2 dt_last), in T */
} ocu_ns_step_params;

/* Status codes of every ocu_* entry point (0 = success). */
enum ocu_status {
  OCU_OK = 0,
  OCU_ERR_INVALID = -1,     /* bad argument: shape, layout, ghost width, bc type */
  OCU_ERR_RUNTIME = -2,     /* the HIP runtime reported an error */
  OCU_ERR_NOMEM = -3,       /* allocation failed */
  OCU_ERR_UNSUPPORTED = -4, /* valid request this build does not implement */
  OCU_ERR_NOT_CONVERGED = -5
};

#ifdef __cplusplus
}
#endif

#endif /* OCU_TYPES_H */
