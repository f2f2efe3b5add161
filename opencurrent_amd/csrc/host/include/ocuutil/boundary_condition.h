// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// BoundaryCondition / BoundaryConditionSet with the field order and widths of
// ocuutil/boundary_condition.h:26-112, so they are interchangeable with the
// C ABI's ocu_bc / ocu_bc_set (static_asserts in grid3dboundary.cpp).
#ifndef __OCU_UTIL_BOUNDARY_CONDITION_H__
#define __OCU_UTIL_BOUNDARY_CONDITION_H__

#include "ocuutil/defines.h"

namespace ocu {

enum BoundaryConditionType {
  BC_INVALID,
  BC_NONE,
  BC_PERIODIC,
  BC_DIRICHELET,
  BC_NEUMANN,
  BC_SECOND_DERIV,
  BC_DIRICHELET_AND_NEUMANN,
  BC_FORCED_INFLOW_VARIABLE_SLIP,
  BC_SCALAR_SLIP,
};

struct BoundaryCondition {
  BoundaryConditionType type;
  float value;
  float aux_value;

  BoundaryCondition() : type(BC_INVALID), value(0), aux_value(0) {}

  bool check_type(BoundaryConditionType t) const { return t == type; }
  bool check_type(BoundaryConditionType t1, BoundaryConditionType t2) const { return t1 == type || t2 == type; }
  bool check_type(BoundaryConditionType t1, BoundaryConditionType t2, BoundaryConditionType t3) const {
    return t1 == type || t2 == type || t3 == type;
  }
  bool check_type(BoundaryConditionType t1, BoundaryConditionType t2, BoundaryConditionType t3,
                  BoundaryConditionType t4) const {
    return t1 == type || t2 == type || t3 == type || t4 == type;
  }
};

struct BoundaryConditionSet {
  BoundaryCondition xpos, xneg, ypos, yneg, zpos, zneg;

  BoundaryConditionSet() {}
  BoundaryConditionSet(const BoundaryCondition &xpos_val, const BoundaryCondition &xneg_val,
                       const BoundaryCondition &ypos_val, const BoundaryCondition &yneg_val,
                       const BoundaryCondition &zpos_val, const BoundaryCondition &zneg_val)
      : xpos(xpos_val), xneg(xneg_val), ypos(ypos_val), yneg(yneg_val), zpos(zpos_val), zneg(zneg_val) {}
  BoundaryConditionSet(const BoundaryCondition &val)
      : xpos(val), xneg(val), ypos(val), yneg(val), zpos(val), zneg(val) {}

  bool check_type(BoundaryConditionType t) const {
    return xpos.check_type(t) && xneg.check_type(t) && ypos.check_type(t) && yneg.check_type(t) &&
           zpos.check_type(t) && zneg.check_type(t);
  }
  bool check_type(BoundaryConditionType t1, BoundaryConditionType t2) const {
    return xpos.check_type(t1, t2) && xneg.check_type(t1, t2) && ypos.check_type(t1, t2) &&
           yneg.check_type(t1, t2) && zpos.check_type(t1, t2) && zneg.check_type(t1, t2);
  }
  bool check_type(BoundaryConditionType t1, BoundaryConditionType t2, BoundaryConditionType t3) const {
    return xpos.check_type(t1, t2, t3) && xneg.check_type(t1, t2, t3) && ypos.check_type(t1, t2, t3) &&
           yneg.check_type(t1, t2, t3) && zpos.check_type(t1, t2, t3) && zneg.check_type(t1, t2, t3);
  }
  bool check_type(BoundaryConditionType t1, BoundaryConditionType t2, BoundaryConditionType t3,
                  BoundaryConditionType t4) const {
    return xpos.check_type(t1, t2, t3, t4) && xneg.check_type(t1, t2, t3, t4) &&
           ypos.check_type(t1, t2, t3, t4) && yneg.check_type(t1, t2, t3, t4) &&
           zpos.check_type(t1, t2, t3, t4) && zneg.check_type(t1, t2, t3, t4);
  }

  // no reference counterpart: bit a (x=1, y=2, z=4) set when both faces of
  // axis a are periodic
  int periodic_axes() const {
    auto pp = [](const BoundaryCondition &a, const BoundaryCondition &b) {
      return a.type == BC_PERIODIC && b.type == BC_PERIODIC;
    };
    return (pp(xneg, xpos) ? 1 : 0) | (pp(yneg, ypos) ? 2 : 0) | (pp(zneg, zpos) ? 4 : 0);
  }

  void make_homogeneous() {
    BoundaryCondition *all[6] = {&xpos, &xneg, &ypos, &yneg, &zpos, &zneg};
    for (auto b : all) { b->value = 0; b->aux_value = 0; }
  }
};

}  // namespace ocu

#endif
