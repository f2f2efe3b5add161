// ocuequation/parameters.h
#ifndef __OCU_EQUATION_PARAMETERS_H__
#define __OCU_EQUATION_PARAMETERS_H__

namespace ocu {
class Parameters {};
}  // namespace ocu

#endif
