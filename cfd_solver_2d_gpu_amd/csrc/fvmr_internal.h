// Internal declarations shared by the host mesh layer and the device solver.
#pragma once
#include <string>

namespace fvmr {

// Records the message of the last failure of the calling thread and returns code.
int set_error(int code, const std::string &msg);

// Physical limits of FVMR/src/math/CommonValue.h:16-31.
constexpr double kEPS = 1e-10;
constexpr double kRHO_MIN = kEPS;
constexpr double kRHO_MAX = 300.0;
constexpr double kP_MIN = kEPS;
constexpr double kP_MAX = 400.0 * 101325.0;
// PI as the reference writes it (CommonValue.h:19), used by the initialisers.
constexpr double kPI_REF = 3.1415926535897;

}  // namespace fvmr
