// Error codes of the reference (utils/error_handling.h:25-52,
// utils/exception.h): insufficient resources 1-2, invalid input data
// (ConfigurationException family) 101-109, runtime errors 201-206, file IO
// 301-302.  Codes from 401 have no reference counterpart: the device or its
// runtime failed, or an unexpected failure inside the engine.
#pragma once
#include <stdexcept>
#include <string>

namespace qiotk {

enum ErrorCode {
  kMemoryLimited = 1,
  kTimeoutInsufficient = 2,
  kDuplicatedVariable = 101,
  kMissingInput = 102,
  kInvalidTypes = 103,
  kInitialConfig = 104,
  kParsingError = 105,
  kFeatureSwitch = 106,
  kValueError = 107,
  kIndexOutOfRange = 108,
  kNoFeasibleSolution = 109,
  kNotImplemented = 201,
  kKeyDoesNotExist = 202,
  kAccessAsIncorrectType = 203,
  kPopulationIsEmpty = 204,
  kInconsistentValue = 205,
  kUnmetPrecondition = 206,
  kFileReadError = 301,
  kFileWriteError = 302,
  kDeviceError = 401,
  kRuntimeError = 402,
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void device_check(int rc, const char* what);

}  // namespace qiotk

#include "../../../include/qmc.h"
namespace qiotk {
inline void device_check(int rc, const char* what) {
  if (rc != QMC_OK) {
    int code = rc == QMC_MEMORY_LIMITED ? kMemoryLimited
             : (rc >= 101 && rc < 400) ? rc
                                       : kDeviceError;
    throw Error(code, std::string(what) + ": " + qmc_last_error());
  }
}
}  // namespace qiotk
