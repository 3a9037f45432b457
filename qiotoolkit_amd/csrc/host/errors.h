// Error taxonomy of the reference (utils/error_handling.h:25-52,
// utils/exception.h): user errors (ConfigurationException family, codes
// 101-109) and runtime errors (201-206); MemoryLimited = 1.
#pragma once
#include <stdexcept>
#include <string>

namespace qiotk {

enum ErrorCode {
  kMemoryLimited = 1,
  kMissingInput = 101,
  kInitialConfig = 102,
  kValueError = 103,
  kIndexOutOfRange = 104,
  kDuplicatedVariable = 105,
  kNotImplemented = 106,
  kRuntimeError = 201,
  kDeviceError = 202,
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
             : (rc >= 101 && rc < 200) ? rc
                                       : kDeviceError;
    throw Error(code, std::string(what) + ": " + qmc_last_error());
  }
}
}  // namespace qiotk
