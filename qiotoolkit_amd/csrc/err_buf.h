// Message of the last failing call on the calling thread.
#pragma once
#include <cstring>
#include <string>

namespace qmc {

// A plain buffer, not a std::string: a thread_local with a destructor makes
// glibc keep the library loaded until the thread ends, and unloading the
// library is what lets its code-object unregistration run while the HIP
// runtime is still intact (see the Python handle registry).
struct ErrBuf {
  char s[1024];
  ErrBuf& operator=(const char* m) {
    std::strncpy(s, m ? m : "", sizeof(s) - 1);
    s[sizeof(s) - 1] = 0;
    return *this;
  }
  ErrBuf& operator=(const std::string& m) { return *this = m.c_str(); }
  const char* c_str() const { return s; }
  operator std::string() const { return std::string(s); }
};

}  // namespace qmc
