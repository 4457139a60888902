// Error convention of the host API: failures throw Plaquette::Exception
// (surfaced to Python as RuntimeError), with the message format and macro
// names of the reference's plaquette_simulator/src/Error.hpp:12-98.
#pragma once

#include <exception>
#include <sstream>
#include <string>

#define PLAQUETTE_ABORT(message) \
  Plaquette::Abort(message, __FILE__, __LINE__, __func__)
#define PLAQUETTE_ABORT_IF(expression, message) \
  if ((expression)) {                           \
    PLAQUETTE_ABORT(message);                   \
  }
#define PLAQUETTE_ABORT_IF_NOT(expression, message) \
  if (!(expression)) {                              \
    PLAQUETTE_ABORT(message);                       \
  }
#define PLAQUETTE_ASSERT(expression) \
  PLAQUETTE_ABORT_IF_NOT(expression, "Assertion failed: " #expression)

namespace Plaquette {

class Exception : public std::exception {
 public:
  explicit Exception(const std::string &err_msg) noexcept : err_msg_(err_msg) {}
  ~Exception() override = default;
  const char *what() const noexcept override { return err_msg_.c_str(); }

 private:
  std::string err_msg_;
};

[[noreturn]] inline void Abort(const std::string &message, const char *file_name,
                               int line, const char *function_name) {
  std::ostringstream m;
  m << "[" << file_name << "][Line:" << line << "][Method:" << function_name
    << "]: Error in Plaquette: " << message;
  throw Exception(m.str());
}

}  // namespace Plaquette
