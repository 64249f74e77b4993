// errors.hpp -- C-ABI error convention: no exception crosses the boundary;
// failures return a negative status and leave a thread-local message.
#pragma once

#include <stdexcept>
#include <string>

namespace dpg {

void set_last_error(const std::string& msg);

template <class Fn>
int guarded(Fn&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::invalid_argument& e) {
    set_last_error(std::string("invalid argument: ") + e.what());
    return -1;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return -3;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -2;
  }
}

// same as guarded, but fn returns the (non-negative) status to forward
template <class Fn>
int guarded_status(Fn&& fn) {
  try {
    return fn();
  } catch (const std::invalid_argument& e) {
    set_last_error(std::string("invalid argument: ") + e.what());
    return -1;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return -3;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -2;
  }
}

}  // namespace dpg
