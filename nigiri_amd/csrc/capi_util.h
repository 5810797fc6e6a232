// Error plumbing shared by the C ABI translation units: no exception may
// cross an extern "C" boundary, so every entry point runs its body through
// guarded(), which maps exceptions to status codes and records the message.
#pragma once

#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/nigiri_amd.h"

namespace nigiri_amd {

void set_error(std::string msg);

struct status_error : std::runtime_error {
  status_error(int code, std::string const& msg)
      : std::runtime_error{msg}, code_{code} {}
  int code_;
};

template <typename Fn>
int guarded(Fn&& fn) noexcept {
  try {
    return fn();
  } catch (status_error const& e) {
    set_error(e.what());
    return e.code_;
  } catch (std::bad_alloc const&) {
    set_error("out of host memory");
    return NIGIRI_AMD_ENOMEM;
  } catch (std::invalid_argument const& e) {
    set_error(e.what());
    return NIGIRI_AMD_EINVAL;
  } catch (std::exception const& e) {
    set_error(e.what());
    return NIGIRI_AMD_EINVAL;
  } catch (...) {
    set_error("unknown error");
    return NIGIRI_AMD_EINVAL;
  }
}

}  // namespace nigiri_amd
