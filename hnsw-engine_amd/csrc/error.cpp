#include "error.hpp"

#include <exception>
#include <new>

#include "../../include/hnswgpu.h"

namespace hnswamd {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& message) noexcept {
  try {
    g_last_error = message;
  } catch (...) {
  }
}

const std::string& last_error() noexcept { return g_last_error; }

int status_from_current_exception() noexcept {
  try {
    throw;
  } catch (const DeviceError& e) {
    set_last_error(e.what());
    return HNSWGPU_EDEVICE;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return HNSWGPU_EINVAL;
  } catch (const std::out_of_range& e) {
    set_last_error(e.what());
    return HNSWGPU_ERANGE;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return HNSWGPU_ERUNTIME;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return HNSWGPU_ERUNTIME;
  } catch (...) {
    set_last_error("unknown error");
    return HNSWGPU_ERUNTIME;
  }
}

[[noreturn]] void throw_status(int status, const std::string& message) {
  switch (status) {
    case HNSWGPU_EINVAL:
      throw std::invalid_argument(message);
    case HNSWGPU_ERANGE:
      throw std::out_of_range(message);
    case HNSWGPU_EDEVICE:
      throw DeviceError(message);
    default:
      throw std::runtime_error(message);
  }
}

}  // namespace hnswamd

extern "C" const char* hnswgpu_last_error(void) { return hnswamd::last_error().c_str(); }
