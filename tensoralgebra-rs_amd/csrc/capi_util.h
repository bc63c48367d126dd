// The entry wrapper every extern "C" function uses: exceptions never cross the
// C boundary; a tp::Error becomes its status plus the thread-local message,
// anything else TPRIMS_ERR_PANIC (crates/tprims-capi/src/status.rs:139-162).
#pragma once
#include <exception>
#include <new>
#include <string>

#include "common.h"

namespace tp {

void set_last_error(const std::string& m);
void clear_last_error();
// Reads alpha or beta (`name`) as the plan's element type into (re, im);
// null is TPRIMS_ERR_INVALID_ARGUMENT.
void read_scalar(const void* p, DType t, const char* name, double out[2]);

template <class F>
int ffi(F&& body) {
  try {
    body();
    clear_last_error();
    return OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("allocation failure");
    return ERR_INTERNAL;
  } catch (const std::exception& e) {
    set_last_error(std::string("panic: ") + e.what());
    return ERR_PANIC;
  } catch (...) {
    set_last_error("panic");
    return ERR_PANIC;
  }
}

}  // namespace tp
