// Helpers of the C ABI entry points: errors become a status with (operation, detail).
#pragma once
#include <cstdio>
#include <new>

#include "kernels.hpp"

namespace ddgrs {

inline void clear_error(ddgrs_error* err) {
  if (!err) return;
  err->status = DDGRS_OK;
  err->operation[0] = 0;
  err->detail[0] = 0;
  err->a = err->b = -1;
  err->value = 0.0;
}

inline int report_error(ddgrs_error* err, int status, const std::string& op, const std::string& detail, int64_t a = -1,
                        int64_t b = -1, double value = 0.0) {
  if (err) {
    err->status = status;
    std::snprintf(err->operation, sizeof err->operation, "%s", op.c_str());
    std::snprintf(err->detail, sizeof err->detail, "%s", detail.c_str());
    err->a = a;
    err->b = b;
    err->value = value;
  }
  return status;
}

template <class F>
int guarded(ddgrs_error* err, const char* op, F&& f) {
  clear_error(err);
  try {
    f();
    return DDGRS_OK;
  } catch (const Error& e) {
    return report_error(err, e.status, e.operation.empty() ? op : e.operation, e.what(), e.a, e.b, e.value);
  } catch (const std::bad_alloc&) {
    return report_error(err, DDGRS_GPU, op, "host allocation failed");
  } catch (const std::exception& e) {
    return report_error(err, DDGRS_GPU, op, e.what());
  }
}

inline void need(const void* p, const char* name) {
  if (!p) invalid(name, "null", "a valid pointer");
}

}  // namespace ddgrs
