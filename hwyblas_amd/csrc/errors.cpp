// Error reporting for invalid arguments (the BLAS convention) and for HIP
// failures.
//
// Invalid arguments follow hwyBLAS (src/common.cc:19-44, :68-84,
// src/cblas.cc:714-726): the first invalid argument is reported once, through
// the handler installed with hwyblas_set_error_handler if there is one, else
// through cblas_xerbla (with the CBLAS name and position) while a CBLAS
// wrapper runs on this thread, else through xerbla_. Nothing aborts and the
// routine returns with its outputs untouched.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>

#include "common.h"

namespace hwy_amd {

namespace {
thread_local const CblasContext* t_cblas_context = nullptr;
std::atomic<hwyblas_error_handler> g_handler{nullptr};
}  // namespace

void SetCblasContext(const CblasContext* context) { t_cblas_context = context; }

bool CallErrorHandler(const char* name, blasint position) {
  const hwyblas_error_handler h = g_handler.load(std::memory_order_acquire);
  if (h == nullptr) return false;
  h(name, position);
  return true;
}

void ReportError(const char* name, blasint info) {
  if (const CblasContext* c = t_cblas_context) {
    const blasint position = c->positions[info - 1];
    if (!CallErrorHandler(c->name, position)) cblas_xerbla(position, c->name, "");
    return;
  }
  if (!CallErrorHandler(name, info)) xerbla_(name, &info, strlen(name));
}

void FatalHip(hipError_t e, const char* what, const char* file, int line) {
  fprintf(stderr, "hwyblas_amd: HIP error %d (%s) in %s at %s:%d\n",
          static_cast<int>(e), hipGetErrorString(e), what, file, line);
  fflush(stderr);
  abort();
}

}  // namespace hwy_amd

extern "C" {

HWYBLAS_AMD_API void xerbla_(const char* srname, const hwyblas_int* info,
                             size_t srname_len) {
  // A Fortran string: not necessarily NUL-terminated, blank padded.
  size_t len = 0;
  while (len < srname_len && srname[len] != '\0' && srname[len] != ' ') ++len;
  fprintf(stderr, " ** On entry to %.*s parameter number %d had an illegal value\n",
          static_cast<int>(len), srname, static_cast<int>(*info));
}

HWYBLAS_AMD_API void cblas_xerbla(hwyblas_int p, const char* rout,
                                  const char* form, ...) {
  if (p != 0) fprintf(stderr, "Parameter %d to routine %s was incorrect\n", p, rout);
  if (form != nullptr && *form != '\0') {
    va_list args;
    va_start(args, form);
    vfprintf(stderr, form, args);
    va_end(args);
  }
}

HWYBLAS_AMD_API hwyblas_error_handler
hwyblas_set_error_handler(hwyblas_error_handler handler) {
  return hwy_amd::g_handler.exchange(handler, std::memory_order_acq_rel);
}

}  // extern "C"
