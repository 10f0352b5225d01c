// Invalid-argument reports and HIP failures.
//
// There is no per-thread error state: each routine receives the identity of
// its caller (a Caller, common.h) and reports under it. A Fortran caller's
// identity is the routine name, and positions are Fortran argument numbers.
// A CBLAS caller's identity carries the CBLAS name and, for every Fortran
// argument, the CBLAS argument it was made from. The behaviour seen by a
// program is hwyBLAS's (src/common.cc:33-44, :68-84; src/cblas.cc:714-726):
// the first invalid argument goes to the handler installed with
// hwyblas_set_error_handler, else is printed by cblas_xerbla (CBLAS calls)
// or xerbla_ (Fortran calls); nothing aborts, and the routine returns with
// its outputs untouched.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>

#include "common.h"

namespace hwy_amd {
namespace {

// The installed handler; null means "print". Swapped atomically so that a
// handler installed on one thread is seen by routines running on others.
std::atomic<hwyblas_error_handler> installed_handler{nullptr};

}  // namespace

blasint Caller::Position(blasint info) const {
  return cblas_arg == nullptr ? info : blasint(cblas_arg[info - 1]);
}

void Reject(const Caller& who, blasint info) {
  const blasint position = who.Position(info);
  if (hwyblas_error_handler h = installed_handler.load()) {
    h(who.name, position);
  } else if (who.cblas_arg != nullptr) {
    cblas_xerbla(position, who.name, "");
  } else {
    xerbla_(who.name, &position, strlen(who.name));
  }
}

void RejectLayout(const char* name, int layout) {
  if (hwyblas_error_handler h = installed_handler.load()) {
    h(name, 1);
  } else {
    cblas_xerbla(1, name, "Illegal layout setting, %d\n", layout);
  }
}

void ReportInternal(const char* name, int code) {
  if (hwyblas_error_handler h = installed_handler.load()) h(name, code);
}

void FatalHip(hipError_t e, const char* what, const char* file, int line) {
  fprintf(stderr, "hwyblas_amd: HIP error %d (%s) in %s at %s:%d\n",
          static_cast<int>(e), hipGetErrorString(e), what, file, line);
  fflush(stderr);
  abort();
}

}  // namespace hwy_amd

extern "C" {

// The reference-BLAS message, without the reference BLAS's STOP. srname is a
// Fortran CHARACTER string: srname_len bytes, blank padded, possibly without
// a terminating NUL.
HWYBLAS_AMD_API void xerbla_(const char* srname, const hwyblas_int* info,
                             size_t srname_len) {
  const char* nul = static_cast<const char*>(memchr(srname, '\0', srname_len));
  size_t len = nul != nullptr ? size_t(nul - srname) : srname_len;
  while (len > 0 && srname[len - 1] == ' ') --len;
  fprintf(stderr, " ** On entry to %.*s parameter number %d had an illegal value\n",
          static_cast<int>(len), srname, static_cast<int>(*info));
}

// The CBLAS reference message for position p (none for p == 0), then the
// optional printf-style detail.
HWYBLAS_AMD_API void cblas_xerbla(hwyblas_int p, const char* rout,
                                  const char* form, ...) {
  if (p != 0) fprintf(stderr, "Parameter %d to routine %s was incorrect\n", p, rout);
  if (form == nullptr || form[0] == '\0') return;
  va_list args;
  va_start(args, form);
  vfprintf(stderr, form, args);
  va_end(args);
}

HWYBLAS_AMD_API hwyblas_error_handler
hwyblas_set_error_handler(hwyblas_error_handler handler) {
  return hwy_amd::installed_handler.exchange(handler);
}

}  // extern "C"
