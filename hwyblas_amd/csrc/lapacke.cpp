// LAPACKE (C interface) for the Cholesky routines (LAPACKE_{s,d}potrf,
// potrs, posv and their _work forms) and the LU solver (LAPACKE_dgetrf,
// dgetrs, dgesv and their _work forms), with LAPACKE_xerbla and the
// NaN-check switch (hwyBLAS src/lapack/lapacke.cc:112-133, 2065-2092;
// src/lapack/lapacke_symmetric.cc:22-70; src/lapack/lapacke_util.h:89-119,
// 395-480).
//
// LAPACKE_x: an invalid layout is argument -1; then, unless the NaN check is
// off (LAPACKE_set_nancheck(0) or LAPACKE_NANCHECK=0 in the environment),
// a NaN in the referenced triangle of A or in B returns minus that
// argument's position; then LAPACKE_x_work. LAPACKE_x_work: column-major
// calls the Fortran routine and returns its info, minus one more for an
// invalid argument (the layout is argument 1). Row-major checks each leading
// dimension against the columns (returning minus its position), copies the
// referenced triangle of A and all of B into column-major temporaries, calls
// the routine and copies back the outputs (A's triangle only). Operands are
// host memory, as in the reference; the Fortran routines run on the GPU.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>

#include "common.h"

#define HWY_LAPACK_ROW_MAJOR 101
#define HWY_LAPACK_COL_MAJOR 102
#define HWY_LAPACK_TRANSPOSE_MEMORY_ERROR (-1011)

using hwy_amd::blasint;

namespace hwy_amd {
namespace {

std::atomic<int> g_nancheck{-1};  // -1: not read from the environment yet

bool Lsame(char a, char b) { return (a | 0x20) == (b | 0x20); }

template <typename T>
struct Lapack;
template <>
struct Lapack<double> {
  static constexpr auto potrf = dpotrf_;
  static constexpr auto potrs = dpotrs_;
  static constexpr auto posv = dposv_;
  static constexpr char kPre = 'd';
};
template <>
struct Lapack<float> {
  static constexpr auto potrf = spotrf_;
  static constexpr auto potrs = spotrs_;
  static constexpr auto posv = sposv_;
  static constexpr char kPre = 's';
};

template <typename T>
std::string Name(const char* rest) {
  return std::string("LAPACKE_") + Lapack<T>::kPre + rest;
}

void Xerbla(const std::string& name, blasint info) { LAPACKE_xerbla(name.c_str(), info); }

bool ValidLayout(int layout) {
  return layout == HWY_LAPACK_ROW_MAJOR || layout == HWY_LAPACK_COL_MAJOR;
}

// Whether the triangle uplo of the n x n matrix (either layout) holds a NaN.
// Column-major upper and row-major lower are the same walk.
template <typename T>
bool TriangleHasNan(int layout, char uplo, blasint n, const T* a, blasint lda) {
  if (a == nullptr) return false;
  const bool col = layout == HWY_LAPACK_COL_MAJOR;
  const bool lower = Lsame(uplo, 'L');
  if (!lower && !Lsame(uplo, 'U')) return false;
  if (col != lower) {
    for (blasint j = 0; j < n; ++j)
      for (blasint i = 0; i < std::min(j + 1, lda); ++i)
        if (std::isnan(a[i + size_t(j) * lda])) return true;
  } else {
    for (blasint j = 0; j < n; ++j)
      for (blasint i = j; i < std::min(n, lda); ++i)
        if (std::isnan(a[i + size_t(j) * lda])) return true;
  }
  return false;
}

// Whether the m x n matrix (either layout) holds a NaN.
template <typename T>
bool GeneralHasNan(int layout, blasint m, blasint n, const T* a, blasint lda) {
  if (a == nullptr) return false;
  const blasint outer = layout == HWY_LAPACK_COL_MAJOR ? n : m;
  const blasint inner = layout == HWY_LAPACK_COL_MAJOR ? m : n;
  for (blasint j = 0; j < outer; ++j)
    for (blasint i = 0; i < std::min(inner, lda); ++i)
      if (std::isnan(a[i + size_t(j) * lda])) return true;
  return false;
}

// Row-major <-> column-major copies: to_col copies row-major `row` into
// column-major `col`, otherwise back. tri: only the triangle uplo of the
// n x n matrix ('U' unless uplo is 'L').
template <typename T>
void Copy(bool to_col, bool tri, char uplo, blasint rows, blasint cols, T* row, blasint ldr,
          T* col, blasint ldc) {
  const bool lower = Lsame(uplo, 'L');
  for (blasint j = 0; j < cols; ++j) {
    blasint i0 = 0, i1 = rows;
    if (tri) {
      i0 = lower ? j : 0;
      i1 = lower ? rows : std::min(j + 1, rows);
    }
    for (blasint i = i0; i < i1; ++i) {
      T& r = row[size_t(i) * ldr + j];
      T& c = col[i + size_t(j) * ldc];
      if (to_col) c = r;
      else r = c;
    }
  }
}

// A matrix argument of a _work function (rows x cols as the routine sees
// it), with the position reported for a short row-major leading dimension.
template <typename T>
struct Arg {
  T* data;
  blasint rows, cols, ld, ld_info;
  bool tri, out;
};

template <typename T, int N, typename Call>
blasint CallWork(int layout, const std::string& name, char uplo, Arg<T> (&args)[N],
                 const Call& call) {
  T* p[N];
  blasint ld[N];
  if (layout == HWY_LAPACK_COL_MAJOR) {
    for (int i = 0; i < N; ++i) {
      p[i] = args[i].data;
      ld[i] = args[i].ld;
    }
    const blasint info = call(p, ld);
    return info < 0 ? info - 1 : info;
  }
  if (layout != HWY_LAPACK_ROW_MAJOR) {
    Xerbla(name, -1);
    return -1;
  }
  for (int i = 0; i < N; ++i) {
    if (args[i].ld < args[i].cols) {
      Xerbla(name, args[i].ld_info);
      return args[i].ld_info;
    }
  }
  const char tu = Lsame(uplo, 'U') ? 'U' : 'L';
  for (int i = 0; i < N; ++i) {
    ld[i] = std::max<blasint>(1, args[i].rows);
    p[i] = static_cast<T*>(malloc(sizeof(T) * size_t(ld[i]) *
                                  size_t(std::max<blasint>(1, args[i].cols))));
    if (p[i] == nullptr) {
      for (int q = 0; q < i; ++q) free(p[q]);
      Xerbla(name, HWY_LAPACK_TRANSPOSE_MEMORY_ERROR);
      return HWY_LAPACK_TRANSPOSE_MEMORY_ERROR;
    }
    Copy(true, args[i].tri, tu, args[i].rows, args[i].cols, args[i].data, args[i].ld, p[i],
         ld[i]);
  }
  blasint info = call(p, ld);
  if (info < 0) info -= 1;
  for (int i = 0; i < N; ++i) {
    if (args[i].out)
      Copy(false, args[i].tri, tu, args[i].rows, args[i].cols, args[i].data, args[i].ld, p[i],
           ld[i]);
    free(p[i]);
  }
  return info;
}

bool NanCheck() { return LAPACKE_get_nancheck() != 0; }

template <typename T>
blasint PotrfWork(int layout, char uplo, blasint n, T* a, blasint lda) {
  Arg<T> args[1] = {{a, n, n, lda, -5, true, true}};
  return CallWork<T, 1>(layout, Name<T>("potrf_work"), uplo, args, [&](T** p, blasint* ld) {
    blasint info = 0;
    Lapack<T>::potrf(&uplo, &n, p[0], &ld[0], &info);
    return info;
  });
}

template <typename T>
blasint Potrf(int layout, char uplo, blasint n, T* a, blasint lda) {
  if (!ValidLayout(layout)) {
    Xerbla(Name<T>("potrf"), -1);
    return -1;
  }
  if (NanCheck() && TriangleHasNan(layout, uplo, n, a, lda)) return -4;
  return PotrfWork(layout, uplo, n, a, lda);
}

template <typename T>
blasint PotrsWork(int layout, char uplo, blasint n, blasint nrhs, const T* a, blasint lda, T* b,
                  blasint ldb) {
  Arg<T> args[2] = {{const_cast<T*>(a), n, n, lda, -6, true, false},
                    {b, n, nrhs, ldb, -8, false, true}};
  return CallWork<T, 2>(layout, Name<T>("potrs_work"), uplo, args, [&](T** p, blasint* ld) {
    blasint info = 0;
    Lapack<T>::potrs(&uplo, &n, &nrhs, p[0], &ld[0], p[1], &ld[1], &info);
    return info;
  });
}

template <typename T>
blasint Potrs(int layout, char uplo, blasint n, blasint nrhs, const T* a, blasint lda, T* b,
              blasint ldb) {
  if (!ValidLayout(layout)) {
    Xerbla(Name<T>("potrs"), -1);
    return -1;
  }
  if (NanCheck()) {
    if (TriangleHasNan(layout, uplo, n, a, lda)) return -5;
    if (GeneralHasNan(layout, n, nrhs, b, ldb)) return -7;
  }
  return PotrsWork(layout, uplo, n, nrhs, a, lda, b, ldb);
}

template <typename T>
blasint PosvWork(int layout, char uplo, blasint n, blasint nrhs, T* a, blasint lda, T* b,
                 blasint ldb) {
  Arg<T> args[2] = {{a, n, n, lda, -6, true, true}, {b, n, nrhs, ldb, -8, false, true}};
  return CallWork<T, 2>(layout, Name<T>("posv_work"), uplo, args, [&](T** p, blasint* ld) {
    blasint info = 0;
    Lapack<T>::posv(&uplo, &n, &nrhs, p[0], &ld[0], p[1], &ld[1], &info);
    return info;
  });
}

template <typename T>
blasint Posv(int layout, char uplo, blasint n, blasint nrhs, T* a, blasint lda, T* b,
             blasint ldb) {
  if (!ValidLayout(layout)) {
    Xerbla(Name<T>("posv"), -1);
    return -1;
  }
  if (NanCheck()) {
    if (TriangleHasNan(layout, uplo, n, a, lda)) return -5;
    if (GeneralHasNan(layout, n, nrhs, b, ldb)) return -7;
  }
  return PosvWork(layout, uplo, n, nrhs, a, lda, b, ldb);
}

// LU (double only, as the Fortran entries): hwyBLAS
// src/lapack/lapacke_linear.cc:14-75, lapacke.cc:40-58.
blasint GetrfWork(int layout, blasint m, blasint n, double* a, blasint lda, blasint* ipiv) {
  Arg<double> args[1] = {{a, m, n, lda, -5, false, true}};
  return CallWork<double, 1>(layout, "LAPACKE_dgetrf_work", 'L', args,
                             [&](double** p, blasint* ld) {
                               blasint info = 0;
                               dgetrf_(&m, &n, p[0], &ld[0], ipiv, &info);
                               return info;
                             });
}

blasint Getrf(int layout, blasint m, blasint n, double* a, blasint lda, blasint* ipiv) {
  if (!ValidLayout(layout)) {
    Xerbla("LAPACKE_dgetrf", -1);
    return -1;
  }
  if (NanCheck() && GeneralHasNan(layout, m, n, a, lda)) return -4;
  return GetrfWork(layout, m, n, a, lda, ipiv);
}

blasint GetrsWork(int layout, char trans, blasint n, blasint nrhs, const double* a, blasint lda,
                  const blasint* ipiv, double* b, blasint ldb) {
  Arg<double> args[2] = {{const_cast<double*>(a), n, n, lda, -6, false, false},
                         {b, n, nrhs, ldb, -9, false, true}};
  return CallWork<double, 2>(layout, "LAPACKE_dgetrs_work", 'L', args,
                             [&](double** p, blasint* ld) {
                               blasint info = 0;
                               dgetrs_(&trans, &n, &nrhs, p[0], &ld[0], ipiv, p[1], &ld[1],
                                       &info);
                               return info;
                             });
}

blasint Getrs(int layout, char trans, blasint n, blasint nrhs, const double* a, blasint lda,
              const blasint* ipiv, double* b, blasint ldb) {
  if (!ValidLayout(layout)) {
    Xerbla("LAPACKE_dgetrs", -1);
    return -1;
  }
  if (NanCheck()) {
    if (GeneralHasNan(layout, n, n, a, lda)) return -5;
    if (GeneralHasNan(layout, n, nrhs, b, ldb)) return -8;
  }
  return GetrsWork(layout, trans, n, nrhs, a, lda, ipiv, b, ldb);
}

blasint GesvWork(int layout, blasint n, blasint nrhs, double* a, blasint lda, blasint* ipiv,
                 double* b, blasint ldb) {
  Arg<double> args[2] = {{a, n, n, lda, -5, false, true}, {b, n, nrhs, ldb, -8, false, true}};
  return CallWork<double, 2>(layout, "LAPACKE_dgesv_work", 'L', args,
                             [&](double** p, blasint* ld) {
                               blasint info = 0;
                               dgesv_(&n, &nrhs, p[0], &ld[0], ipiv, p[1], &ld[1], &info);
                               return info;
                             });
}

blasint Gesv(int layout, blasint n, blasint nrhs, double* a, blasint lda, blasint* ipiv,
             double* b, blasint ldb) {
  if (!ValidLayout(layout)) {
    Xerbla("LAPACKE_dgesv", -1);
    return -1;
  }
  if (NanCheck()) {
    if (GeneralHasNan(layout, n, n, a, lda)) return -4;
    if (GeneralHasNan(layout, n, nrhs, b, ldb)) return -7;
  }
  return GesvWork(layout, n, nrhs, a, lda, ipiv, b, ldb);
}

}  // namespace
}  // namespace hwy_amd

extern "C" {

HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrf(int layout, hwyblas_int m, hwyblas_int n, double* a,
                                           hwyblas_int lda, hwyblas_int* ipiv) {
  return hwy_amd::Getrf(layout, m, n, a, lda, ipiv);
}
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrf_work(int layout, hwyblas_int m, hwyblas_int n,
                                                double* a, hwyblas_int lda, hwyblas_int* ipiv) {
  return hwy_amd::GetrfWork(layout, m, n, a, lda, ipiv);
}
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrs(int layout, char trans, hwyblas_int n,
                                           hwyblas_int nrhs, const double* a, hwyblas_int lda,
                                           const hwyblas_int* ipiv, double* b, hwyblas_int ldb) {
  return hwy_amd::Getrs(layout, trans, n, nrhs, a, lda, ipiv, b, ldb);
}
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrs_work(int layout, char trans, hwyblas_int n,
                                                hwyblas_int nrhs, const double* a,
                                                hwyblas_int lda, const hwyblas_int* ipiv,
                                                double* b, hwyblas_int ldb) {
  return hwy_amd::GetrsWork(layout, trans, n, nrhs, a, lda, ipiv, b, ldb);
}
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgesv(int layout, hwyblas_int n, hwyblas_int nrhs,
                                          double* a, hwyblas_int lda, hwyblas_int* ipiv,
                                          double* b, hwyblas_int ldb) {
  return hwy_amd::Gesv(layout, n, nrhs, a, lda, ipiv, b, ldb);
}
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgesv_work(int layout, hwyblas_int n, hwyblas_int nrhs,
                                               double* a, hwyblas_int lda, hwyblas_int* ipiv,
                                               double* b, hwyblas_int ldb) {
  return hwy_amd::GesvWork(layout, n, nrhs, a, lda, ipiv, b, ldb);
}

HWYBLAS_AMD_API void LAPACKE_xerbla(const char* name, hwyblas_int info) {
  if (info == -1010) {
    printf("Not enough memory to allocate work array in %s\n", name);
  } else if (info == HWY_LAPACK_TRANSPOSE_MEMORY_ERROR) {
    printf("Not enough memory to transpose matrix in %s\n", name);
  } else if (info < 0) {
    printf("Wrong parameter %d in %s\n", -static_cast<int>(info), name);
  }
  fflush(stdout);
}

HWYBLAS_AMD_API void LAPACKE_set_nancheck(int flag) {
  hwy_amd::g_nancheck.store(flag ? 1 : 0);
}

HWYBLAS_AMD_API int LAPACKE_get_nancheck(void) {
  const int flag = hwy_amd::g_nancheck.load();
  if (flag != -1) return flag;
  const char* env = getenv("LAPACKE_NANCHECK");
  const int value = (env == nullptr || atoi(env) != 0) ? 1 : 0;
  int expected = -1;
  hwy_amd::g_nancheck.compare_exchange_strong(expected, value);
  return hwy_amd::g_nancheck.load();
}

#define HWY_LAPACKE_CHOL(p, T)                                                                  \
  HWYBLAS_AMD_API hwyblas_int LAPACKE_##p##potrf(int layout, char uplo, hwyblas_int n, T* a,    \
                                                 hwyblas_int lda) {                             \
    return hwy_amd::Potrf<T>(layout, uplo, n, a, lda);                                          \
  }                                                                                             \
  HWYBLAS_AMD_API hwyblas_int LAPACKE_##p##potrf_work(int layout, char uplo, hwyblas_int n,     \
                                                      T* a, hwyblas_int lda) {                  \
    return hwy_amd::PotrfWork<T>(layout, uplo, n, a, lda);                                      \
  }                                                                                             \
  HWYBLAS_AMD_API hwyblas_int LAPACKE_##p##potrs(int layout, char uplo, hwyblas_int n,          \
                                                 hwyblas_int nrhs, const T* a, hwyblas_int lda, \
                                                 T* b, hwyblas_int ldb) {                       \
    return hwy_amd::Potrs<T>(layout, uplo, n, nrhs, a, lda, b, ldb);                            \
  }                                                                                             \
  HWYBLAS_AMD_API hwyblas_int LAPACKE_##p##potrs_work(int layout, char uplo, hwyblas_int n,     \
                                                      hwyblas_int nrhs, const T* a,             \
                                                      hwyblas_int lda, T* b, hwyblas_int ldb) { \
    return hwy_amd::PotrsWork<T>(layout, uplo, n, nrhs, a, lda, b, ldb);                        \
  }                                                                                             \
  HWYBLAS_AMD_API hwyblas_int LAPACKE_##p##posv(int layout, char uplo, hwyblas_int n,           \
                                                hwyblas_int nrhs, T* a, hwyblas_int lda, T* b,  \
                                                hwyblas_int ldb) {                              \
    return hwy_amd::Posv<T>(layout, uplo, n, nrhs, a, lda, b, ldb);                             \
  }                                                                                             \
  HWYBLAS_AMD_API hwyblas_int LAPACKE_##p##posv_work(int layout, char uplo, hwyblas_int n,      \
                                                     hwyblas_int nrhs, T* a, hwyblas_int lda,   \
                                                     T* b, hwyblas_int ldb) {                   \
    return hwy_amd::PosvWork<T>(layout, uplo, n, nrhs, a, lda, b, ldb);                         \
  }

HWY_LAPACKE_CHOL(s, float)
HWY_LAPACKE_CHOL(d, double)

}  // extern "C"
