// CPython binding of the C ABI in include/pavelsvd.h (module `_hip` of the package).
//
// Every function takes plain integers (device pointers, sizes, a hipStream_t handle) and returns
// the C status code; no PyTorch types cross this boundary.  Built with g++ against
// libpavelsvd.so, which holds the gfx950 kernels.
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include "pavelsvd.h"

#define PTR(x) reinterpret_cast<void *>(static_cast<uintptr_t>(x))

static PyObject *py_version(PyObject *, PyObject *) { return PyUnicode_FromString(psvd_version()); }

static PyObject *py_jacobi_lds_workspace_bytes(PyObject *, PyObject *args) {
  long long batch;
  if (!PyArg_ParseTuple(args, "L", &batch)) return nullptr;
  return PyLong_FromSize_t(psvd_jacobi_lds_workspace_bytes(batch));
}

static PyObject *py_jacobi_lds_smem_bytes(PyObject *, PyObject *args) {
  int dtype, m, n, flags;
  if (!PyArg_ParseTuple(args, "iiii", &dtype, &m, &n, &flags)) return nullptr;
  size_t bytes = 0;
  int st = psvd_jacobi_lds_smem_bytes(dtype, m, n, flags, &bytes);
  return Py_BuildValue("(in)", st, (Py_ssize_t)bytes);
}

static PyObject *py_jacobi_lds(PyObject *, PyObject *args) {
  int dtype, mu, nu, max_m, max_n, max_sweeps, flags, lanes, threads;
  long long batch;
  unsigned long long dims, offs, order, a, u, s, vh, info, ws, ws_bytes, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiiKKKiiKKKKKdiiiiKKK", &dtype, &batch, &mu, &nu, &dims, &offs, &order,
                        &max_m, &max_n, &a, &u, &s, &vh, &info, &tol, &max_sweeps, &flags, &lanes,
                        &threads, &ws, &ws_bytes, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_jacobi_lds(dtype, batch, mu, nu, static_cast<const int32_t *>(PTR(dims)),
                       static_cast<const int64_t *>(PTR(offs)), static_cast<const int32_t *>(PTR(order)),
                       max_m, max_n, PTR(a), PTR(u), PTR(s), PTR(vh), static_cast<int32_t *>(PTR(info)),
                       tol, max_sweeps, flags, lanes, threads, PTR(ws), (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_jacobi_lds_noalloc(PyObject *, PyObject *args) {
  int dtype, mu, nu, max_m, max_n, max_sweeps, flags, lanes, threads;
  long long batch;
  unsigned long long dims, offs, order, a, u, s, vh, info, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiiKKKiiKKKKKdiiiiK", &dtype, &batch, &mu, &nu, &dims, &offs, &order, &max_m,
                        &max_n, &a, &u, &s, &vh, &info, &tol, &max_sweeps, &flags, &lanes, &threads, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_jacobi_lds_noalloc(dtype, batch, mu, nu, static_cast<const int32_t *>(PTR(dims)),
                               static_cast<const int64_t *>(PTR(offs)), static_cast<const int32_t *>(PTR(order)),
                               max_m, max_n, PTR(a), PTR(u), PTR(s), PTR(vh), static_cast<int32_t *>(PTR(info)), tol,
                               max_sweeps, flags, lanes, threads, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_jacobi_grid_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  if (!PyArg_ParseTuple(args, "iii", &dtype, &m, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_jacobi_grid_workspace_bytes(dtype, m, n));
}

static PyObject *py_jacobi_grid(PyObject *, PyObject *args) {
  int dtype, m, n, max_sweeps, flags;
  unsigned long long a, u, s, vh, info, ws, ws_bytes, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iiiKKKKKdiiKKK", &dtype, &m, &n, &a, &u, &s, &vh, &info, &tol, &max_sweeps,
                        &flags, &ws, &ws_bytes, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_jacobi_grid(dtype, m, n, PTR(a), PTR(u), PTR(s), PTR(vh), static_cast<int32_t *>(PTR(info)), tol,
                        max_sweeps, flags, PTR(ws), (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_jacobi_blocked_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  if (!PyArg_ParseTuple(args, "iii", &dtype, &m, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_jacobi_blocked_workspace_bytes(dtype, m, n));
}

static PyObject *py_jacobi_blocked(PyObject *, PyObject *args) {
  int dtype, m, n, max_sweeps, flags, inner;
  unsigned long long a, u, s, vh, info, ws, ws_bytes, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iiiKKKKKdiiiKKK", &dtype, &m, &n, &a, &u, &s, &vh, &info, &tol, &max_sweeps, &flags,
                        &inner, &ws, &ws_bytes, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_jacobi_blocked(dtype, m, n, PTR(a), PTR(u), PTR(s), PTR(vh), static_cast<int32_t *>(PTR(info)), tol,
                           max_sweeps, flags, inner, PTR(ws), (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_gemm(PyObject *, PyObject *args) {
  int mode, ta, tb;
  long long m, n, k, lda, sa, ldb, sb, ldc, sc, batch;
  unsigned long long a, b, c, stream;
  double alpha, beta;
  if (!PyArg_ParseTuple(args, "iiiLLLdKLLKLLdKLLLK", &mode, &ta, &tb, &m, &n, &k, &alpha, &a, &lda, &sa, &b, &ldb,
                        &sb, &beta, &c, &ldc, &sc, &batch, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_gemm(mode, ta, tb, m, n, k, alpha, PTR(a), lda, sa, PTR(b), ldb, sb, beta, PTR(c), ldc, sc, batch,
                 PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_qr_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  if (!PyArg_ParseTuple(args, "iLii", &dtype, &batch, &m, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_qr_workspace_bytes(dtype, batch, m, n));
}

static PyObject *py_qr(PyObject *, PyObject *args) {
  int dtype, gmode, m, n;
  long long batch, lda, sa;
  unsigned long long a, q, r, ws, ws_bytes, stream;
  if (!PyArg_ParseTuple(args, "iiLiiKLLKKKKK", &dtype, &gmode, &batch, &m, &n, &a, &lda, &sa, &q, &r, &ws,
                        &ws_bytes, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_qr(dtype, gmode, batch, m, n, PTR(a), lda, sa, PTR(q), PTR(r), PTR(ws), (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_reduce(PyObject *, PyObject *args) {
  int dtype, op, rows, cols;
  long long batch, ldx, sx, ldy, sy;
  unsigned long long x, y, out, stream;
  if (!PyArg_ParseTuple(args, "iiLiiKLLKLLKK", &dtype, &op, &batch, &rows, &cols, &x, &ldx, &sx, &y, &ldy, &sy, &out,
                        &stream))
    return nullptr;
  return PyLong_FromLong(psvd_reduce(dtype, op, batch, rows, cols, PTR(x), ldx, sx, PTR(y), ldy, sy,
                                     static_cast<double *>(PTR(out)), PTR(stream)));
}

static PyObject *py_scale(PyObject *, PyObject *args) {
  int dtype, rows, cols, invert;
  long long batch, ldx, sx, ldy, sy;
  unsigned long long x, y, d, stream;
  double s;
  if (!PyArg_ParseTuple(args, "iLiiKLLKLLKidK", &dtype, &batch, &rows, &cols, &x, &ldx, &sx, &y, &ldy, &sy, &d,
                        &invert, &s, &stream))
    return nullptr;
  return PyLong_FromLong(psvd_scale(dtype, batch, rows, cols, PTR(x), ldx, sx, PTR(y), ldy, sy,
                                    static_cast<const double *>(PTR(d)), invert, s, PTR(stream)));
}

static PyObject *py_symmetrize(PyObject *, PyObject *args) {
  int dtype, n;
  long long batch, lda, sa;
  unsigned long long a, stream;
  if (!PyArg_ParseTuple(args, "iLiKLLK", &dtype, &batch, &n, &a, &lda, &sa, &stream)) return nullptr;
  return PyLong_FromLong(psvd_symmetrize(dtype, batch, n, PTR(a), lda, sa, PTR(stream)));
}

static PyObject *py_eigh_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, n;
  if (!PyArg_ParseTuple(args, "ii", &dtype, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_eigh_workspace_bytes(dtype, n));
}

static PyObject *py_eigh(PyObject *, PyObject *args) {
  int dtype, n, max_sweeps;
  long long batch;
  unsigned long long h, w, v, info, ws, ws_bytes, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiKKKKdiKKK", &dtype, &batch, &n, &h, &w, &v, &info, &tol, &max_sweeps, &ws, &ws_bytes,
                        &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_eigh(dtype, batch, n, PTR(h), PTR(w), PTR(v), static_cast<int32_t *>(PTR(info)), tol, max_sweeps, PTR(ws),
                 (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_eigh_blocked_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, n;
  if (!PyArg_ParseTuple(args, "ii", &dtype, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_eigh_blocked_workspace_bytes(dtype, n));
}

static PyObject *py_eigh_blocked(PyObject *, PyObject *args) {
  int dtype, n, max_sweeps, inner;
  long long batch;
  unsigned long long h, w, v, info, ws, ws_bytes, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiKKKKdiiKKK", &dtype, &batch, &n, &h, &w, &v, &info, &tol, &max_sweeps, &inner, &ws,
                        &ws_bytes, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_eigh_blocked(dtype, batch, n, PTR(h), PTR(w), PTR(v), static_cast<int32_t *>(PTR(info)), tol, max_sweeps,
                         inner, PTR(ws), (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_cholesky(PyObject *, PyObject *args) {
  int dtype, gmode, n;
  long long batch, lda, sa;
  unsigned long long a, fail, stream;
  if (!PyArg_ParseTuple(args, "iiLiKLLKK", &dtype, &gmode, &batch, &n, &a, &lda, &sa, &fail, &stream)) return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_cholesky(dtype, gmode, batch, n, PTR(a), lda, sa, static_cast<int32_t *>(PTR(fail)), PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_trsm_right(PyObject *, PyObject *args) {
  int dtype, gmode, trans, n, rows;
  long long batch, ldl, sl, ldb, sb;
  unsigned long long l, b, stream;
  if (!PyArg_ParseTuple(args, "iiLiiiKLLKLLK", &dtype, &gmode, &batch, &trans, &n, &rows, &l, &ldl, &sl, &b, &ldb, &sb,
                        &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_trsm_right(dtype, gmode, batch, trans, n, rows, PTR(l), ldl, sl, PTR(b), ldb, sb, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_axpby(PyObject *, PyObject *args) {
  int dtype, rows, cols;
  long long batch, ldx, sx, ldy, sy;
  unsigned long long x, y, stream;
  double alpha, beta;
  if (!PyArg_ParseTuple(args, "iLiidKLLdKLLK", &dtype, &batch, &rows, &cols, &alpha, &x, &ldx, &sx, &beta, &y, &ldy,
                        &sy, &stream))
    return nullptr;
  return PyLong_FromLong(psvd_axpby(dtype, batch, rows, cols, alpha, PTR(x), ldx, sx, beta, PTR(y), ldy, sy, PTR(stream)));
}

static PyObject *py_sort_desc(PyObject *, PyObject *args) {
  int dtype, rows, n;
  long long batch, ldx, sx, ldy, sy;
  unsigned long long keys, s, x, y, perm, stream;
  if (!PyArg_ParseTuple(args, "iLiiKKKLLKLLKK", &dtype, &batch, &rows, &n, &keys, &s, &x, &ldx, &sx, &y, &ldy, &sy,
                        &perm, &stream))
    return nullptr;
  return PyLong_FromLong(psvd_sort_desc(dtype, batch, rows, n, static_cast<const double *>(PTR(keys)), PTR(s), PTR(x),
                                        ldx, sx, PTR(y), ldy, sy, static_cast<int32_t *>(PTR(perm)), PTR(stream)));
}

static PyObject *py_scale_columns(PyObject *, PyObject *args) {
  int dtype, rows, cols;
  long long batch, ldx, sx, sd;
  unsigned long long x, d, stream;
  if (!PyArg_ParseTuple(args, "iLiiKLLKLK", &dtype, &batch, &rows, &cols, &x, &ldx, &sx, &d, &sd, &stream))
    return nullptr;
  return PyLong_FromLong(psvd_scale_columns(dtype, batch, rows, cols, PTR(x), ldx, sx, PTR(d), sd, PTR(stream)));
}

static PyObject *py_cqr_lds_smem_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  if (!PyArg_ParseTuple(args, "iii", &dtype, &m, &n)) return nullptr;
  size_t bytes = 0;
  int st = psvd_cqr_lds_smem_bytes(dtype, m, n, &bytes);
  return Py_BuildValue("(in)", st, (Py_ssize_t)bytes);
}

static PyObject *py_cqr_lds(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  unsigned long long a, q, r, stream;
  if (!PyArg_ParseTuple(args, "iLiiKKKK", &dtype, &batch, &m, &n, &a, &q, &r, &stream)) return nullptr;
  return PyLong_FromLong(psvd_cqr_lds(dtype, batch, m, n, PTR(a), PTR(q), PTR(r), PTR(stream)));
}

static PyObject *py_cmatmul(PyObject *, PyObject *args) {
  int dtype, ta, tb, m, k, n;
  long long batch;
  unsigned long long a, b, c, stream;
  if (!PyArg_ParseTuple(args, "iiiLiiiKKKK", &dtype, &ta, &tb, &batch, &m, &k, &n, &a, &b, &c, &stream))
    return nullptr;
  return PyLong_FromLong(psvd_cmatmul(dtype, ta, tb, batch, m, k, n, PTR(a), PTR(b), PTR(c), PTR(stream)));
}

static PyObject *py_ccomplete_columns(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  unsigned long long u, nk, status, stream;
  if (!PyArg_ParseTuple(args, "iLiiKKKK", &dtype, &batch, &m, &n, &u, &nk, &status, &stream)) return nullptr;
  return PyLong_FromLong(psvd_ccomplete_columns(dtype, batch, m, n, PTR(u), static_cast<const int32_t *>(PTR(nk)),
                                                static_cast<int32_t *>(PTR(status)), PTR(stream)));
}

static PyObject *py_chermitize(PyObject *, PyObject *args) {
  int dtype, n;
  long long batch;
  unsigned long long a, stream;
  if (!PyArg_ParseTuple(args, "iLiKK", &dtype, &batch, &n, &a, &stream)) return nullptr;
  return PyLong_FromLong(psvd_chermitize(dtype, batch, n, PTR(a), PTR(stream)));
}

static PyObject *py_block_jacobi_workspace_bytes(PyObject *, PyObject *args) {
  long long batch;
  int n;
  if (!PyArg_ParseTuple(args, "Li", &batch, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_block_jacobi_workspace_bytes(batch, n));
}

static PyObject *py_block_jacobi_sweep(PyObject *, PyObject *args) {
  int dtype, m, n, width, max_sweeps;
  long long batch, ldw, sw, ldv, sv;
  double tol;
  unsigned long long w, v, info, corr, sigma, ws, wsb, stream;
  if (!PyArg_ParseTuple(args, "iLiiiKLLKLLdiKKKKKK", &dtype, &batch, &m, &n, &width, &w, &ldw, &sw, &v, &ldv, &sv,
                        &tol, &max_sweeps, &info, &corr, &sigma, &ws, &wsb, &stream))
    return nullptr;
  return PyLong_FromLong(psvd_block_jacobi_sweep(dtype, batch, m, n, width, PTR(w), ldw, sw, PTR(v), ldv, sv, tol,
                                                 max_sweeps, static_cast<int32_t *>(PTR(info)),
                                                 static_cast<double *>(PTR(corr)), static_cast<double *>(PTR(sigma)),
                                                 PTR(ws), (size_t)wsb, PTR(stream)));
}

static PyObject *py_mlp_param_count(PyObject *, PyObject *) { return PyLong_FromLong(psvd_mlp_param_count()); }

static PyObject *py_mlp_workspace_bytes(PyObject *, PyObject *args) {
  long long batch;
  if (!PyArg_ParseTuple(args, "L", &batch)) return nullptr;
  return PyLong_FromSize_t(psvd_mlp_workspace_bytes(batch));
}

static PyObject *py_mlp_forward(PyObject *, PyObject *args) {
  unsigned long long prm, x, y, stream;
  long long batch;
  if (!PyArg_ParseTuple(args, "KKLKK", &prm, &x, &batch, &y, &stream)) return nullptr;
  return PyLong_FromLong(psvd_mlp_forward(static_cast<const float *>(PTR(prm)), static_cast<const float *>(PTR(x)), batch,
                                          static_cast<float *>(PTR(y)), PTR(stream)));
}

static PyObject *py_mlp_grad(PyObject *, PyObject *args) {
  unsigned long long prm, x, t, y, g, loss, ws, ws_bytes, stream;
  long long batch;
  if (!PyArg_ParseTuple(args, "KKKLKKKKKK", &prm, &x, &t, &batch, &y, &g, &loss, &ws, &ws_bytes, &stream)) return nullptr;
  return PyLong_FromLong(psvd_mlp_grad(static_cast<const float *>(PTR(prm)), static_cast<const float *>(PTR(x)),
                                       static_cast<const float *>(PTR(t)), batch, static_cast<float *>(PTR(y)),
                                       static_cast<float *>(PTR(g)), static_cast<double *>(PTR(loss)), PTR(ws),
                                       (size_t)ws_bytes, PTR(stream)));
}

static PyObject *py_adam_step(PyObject *, PyObject *args) {
  unsigned long long prm, g, m, v, stream;
  int n, step;
  float lr, b1, b2, eps;
  if (!PyArg_ParseTuple(args, "KKKKiffffiK", &prm, &g, &m, &v, &n, &lr, &b1, &b2, &eps, &step, &stream)) return nullptr;
  return PyLong_FromLong(psvd_adam_step(static_cast<float *>(PTR(prm)), static_cast<const float *>(PTR(g)),
                                        static_cast<float *>(PTR(m)), static_cast<float *>(PTR(v)), n, lr, b1, b2, eps,
                                        step, PTR(stream)));
}

static PyObject *py_colstats(PyObject *, PyObject *args) {
  int dtype, rows, cols;
  long long batch, lda, sa;
  unsigned long long a, out, stream;
  if (!PyArg_ParseTuple(args, "iLiiKLLKK", &dtype, &batch, &rows, &cols, &a, &lda, &sa, &out, &stream)) return nullptr;
  return PyLong_FromLong(
      psvd_colstats(dtype, batch, rows, cols, PTR(a), lda, sa, static_cast<float *>(PTR(out)), PTR(stream)));
}

static PyObject *py_complete_workspace_bytes(PyObject *, PyObject *args) {
  long long batch;
  int m, n;
  if (!PyArg_ParseTuple(args, "Lii", &batch, &m, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_complete_workspace_bytes(batch, m, n));
}

static PyObject *py_complete_columns(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch, ldu, su;
  unsigned long long u, nk, status, ws, ws_bytes, stream;
  if (!PyArg_ParseTuple(args, "iLiiKLLKKKKK", &dtype, &batch, &m, &n, &u, &ldu, &su, &nk, &status, &ws, &ws_bytes,
                        &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_complete_columns(dtype, batch, m, n, PTR(u), ldu, su, static_cast<const int32_t *>(PTR(nk)),
                             static_cast<int32_t *>(PTR(status)), PTR(ws), (size_t)ws_bytes, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_ceigh_lds_smem_bytes(PyObject *, PyObject *args) {
  int dtype, n;
  if (!PyArg_ParseTuple(args, "ii", &dtype, &n)) return nullptr;
  size_t bytes = 0;
  int st = psvd_ceigh_lds_smem_bytes(dtype, n, &bytes);
  return Py_BuildValue("(in)", st, (Py_ssize_t)bytes);
}

static PyObject *py_ceigh_lds(PyObject *, PyObject *args) {
  int dtype, n, max_sweeps;
  long long batch;
  unsigned long long h, w, v, info, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiKKKKdiK", &dtype, &batch, &n, &h, &w, &v, &info, &tol, &max_sweeps, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_ceigh_lds(dtype, batch, n, PTR(h), PTR(w), PTR(v), static_cast<int32_t *>(PTR(info)), tol, max_sweeps,
                      PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_ccholesky_lds(PyObject *, PyObject *args) {
  int dtype, n;
  long long batch;
  unsigned long long g, l, fail, stream;
  if (!PyArg_ParseTuple(args, "iLiKKKK", &dtype, &batch, &n, &g, &l, &fail, &stream)) return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_ccholesky_lds(dtype, batch, n, PTR(g), PTR(l), static_cast<int32_t *>(PTR(fail)), PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_csolve_lds_smem_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  if (!PyArg_ParseTuple(args, "iii", &dtype, &m, &n)) return nullptr;
  size_t bytes = 0;
  int st = psvd_csolve_lds_smem_bytes(dtype, m, n, &bytes);
  return Py_BuildValue("(in)", st, (Py_ssize_t)bytes);
}

static PyObject *py_csolve_chol_lds(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  unsigned long long l, x, stream;
  if (!PyArg_ParseTuple(args, "iLiiKKK", &dtype, &batch, &m, &n, &l, &x, &stream)) return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_csolve_chol_lds(dtype, batch, m, n, PTR(l), PTR(x), PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_cqr_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  if (!PyArg_ParseTuple(args, "iLii", &dtype, &batch, &m, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_cqr_workspace_bytes(dtype, batch, m, n));
}

static PyObject *py_cqr(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  unsigned long long a, q, r, ws, wsb, stream;
  if (!PyArg_ParseTuple(args, "iLiiKKKKKK", &dtype, &batch, &m, &n, &a, &q, &r, &ws, &wsb, &stream)) return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_cqr(dtype, batch, m, n, PTR(a), PTR(q), PTR(r), PTR(ws), (size_t)wsb, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_ceigh_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, n;
  long long batch;
  if (!PyArg_ParseTuple(args, "iLi", &dtype, &batch, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_ceigh_workspace_bytes(dtype, batch, n));
}

static PyObject *py_ceigh(PyObject *, PyObject *args) {
  int dtype, n, max_sweeps;
  long long batch;
  unsigned long long h, w, v, info, ws, wsb, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiKKKKdiKKK", &dtype, &batch, &n, &h, &w, &v, &info, &tol, &max_sweeps, &ws, &wsb,
                        &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_ceigh(dtype, batch, n, PTR(h), PTR(w), PTR(v), static_cast<int32_t *>(PTR(info)), tol, max_sweeps,
                  PTR(ws), (size_t)wsb, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_ccholesky_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, n;
  long long batch;
  if (!PyArg_ParseTuple(args, "iLi", &dtype, &batch, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_ccholesky_workspace_bytes(dtype, batch, n));
}

static PyObject *py_ccholesky(PyObject *, PyObject *args) {
  int dtype, n;
  long long batch;
  unsigned long long g, l, fail, ws, wsb, stream;
  if (!PyArg_ParseTuple(args, "iLiKKKKKK", &dtype, &batch, &n, &g, &l, &fail, &ws, &wsb, &stream)) return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_ccholesky(dtype, batch, n, PTR(g), PTR(l), static_cast<int32_t *>(PTR(fail)), PTR(ws), (size_t)wsb,
                      PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_csolve_chol_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  if (!PyArg_ParseTuple(args, "iLii", &dtype, &batch, &m, &n)) return nullptr;
  return PyLong_FromSize_t(psvd_csolve_chol_workspace_bytes(dtype, batch, m, n));
}

static PyObject *py_csolve_chol(PyObject *, PyObject *args) {
  int dtype, m, n;
  long long batch;
  unsigned long long l, x, ws, wsb, stream;
  if (!PyArg_ParseTuple(args, "iLiiKKKKK", &dtype, &batch, &m, &n, &l, &x, &ws, &wsb, &stream)) return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_csolve_chol(dtype, batch, m, n, PTR(l), PTR(x), PTR(ws), (size_t)wsb, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyObject *py_cjacobi_batched_workspace_bytes(PyObject *, PyObject *args) {
  int dtype, m, n, flags;
  long long batch;
  if (!PyArg_ParseTuple(args, "iLiii", &dtype, &batch, &m, &n, &flags)) return nullptr;
  return PyLong_FromSize_t(psvd_cjacobi_batched_workspace_bytes(dtype, batch, m, n, flags));
}

static PyObject *py_cjacobi_batched(PyObject *, PyObject *args) {
  int dtype, m, n, max_sweeps, flags;
  long long batch;
  unsigned long long a, u, s, vh, info, ws, wsb, stream;
  double tol;
  if (!PyArg_ParseTuple(args, "iLiiKKKKKdiiKKK", &dtype, &batch, &m, &n, &a, &u, &s, &vh, &info, &tol, &max_sweeps,
                        &flags, &ws, &wsb, &stream))
    return nullptr;
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = psvd_cjacobi_batched(dtype, batch, m, n, PTR(a), PTR(u), PTR(s), PTR(vh), static_cast<int32_t *>(PTR(info)),
                            tol, max_sweeps, flags, PTR(ws), (size_t)wsb, PTR(stream));
  Py_END_ALLOW_THREADS
  return PyLong_FromLong(st);
}

static PyMethodDef kMethods[] = {
    {"version", py_version, METH_NOARGS, "library version string"},
    {"jacobi_lds_workspace_bytes", py_jacobi_lds_workspace_bytes, METH_VARARGS, nullptr},
    {"jacobi_lds_smem_bytes", py_jacobi_lds_smem_bytes, METH_VARARGS, nullptr},
    {"jacobi_lds", py_jacobi_lds, METH_VARARGS, nullptr},
    {"jacobi_lds_noalloc", py_jacobi_lds_noalloc, METH_VARARGS, nullptr},
    {"jacobi_grid_workspace_bytes", py_jacobi_grid_workspace_bytes, METH_VARARGS, nullptr},
    {"jacobi_grid", py_jacobi_grid, METH_VARARGS, nullptr},
    {"jacobi_blocked_workspace_bytes", py_jacobi_blocked_workspace_bytes, METH_VARARGS, nullptr},
    {"jacobi_blocked", py_jacobi_blocked, METH_VARARGS, nullptr},
    {"gemm", py_gemm, METH_VARARGS, nullptr},
    {"eigh_workspace_bytes", py_eigh_workspace_bytes, METH_VARARGS, nullptr},
    {"eigh", py_eigh, METH_VARARGS, nullptr},
    {"eigh_blocked_workspace_bytes", py_eigh_blocked_workspace_bytes, METH_VARARGS, nullptr},
    {"eigh_blocked", py_eigh_blocked, METH_VARARGS, nullptr},
    {"cholesky", py_cholesky, METH_VARARGS, nullptr},
    {"trsm_right", py_trsm_right, METH_VARARGS, nullptr},
    {"reduce", py_reduce, METH_VARARGS, nullptr},
    {"scale", py_scale, METH_VARARGS, nullptr},
    {"symmetrize", py_symmetrize, METH_VARARGS, nullptr},
    {"axpby", py_axpby, METH_VARARGS, nullptr},
    {"sort_desc", py_sort_desc, METH_VARARGS, nullptr},
    {"scale_columns", py_scale_columns, METH_VARARGS, nullptr},
    {"cqr_lds_smem_bytes", py_cqr_lds_smem_bytes, METH_VARARGS, nullptr},
    {"cqr_lds", py_cqr_lds, METH_VARARGS, nullptr},
    {"cjacobi_batched_workspace_bytes", py_cjacobi_batched_workspace_bytes, METH_VARARGS, nullptr},
    {"cjacobi_batched", py_cjacobi_batched, METH_VARARGS, nullptr},
    {"cqr_workspace_bytes", py_cqr_workspace_bytes, METH_VARARGS, nullptr},
    {"cqr", py_cqr, METH_VARARGS, nullptr},
    {"ceigh_workspace_bytes", py_ceigh_workspace_bytes, METH_VARARGS, nullptr},
    {"ceigh", py_ceigh, METH_VARARGS, nullptr},
    {"ccholesky_workspace_bytes", py_ccholesky_workspace_bytes, METH_VARARGS, nullptr},
    {"ccholesky", py_ccholesky, METH_VARARGS, nullptr},
    {"csolve_chol_workspace_bytes", py_csolve_chol_workspace_bytes, METH_VARARGS, nullptr},
    {"csolve_chol", py_csolve_chol, METH_VARARGS, nullptr},
    {"cmatmul", py_cmatmul, METH_VARARGS, nullptr},
    {"chermitize", py_chermitize, METH_VARARGS, nullptr},
    {"ccomplete_columns", py_ccomplete_columns, METH_VARARGS, nullptr},
    {"block_jacobi_workspace_bytes", py_block_jacobi_workspace_bytes, METH_VARARGS, nullptr},
    {"block_jacobi_sweep", py_block_jacobi_sweep, METH_VARARGS, nullptr},
    {"mlp_param_count", py_mlp_param_count, METH_NOARGS, nullptr},
    {"mlp_workspace_bytes", py_mlp_workspace_bytes, METH_VARARGS, nullptr},
    {"mlp_forward", py_mlp_forward, METH_VARARGS, nullptr},
    {"mlp_grad", py_mlp_grad, METH_VARARGS, nullptr},
    {"adam_step", py_adam_step, METH_VARARGS, nullptr},
    {"colstats", py_colstats, METH_VARARGS, nullptr},
    {"complete_workspace_bytes", py_complete_workspace_bytes, METH_VARARGS, nullptr},
    {"complete_columns", py_complete_columns, METH_VARARGS, nullptr},
    {"qr_workspace_bytes", py_qr_workspace_bytes, METH_VARARGS, nullptr},
    {"qr", py_qr, METH_VARARGS, nullptr},
    {"ceigh_lds_smem_bytes", py_ceigh_lds_smem_bytes, METH_VARARGS, nullptr},
    {"ceigh_lds", py_ceigh_lds, METH_VARARGS, nullptr},
    {"ccholesky_lds", py_ccholesky_lds, METH_VARARGS, nullptr},
    {"csolve_lds_smem_bytes", py_csolve_lds_smem_bytes, METH_VARARGS, nullptr},
    {"csolve_chol_lds", py_csolve_chol_lds, METH_VARARGS, nullptr},
    {nullptr, nullptr, 0, nullptr}};

static struct PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "_hip", "gfx950 SVD kernels (C ABI binding)",
                                     -1, kMethods};

PyMODINIT_FUNC PyInit__hip(void) { return PyModule_Create(&kModule); }
