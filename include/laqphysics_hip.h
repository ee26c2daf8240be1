/*
 * laqphysics_hip -- MI355X (gfx950) propagation engine for the 3D helium TDSE.
 *
 * C ABI.  Plain pointers and sizes only.  Complex numbers are complex128 stored
 * as interleaved (re, im) doubles, i.e. C99 `double _Complex` / numpy
 * complex128 / torch.complex128.  Wavefunctions are F-ordered psi(i, j, c):
 * flat index i + n*j + n*n*c (i = electron-1 radial index, fastest).
 *
 * Every entry point returns 0 on success or a negative code; hex_last_error()
 * returns the message of the last failure on the calling thread.
 *
 * Codes:  HEX_EINVAL  bad argument (the Python layer raises ValueError)
 *         HEX_EDEVICE HIP runtime failure (RuntimeError)
 *         HEX_ENOTCONV Krylov / GMRES non-convergence (RuntimeError whose message
 *                      contains "did not converge", which callers match on)
 */
#ifndef LAQPHYSICS_HIP_H
#define LAQPHYSICS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEX_OK        0
#define HEX_EINVAL   -1
#define HEX_EDEVICE  -2
#define HEX_ENOTCONV -3

const char *hex_last_error(void);
int hex_version(void);

/* ---------------------------------------------------------------------------
 * Drop-in replacement for the reference's native H.psi kernel.
 * Replaces: src/kernels.f90:119-176 `tensor_apply` (bind(C)), as bound by
 * python/native.py:24-38.  Same argument list and meaning: host pointers,
 * 0-based CSR/edge indices, dcoef F-ordered (3, nd), psi/out F-ordered
 * (n, n, nc).  The apply runs on the current HIP device; buffers are staged
 * through device memory (the reference kernel has no return value; this one
 * returns a status code, which callers that ignore it may drop).
 * ------------------------------------------------------------------------- */
int tensor_apply(int n, int nc, int nt,
                 const int32_t *ptr, const int32_t *col, const double *tv,
                 const double *diag,
                 int nv, const int32_t *vptr, const int32_t *vc, const int32_t *vl,
                 const double *vcoef, const double *rad,
                 int nd, const int32_t *dptr, const int32_t *dc, const int32_t *de,
                 const double *dcoef, const double *ldiff,
                 const double *r, const double *field, int velocity,
                 const double *psi, double *out);

/* ---------------------------------------------------------------------------
 * Device-resident operator ("engine").
 * Replaces: python/torch_backend.py:11-75 `TorchHamiltonian.__init__` (host ->
 * device copy of all operator tables, once) and :109-131 `.apply`.
 *
 * Tables are the ones python/helium3d.py:13-50 (and the coupled/bipolar
 * variants) build, except that the diagonal is passed in its separable form:
 * diag(i, j, c) = d1[i + n*c] + d2[j + n*c] (bit-identical to the reference's
 * u[:,None] + v[None,:]), and the multipoles rad(i, j, lam) =
 * r_<^lam / r_>^(lam+1) * cut_i * cut_j are regenerated on the fly from r and
 * cut (python/helium3d.py:29-30) instead of being streamed.
 * ------------------------------------------------------------------------- */
typedef struct hex_tables {
    int n, nc, lam_count;
    int nt;              const int32_t *ptr, *col; const double *tv;   /* complex[nt] */
    const double *r;     /* complex[n] */
    const double *cut;   /* real[n]    */
    const double *d1, *d2;                                              /* complex[n*nc] */
    int nv;  const int32_t *vptr, *vc, *vl; const double *vcoef;        /* complex[nv] */
    int nd;  const int32_t *dptr, *dc, *de; const double *dcoef;        /* complex[3*nd], F (3, nd) */
    const double *ldiff; /* real[nd] */
} hex_tables;

typedef struct hex_engine hex_engine;

int hex_engine_create(const hex_tables *tables, int device, hex_engine **engine);
int hex_engine_destroy(hex_engine *engine);
/* A table-free handle that only owns reduction scratch, for the Krylov entry
 * points when no operator is involved (destroy with hex_engine_destroy). */
int hex_workspace_create(int device, hex_engine **engine);
/* Device bytes held by the engine's tables and scratch. */
int64_t hex_engine_bytes(const hex_engine *engine);

/* out = scale * H(field) psi  (device pointers, N = n*n*nc complex each).
 * velocity != 0: velocity gauge with the A^2 term (kernels.f90:153-170);
 * velocity == 0: length gauge (kernels.f90:147-152).
 * stream: a hipStream_t (NULL = default stream). */
int hex_apply(hex_engine *engine, const double field[3], int velocity, double scale,
              const double *psi, double *out, void *stream);

/* ---------------------------------------------------------------------------
 * Krylov building blocks (wavefront reductions, deterministic order).
 * Replace the torch calls of python/torch_backend.py:141-158 and
 * python/implicit.py:46-74: `Q[:, :j+1].mH @ v`, `v -= Q[:, :j+1] @ z` (CGS2),
 * `torch.linalg.vector_norm`, `Q[:, :m] @ c`.
 * A basis is `k` vectors of N complex each, vector m at Q + 2*m*ldq doubles.
 * ------------------------------------------------------------------------- */

/* Two-pass classical Gram-Schmidt of v against Q[0..k-1], in place, then
 * ||v||.  hcol[0..k-1] += z1 + z2 (the Hessenberg column) and
 * hcol[k] = ||v|| are written to the host array `hcol` (complex[k+1]); the call
 * synchronises the stream.  Same algebra and order as propagate.py:17-25. */
int hex_cgs2(hex_engine *engine, const double *Q, int64_t ldq, int k, double *v, int64_t N,
             double *hcol, void *stream);

/* ||x||_2 returned in *norm (host); synchronises. */
int hex_norm(hex_engine *engine, const double *x, int64_t N, double *norm, void *stream);

/* y = x / s  (s real) */
int hex_scale(const double *x, double s, double *y, int64_t N, void *stream);

/* y = sum_m Q_m c[m], c host complex[k] */
int hex_combine(hex_engine *engine, const double *Q, int64_t ldq, int k, const double *c,
                double *y, int64_t N, void *stream);

/* z[m] = <Q_m, v> (host complex[k]); synchronises. */
int hex_dots(hex_engine *engine, const double *Q, int64_t ldq, int k, const double *v, int64_t N,
             double *z, void *stream);

#ifdef __cplusplus
}
#endif
#endif
