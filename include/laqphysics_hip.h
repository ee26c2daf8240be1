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
 *
 * Concurrency: an engine (and a separable preconditioner, and a batch) owns
 * scratch buffers shared by all its calls.  Calls on one object are ordered
 * by the stream they are given; calls on one object from different streams
 * must not overlap in time.  Different objects are independent.
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

/* out = H(field)^dagger psi: the adjoint apply of the costate propagation.
 * No reference counterpart (the reference has no adjoint); H^dagger is H with
 * conjugated r, T, diagonal and multipoles (the reflected irECS contour). */
int hex_apply_adjoint(hex_engine *engine, const double field[3], int velocity,
                      const double *psi, double *out, void *stream);

/* out = a * H_terms(field) psi + b * psi with a selection of H's terms:
 * bit 0 diagonal + kinetic + Coulomb, bit 1 field-coupling (dipole) edges,
 * bit 2 |A|^2; adjoint != 0 applies the adjoint of that selection.  With
 * terms = 2 and field = e_z this is dH/dA_z minus the 2 A_z identity part,
 * which the adjoint gradient needs.  No reference counterpart. */
int hex_apply_terms(hex_engine *engine, const double field[3], int velocity, int adjoint, int terms,
                    const double a[2], const double b[2], const double *psi, double *out, void *stream);

/* out_b = H(fields_b) psi_b for `members` states stored back to back
 * (members x N complex each, device), fields a device array [members][3]:
 * the batched pattern's apply with B right-hand sides, outside the
 * propagation graph.  No reference counterpart (the reference applies one
 * state at a time).  Scratch for the members is allocated on first use. */
int hex_apply_batched(hex_engine *engine, int members, const double *fields, int velocity, const double *psi,
                      double *out, void *stream);

/* Launch shape of the radial pass of the apply: 0 = staged walks (64 x 16
 * tiles through LDS, the default), 1, 2, 4, 8 = direct walks with that many
 * rows j per workgroup; chosen by the autotuner.  No reference counterpart. */
int hex_engine_set_launch(hex_engine *engine, int rows_per_block);

/* out = a * H(field) psi + b * psi with complex a, b (interleaved pairs).
 * The CF4-Pade systems (I + i alpha (H - E_ref)) of python/implicit.py:78-84
 * are one call with a = i alpha, b = 1 - i alpha E_ref. */
int hex_apply_affine(hex_engine *engine, const double field[3], int velocity, const double a[2],
                     const double b[2], const double *psi, double *out, void *stream);

/* ---------------------------------------------------------------------------
 * Mixed length/velocity gauge.
 * Replaces: python/mixed_gauge.py:73-117 `MixedHamiltonian` (__init__, apply).
 * hex_engine_create_mixed builds an engine from the same tables whose
 * field-coupling radial walks use q = r (1 - f) (complex[n]) in place of r;
 * hex_engine_set_mixed gives it the gauge's radial tables -- hq = q / (2 r^2),
 * F = r f (complex[n]), c0 = (1 - f)^2 (real[n]), c2 = q'^2 - (1 - f)^2
 * (complex[n]) -- and the angular-square edges: for output channel c the
 * entries sptr[c] .. sptr[c+1] with source channel ssrc, electron sele (1, 2)
 * and six coefficients scoef[6 k .. 6 k + 5] (complex) of rhat_a rhat_b for
 * the pairs xx, yy, zz, xy, xz, yz.  hex_apply_mixed computes
 * out = a H(c) psi + b psi with c = [A, E, A_a A_b (off-diagonal doubled)]
 * (python/mixed_gauge.py:29-32). */
int hex_engine_create_mixed(const hex_tables *tables, const double *q, int device, hex_engine **engine);
int hex_engine_set_mixed(hex_engine *engine, const double *hq, const double *F, const double *c0, const double *c2,
                         int nsq, const int32_t *sptr, const int32_t *ssrc, const int32_t *sele, const double *scoef);
int hex_apply_mixed(hex_engine *engine, const double coeff[12], const double a[2], const double b[2],
                    const double *psi, double *out, void *stream);

/* ---------------------------------------------------------------------------
 * Separable ionic-resolvent preconditioner (device; complex128, or complex64 below).
 * Replaces: python/implicit.py:13-39 `SeparablePreconditioner.apply`:
 *   out_c = S_{l1} [ (S_{l1}^-1 Psi_c S_{l2}^-T) / (1 + i alpha (E_{l1,i} + E_{l2,j} + shift)) ] S_{l2}^T
 * for every channel c with one-electron labels (l1[c], l2[c]); Psi_c is the
 * F-ordered n x n slab of channel c.  S, Sinv: complex[nl][n*n] column-major;
 * E: complex[nl][n].  apply is asynchronous on `stream`.
 * ------------------------------------------------------------------------- */
typedef struct hex_separable hex_separable;
int hex_separable_create(int device, int n, int nc, int nl, const double *S, const double *Sinv,
                         const double *E, const int32_t *l1, const int32_t *l2, hex_separable **out);
int hex_separable_apply(hex_separable *p, double alpha_re, double alpha_im, double shift,
                        const double *x, double *out, void *stream);
int64_t hex_separable_bytes(const hex_separable *p);
int hex_separable_destroy(hex_separable *p);
/* Number of streams the GEMMs of one apply are spread over (1 or 2), chosen
 * at creation by timing; the extra streams are forked from and joined back to
 * the caller's stream.  set_lanes overrides the choice; the result is the
 * same, bit for bit.  No reference counterpart. */
int hex_separable_lanes(const hex_separable *p);
int hex_separable_set_lanes(hex_separable *p, int lanes);
/* The complex64 form: python/implicit.py:13-39 with precision 'complex64' (the
 * reference's production setting, which its GMRES pairs with the flexible
 * variant, python/implicit.py:45-74).  Same arguments as hex_separable_create;
 * S and Sinv are rounded to complex64 copies.  apply (hex_separable_apply,
 * complex128 in and out) narrows x to complex64, runs the four transforms as
 * complex64 GEMMs and the permutations on complex64 slabs, forms the
 * resolvent's denominator in double and widens the result.
 * hex_separable_single: 1 for this form, 0 for complex128, -1 for NULL. */
int hex_separable_create_c64(int device, int n, int nc, int nl, const double *S, const double *Sinv,
                             const double *E, const int32_t *l1, const int32_t *l2, hex_separable **out);
int hex_separable_single(const hex_separable *p);

/* ---------------------------------------------------------------------------
 * He+ ionic dual propagation (device, complex128, nrhs dual columns).
 * Replaces: python/ionic_adjoint.py:12-52 (`IonicAdjoint`: its CSR operator
 * values = [1, *field] @ data, `apply`, `precondition`), which
 * python/streaming_surface.py:125-146 (`AdjointStepper`, ionic_device
 * cpu/cuda:0) drives through CF4-Pade GMRES steps.
 *   create:  ndof = na * n rows, CSR pattern (ptr[ndof+1], col[nnz]) shared by
 *            nk + 1 complex value arrays data[nk+1][nnz]: the conjugate
 *            transposes of the field-free block Hamiltonian and of one matrix
 *            per field coefficient (nk <= 16).
 *   apply:   y = a G(coef) x + b x, G(coef) = G_0 + sum_k coef[k] G_{k+1}
 *            (coef real[nk]); x, y row-major (ndof, nrhs), y must not alias x.
 *   set_eigen / precondition: per state a, S_a, Sinv_a (complex [na][n][n],
 *            row-major) and E_a (complex [na][n]) of the field-free block;
 *            y_a = S_a diag(1 / (1 + i alpha (E_a + shift))) Sinv_a x_a.
 * One engine's calls share its work buffer: calls on different streams race.
 * ------------------------------------------------------------------------- */
typedef struct hex_ionic hex_ionic;
int hex_ionic_create(int device, int64_t ndof, int nrhs, int nnz, const int32_t *ptr, const int32_t *col, int nk,
                     const double *data, hex_ionic **out);
int hex_ionic_set_eigen(hex_ionic *p, int n, int na, const double *S, const double *Sinv, const double *E);
int hex_ionic_apply(hex_ionic *p, const double *coef, const double a[2], const double b[2], const double *x,
                    double *y, void *stream);
int hex_ionic_precondition(hex_ionic *p, double alpha_re, double alpha_im, double shift, const double *x,
                           double *y, void *stream);
int64_t hex_ionic_bytes(const hex_ionic *p);
int hex_ionic_destroy(hex_ionic *p);

/* ---------------------------------------------------------------------------
 * tSURFF surface source on the device.
 * Replaces: python/helium3d.py:52-77 (`prepare_surface` rows, `flux`) and
 * python/torch_backend.py:133-138 (`TorchHamiltonian.flux`).
 * set_surface: J, JP are the surface rows of [T, Theta] and i[T, Theta] o
 * (r_k - r_s), row-major complex[ns][kc], restricted to the column window
 * [k0, k0 + kc) outside which both are zero.
 * flux: out[c][s][i] (complex, nc*ns*n) = sum_k J[s,k] psi(i,k,c) + the
 * electron-2 dipole edges' coef * sum_k JP[s,k] psi(i,k,dc) (field = A(t)).
 * ------------------------------------------------------------------------- */
int hex_engine_set_surface(hex_engine *engine, int ns, int k0, int kc, const double *J, const double *JP);
int hex_flux(hex_engine *engine, const double field[3], const double *psi, double *out, void *stream);
/* hex_flux restricted to some of its terms: bit 1 the J rows, bit 2 the
 * field-coupling rows (terms = 3 is hex_flux).  The adjoint gradient records
 * the J part and the unit-field part of every frame separately. */
int hex_flux_terms(hex_engine *engine, const double field[3], int terms, const double *psi, double *out,
                   void *stream);
/* Adjoint of the frame map psi -> (F0, Fd), F0 = the J part, Fd = the
 * field-coupling part for the unit field (0, 0, 1): u (state layout, n*n*nc)
 * = J^H g0 + (unit-field rows)^H gd, with g0, gd complex[nc][ns][n] as
 * hex_flux writes frames.  It turns dY/dF^* of a frame-dependent objective
 * into a costate source.  No counterpart in the reference, which has no
 * gradients. */
int hex_flux_adjoint(hex_engine *engine, const double *g0, const double *gd, double *u, void *stream);

/* ---------------------------------------------------------------------------
 * Batched propagation (pulse scans).  No reference counterpart: the reference
 * runs one wavefunction per process.  B members share one engine; member b
 * starts at t0[b], takes counts[b] steps of dt, and is driven by `npulses`
 * sin^2 pulses given as pulses[b][k][6] = {omega, cycles, field, start, cep,
 * polarisation (0 z, 1 sigma+, 2 sigma-)} (python/pulses.py:5-48,
 * python/run3d.py:41-56), evaluated on the device.  Every step is the
 * reference's midpoint Arnoldi step with recursive dt/2, tol/2 subdivision
 * (python/torch_backend.py:140-167), all decisions on the device.  One replay of
 * the recorded hipGraph is one Krylov attempt for every live member.
 * psi: caller-owned device array [B][N] updated in place; it must outlive the
 * batch.  maxdim in [4, 16].  run: replays `attempts` graphs on `stream`.
 * status: host int32[10][B] = alive, active, converged in the last attempt,
 * its Krylov dimension, failures (depth > 8: the member stops), completed
 * steps, split in the last attempt, pending intervals, attempts, accepted
 * substeps.
 * ------------------------------------------------------------------------- */
typedef struct hex_batch hex_batch;
int hex_batch_create(hex_engine *engine, int B, int maxdim, double dt, double tol, int npulses,
                     const double *pulses, const double *t0, const int32_t *counts, double *psi, hex_batch **out);
int hex_batch_set_launch(hex_batch *batch, int rows_per_block);
int hex_batch_run(hex_batch *batch, int attempts, void *stream);
int hex_batch_status(hex_batch *batch, int32_t *out, void *stream);
int64_t hex_batch_bytes(const hex_batch *batch);
int hex_batch_destroy(hex_batch *batch);

/* ---------------------------------------------------------------------------
 * tSURFF channel projection and Volkov accumulation (device pointers).
 * Replaces: python/surface_contract.py:41-56 (`SurfaceContractor.__call__`) and
 * python/streaming_surface.py:54-64, 97-109 (`VolkovAccumulator.add`).
 * contract: q[j][o][s] = sum over entries k in [eptr[j*nout+o], eptr[j*nout+o+1])
 *   of ew[k] * sum_r conj(chi[(ea[k]*n + r)*nu + eu[k]]) * frame[(ec[k]*ns + s)*n + r]
 *   (frame = hex_flux output layout [c][s][r]).
 * volkov_add: amplitude[j][e][d] += dt (previous + current) / 2 with
 *   current = i sqrt2 sum_l (sum_s q[j][l][s] tests[l][e][s]) harm[l][d]
 *             * exp(i (energy[e] time + k[e] (unit[d] . intA) + intA2 / 2));
 *   first != 0 only stores `previous`.  unit is real [nd][3].
 * ------------------------------------------------------------------------- */
int hex_surface_contract(int n, int ns, int nu, int nlab, int nout, const int32_t *eptr, const int32_t *ec,
                         const int32_t *ea, const int32_t *eu, const double *ew, const double *chi,
                         const double *frame, double *q, void *stream);
int hex_volkov_add(int nlab, int nout, int ns, int nE, int nd, const double *q, const double *tests,
                   const double *harm, const double *energy, const double *k, const double *unit,
                   double time, const double intA[3], double intA2, double dt, int first,
                   double *previous, double *amplitude, void *stream);

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

/* The same CGS2 as hex_cgs2 without a host round trip: the Hessenberg column
 * (k + 1 complex entries, the last one ||v||) goes to device memory hdev, and
 * qnext = v / ||v|| (v unscaled when ||v|| < 1e-14).  Lets GMRES read its
 * columns only at its least-squares check points (python/implicit.py:55-72 keeps
 * the same control flow).  qnext must not alias v. */
int hex_cgs2_commit(hex_engine *engine, const double *Q, int64_t ldq, int k, double *v, int64_t N, double *hdev,
                    double *qnext, void *stream);

/* hex_cgs2_commit for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * No counterpart in the reference.  Same kernels, same results bit for bit,
 * but no engine argument: its reduction scratch (1 MiB + 2 KiB: per-block
 * partial sums and the two coefficient vectors) is a fixed device array of the
 * library, resolved once per device and held by that GPU until the process
 * ends; no call allocates or frees.  Limit: k <= 63 basis vectors (HEX_EINVAL
 * beyond); any N.  Calls that overlap on one GPU share the scratch and race:
 * issue them on one stream. */
int hex_cgs2_commit_noalloc(const double *Q, int64_t ldq, int k, double *v, int64_t N, double *hdev,
                            double *qnext, void *stream);

/* Krylov control on the device.  hcols holds the Hessenberg columns written by
 * hex_cgs2_commit (column jj: h[0..jj] then ||v|| at jj + 1, 64 complex per
 * column); columns first..last are gathered into H (device, row-major,
 * (order + 1) x order) and the reference's per-column tests replayed in one
 * wave.  result (device, complex[2 + order]): [0] = (status, m), [1] =
 * (estimate, h_{m,m-1}) of the last test made, [2..2+m) the coefficients;
 * status 1 means the reference would stop (Arnoldi) or form a candidate
 * (GMRES) after column m - 1.
 *
 * hex_arnoldi_check replaces the host expm of python/propagate.py:28-35 and
 * python/torch_backend.py:150-156: breakdown below 1e-13, checks at j >= 3
 * with (j + 1) % 4 == 0 or j + 1 == maxdim, c = expm(-i dt H_m)[:, 0] beta by
 * scaling and squaring, accept when |dt h c_{m-1}| < tol; work (device,
 * complex[8 maxdim^2]) is needed for maxdim > 22.
 * hex_gmres_check replaces the host lstsq of python/implicit.py:62-68: at a
 * breakdown (< 1e-14), every cadence-th column and the restart length it
 * solves min |beta e_1 - H_m y| by Givens rotations; status 1 when the
 * estimate is below tol_b (= tol |b|), at a breakdown or at the restart. */
int hex_arnoldi_check(const double *hcols, int first, int last, int maxdim, double dt, double beta, double tol,
                      double *H, double *work, double *result, void *stream);
/* The same for a step replayed at a forced Krylov dimension `fixed` (the
 * adjoint's finite-difference runs reuse the dimensions of the run they
 * perturb): besides a breakdown the only test is after column fixed - 1, and
 * it always stops there.  No counterpart in the reference, which has no
 * gradients. */
int hex_arnoldi_check_fixed(const double *hcols, int first, int last, int maxdim, int fixed, double dt,
                            double beta, double *H, double *work, double *result, void *stream);
int hex_gmres_check(const double *hcols, int first, int last, int restart, int cadence, double beta, double tol_b,
                    double *H, double *result, void *stream);

/* ||x||_2 returned in *norm (host); synchronises. */
int hex_norm(hex_engine *engine, const double *x, int64_t N, double *norm, void *stream);

/* y = x / s  (s real) */
int hex_scale(const double *x, double s, double *y, int64_t N, void *stream);

/* y = sum_m Q_m c[m], c host complex[k] */
int hex_combine(hex_engine *engine, const double *Q, int64_t ldq, int k, const double *c,
                double *y, int64_t N, void *stream);

/* out = a x + b y, complex a, b; out may alias x or y. */
int hex_axpby(int64_t N, const double a[2], const double *x, const double b[2], const double *y,
              double *out, void *stream);

/* E = expm(A) for a small complex m x m matrix (row-major, m <= 64), on the
 * host, by the scaling-and-squaring Pade code the device Krylov control runs.
 * Replaces the host SciPy `expm` call of python/propagate.py:29 /
 * python/torch_backend.py:153 where no device is involved. */
int hex_expm_host(int m, const double *A, double *E);
/* expm of `count` complex m x m matrices at once on the device (A, E row-major
 * and contiguous, work complex[6 m^2 count]), one wavefront each, by the same
 * code.  The adjoint gradient forms the Frechet derivative of every sub-step's
 * exponential this way (no reference counterpart). */
int hex_expm_batched(int count, int m, const double *A, double *E, double *work, void *stream);

/* z[m] = <Q_m, v> (host complex[k]); synchronises. */
int hex_dots(hex_engine *engine, const double *Q, int64_t ldq, int k, const double *v, int64_t N,
             double *z, void *stream);
/* The same into device memory z (complex[k], k >= 1), without synchronising:
 * a caller that gathers several dot vectors reads them back once. */
int hex_dots_device(hex_engine *engine, const double *Q, int64_t ldq, int k, const double *v, int64_t N,
                    double *z, void *stream);

/* ---------------------------------------------------------------------------
 * He+ channel yield of the adjoint objective (tSURFF, sampled control) and its
 * exact gradient.  No counterpart in the reference, which has no gradients;
 * the value is the reference's online tSURFF chain (python/streaming_surface.py:
 * 39-64 and 198-214, duals by the CF4-Pade steps of python/implicit.py:111-123)
 * evaluated from recorded frames.  create: the dual operator pair h0, pz
 * (row-major complex M x M, M = na nr, already the adjoint ionic matrices),
 * the dual final states chiN (row-major M x nu, nu <= 2), the projection
 * entries E (ej = label * nout + out, ec = channel, ea = ion channel, eu =
 * dual column, ew = weight), the Volkov tables tests[nout][nE][ns],
 * harm[nout][nD], energy[nE], kz[nE][nD], wy[nE][nD], the sample times
 * times[N+1] and the step dt.  Device memory: 4 N M^2 complex for the dual
 * factor inverses plus O(N M) vectors.
 * value: A[N+1] (host) the control A(t_n), a[N+1][2] (host) A at the two CF4
 * stages of every step (row 0 unused), F0, Fd (device, complex[N+1][nc][ns][nr])
 * the frames' J and unit-field parts; *Y (host).  gradient, after value on the
 * same inputs: dA[N+1], da[N+1][2] (host) dY/dA_n and dY/d(stage fields), and
 * G0, Gd (device, frame layout) dY/dconj(F0), dY/dconj(Fd).
 * ------------------------------------------------------------------------- */
typedef struct hex_yield hex_yield;
typedef struct {
    int M, nu, na, nr, N;
    double dt;
    const double *h0, *pz, *chiN;
    int nent, nlab, nout, ns, nc;
    const int32_t *ej, *ec, *ea, *eu;
    const double *ew;
    int nE, nD;
    const double *tests, *harm, *energy, *kz, *wy, *times;
} hex_yield_desc;

int hex_yield_create(int device, const hex_yield_desc *desc, hex_yield **out);
int hex_yield_destroy(hex_yield *yield);
int64_t hex_yield_bytes(const hex_yield *yield);
int hex_yield_value(hex_yield *yield, const double *A, const double *a, const double *F0, const double *Fd,
                    double *Y, void *stream);
int hex_yield_gradient(hex_yield *yield, const double *F0, const double *Fd, double *dA, double *da, double *G0,
                       double *Gd, void *stream);

#ifdef __cplusplus
}
#endif
#endif
