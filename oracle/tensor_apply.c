/*
 * ORACLE -- test infrastructure only.  Nothing in the shipped library links or
 * calls this file; tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it as the checker.
 *
 * Plain-C/OpenMP restatement of the reference H.psi kernel
 * (src/kernels.f90:119-176), same arguments, same F-order layout
 * psi(i,j,c) -> psi[i + n*j + n*n*c], same loop order (c, j outer and
 * collapsed for OpenMP, i inner), same CSR walks, same |coef| < 1e-30 edge skip
 * (kernels.f90:145) and the same velocity-gauge |A|^2 term (kernels.f90:170).
 * Indices are 0-based at the interface as in the reference's C binding.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>

typedef double _Complex zd;

/* Planted-mistake variants, compiled only by the oracle's own sensitivity test
 * (tests/test_oracle_golden.py) with -DORACLE_PLANT=k:
 *   1 drop the |A|^2 term     2 flip the sign of the velocity dipole
 *   3 walk row i for the electron-2 kinetic term   4 use r_i for electron 2
 *   5 ignore the multipole order (always the lam = 0 table) */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

void oracle_tensor_apply(int n, int nc, int nt,
                         const int32_t *ptr, const int32_t *col, const zd *tv,
                         const zd *diag,
                         int nv, const int32_t *vptr, const int32_t *vc, const int32_t *vl,
                         const zd *vcoef, const zd *rad,
                         int nd, const int32_t *dptr, const int32_t *dc, const int32_t *de,
                         const zd *dcoef, const double *ldiff,
                         const zd *r, const double *field, int velocity,
                         const zd *psi, zd *out)
{
    (void)nt; (void)nv; (void)nd;
    const int64_t nn = (int64_t)n * n;
    const double a2 = field[0] * field[0] + field[1] * field[1] + field[2] * field[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 0; c < nc; ++c) {
        for (int j = 0; j < n; ++j) {
            const zd *pc = psi + nn * c;
            for (int i = 0; i < n; ++i) {
                const int64_t at = i + (int64_t)n * j;
                zd val = diag[at + nn * c] * pc[at];
                for (int k = ptr[i]; k < ptr[i + 1]; ++k)
                    val += tv[k] * pc[col[k] + (int64_t)n * j];
                {
                    const int row = ORACLE_PLANT == 3 ? i : j;
                    for (int k = ptr[row]; k < ptr[row + 1]; ++k)
                        val += tv[k] * pc[i + (int64_t)n * col[k]];
                }
                for (int k = vptr[c]; k < vptr[c + 1]; ++k)
                    val += vcoef[k] * rad[at + nn * (ORACLE_PLANT == 5 ? 0 : vl[k])] * psi[at + nn * vc[k]];
                for (int k = dptr[c]; k < dptr[c + 1]; ++k) {
                    const zd coef = field[0] * dcoef[3 * k] + field[1] * dcoef[3 * k + 1]
                                  + field[2] * dcoef[3 * k + 2];
                    if (cabs(coef) < 1e-30) continue;
                    const zd *ps = psi + nn * dc[k];
                    if (!velocity) {
                        val += coef * (de[k] == 1 ? r[i] : r[ORACLE_PLANT == 4 ? i : j]) * ps[at];
                    } else {
                        zd part = 0;
                        if (de[k] == 1) {
                            if (ldiff[k] != 0) part = ldiff[k] / (2 * r[i]) * ps[at];
                            for (int q = ptr[i]; q < ptr[i + 1]; ++q)
                                part += tv[q] * (r[col[q]] - r[i]) * ps[col[q] + (int64_t)n * j];
                        } else {
                            const int rj = ORACLE_PLANT == 4 ? i : j;
                            if (ldiff[k] != 0) part = ldiff[k] / (2 * r[rj]) * ps[at];
                            for (int q = ptr[j]; q < ptr[j + 1]; ++q)
                                part += tv[q] * (r[col[q]] - r[rj]) * ps[i + (int64_t)n * col[q]];
                        }
                        val += (ORACLE_PLANT == 2 ? -I : I) * coef * part;
                    }
                }
                if (velocity && ORACLE_PLANT != 1) val += a2 * pc[at];
                out[at + nn * c] = val;
            }
        }
    }
}
