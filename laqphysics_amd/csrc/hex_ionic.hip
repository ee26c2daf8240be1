// He+ ionic dual propagation on the device: the multi-column operator and the
// field-free resolvent preconditioner of the CF4-Pade dual steps
// (python/ionic_adjoint.py:12-52, driven as in python/streaming_surface.py:125-146).
//
// Operator.  The adjoint ionic Hamiltonian is G(f) = G_0 + sum_k f_k G_k, where
// every G_k (the field-free block Hamiltonian and one matrix per field
// coefficient, all already conjugate-transposed) is stored on one shared CSR
// pattern of the ndof = na * n rows.  A dual state holds nrhs columns, row
// major: element (row, c) at row * nrhs + c.  The combined values of G(f) are
// formed once per coefficient vector (k_ionic_combine); an apply is then one
// multi-column sparse product y = a G x + b x (k_ionic_apply).
//
// Sparse product mapping (wave64): one wave per row.  The lanes split into
// G = 64 / C groups of C lanes, C the power of two >= nrhs (at most 64): lane
// (g, c) accumulates column c over the row's nonzeros g, g + G, ..., and a
// butterfly over the groups sums the partial products in a fixed order
// (deterministic).  The row's column indices and values are read coalesced,
// 64 at a time, and broadcast to the groups by wave shuffles, so the only
// gathers are those of x (16 nrhs contiguous bytes per nonzero), four in
// flight per lane.
//
// Preconditioner.  For every state a (angular label (l, m)) the field-free
// block is G_0's diagonal block of that l, diagonalised on the host as
// S_l diag(E_l) S_l^{-1}.  The resolvent of that block,
//   y_a = S_l diag(1 / (1 + i alpha (E_l + shift))) S_l^{-1} x_a,
// is two complex128 GEMMs batched over the states (rocBLAS strided batched on
// FP64 MFMA) around a pointwise scaling.  The (n, nrhs) row-major slab x_a is
// the column-major nrhs x n matrix X'_a with leading dimension nrhs, so
// S^{-1} x_a = (X'_a S^{-T})^T: with S^{-1} stored row-major the product is
// X'_a times that buffer read column-major (no transposition).  The per-l
// matrices are replicated per state so that one strided batch covers all
// states (na n^2 complex values per matrix set).
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include <rocblas/rocblas.h>

#include <cstring>
#include <vector>

static constexpr int kIonMaxCoef = 16;   // field coefficients beyond G_0

struct IonCoef {
    double c[kIonMaxCoef];
};

struct hex_ionic {
    int device = 0;
    int64_t ndof = 0;
    int nrhs = 0, nnz = 0, nk = 0;
    int n = 0, na = 0;                  // radial size and state count of the preconditioner
    int32_t *ptr = nullptr, *col = nullptr;
    zd *data = nullptr;                 // [nk + 1][nnz]
    zd *values = nullptr;               // [nnz]: G(f) for the cached coefficients
    std::vector<double> key;            // coefficients of `values`; empty: none formed yet
    zd *S = nullptr, *Sinv = nullptr;   // [na][n*n] row-major (replicated per state)
    zd *E = nullptr;                    // [na][n]
    zd *work = nullptr;                 // [ndof * nrhs]
    rocblas_handle blas = nullptr;
};

namespace {

int ion_alloc(size_t count, zd **out) {
    if (hipMalloc(out, sizeof(zd) * (count ? count : 1)) != hipSuccess) {
        hex::set_error("ionic: device allocation failed");
        return HEX_EDEVICE;
    }
    return HEX_OK;
}

__global__ __launch_bounds__(256) void k_ionic_combine(int nnz, int nk, const zd *__restrict__ data, IonCoef coef,
                                                       zd *__restrict__ values) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    zd v = data[k];
    for (int m = 0; m < nk; ++m) {
        const zd d = data[(size_t)(m + 1) * nnz + k];
        v.x += coef.c[m] * d.x;
        v.y += coef.c[m] * d.y;
    }
    values[k] = v;
}

template <int C>
__global__ __launch_bounds__(256) void k_ionic_apply(int64_t ndof, int nrhs, const int32_t *__restrict__ ptr,
                                                     const int32_t *__restrict__ col, const zd *__restrict__ val,
                                                     zd a, zd b, const zd *__restrict__ x, zd *__restrict__ y) {
    constexpr int G = 64 / C;
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= ndof) return;            // uniform over the wave
    const int g = lane / C, c = lane % C;
    const int start = ptr[row], end = ptr[row + 1];
    for (int c0 = 0; c0 < nrhs; c0 += C) {
        const int cc = c0 + c;
        const bool on = cc < nrhs;
        // four independent partial sums keep four gathers of x in flight per lane
        zd acc0 = zmake(0., 0.), acc1 = acc0, acc2 = acc0, acc3 = acc0;
        for (int base = start; base < end; base += 64) {
            // the row's next (up to) 64 nonzeros, one per lane, read coalesced;
            // lanes past the row's end hold column 0 with value 0
            const int cnt = min(64, end - base);
            int jl = 0;
            zd vl = zmake(0., 0.);
            if (lane < cnt) {
                jl = col[base + lane];
                vl = val[base + lane];
            }
            // group g takes entries g, g + G, ...; the trip count is the same on
            // every lane, so every shuffle runs on the whole wave.  Entries
            // q >= cnt read a zero value (their lane index wraps modulo 64 to a
            // valid column).
            const int rounds = (cnt + G - 1) / G;
            for (int u = 0; u < rounds; u += 4) {
                int q[4];
                int j[4];
                zd v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    q[t] = g + (u + t) * G;
                    j[t] = __shfl(jl, q[t] & 63, 64);
                    v[t].x = __shfl(vl.x, q[t] & 63, 64);
                    v[t].y = __shfl(vl.y, q[t] & 63, 64);
                    if (q[t] >= cnt) v[t] = zmake(0., 0.);
                }
                if (on) {
                    const zd x0 = x[(int64_t)j[0] * nrhs + cc], x1 = x[(int64_t)j[1] * nrhs + cc];
                    const zd x2 = x[(int64_t)j[2] * nrhs + cc], x3 = x[(int64_t)j[3] * nrhs + cc];
                    acc0 = zfma(v[0], x0, acc0);
                    acc1 = zfma(v[1], x1, acc1);
                    acc2 = zfma(v[2], x2, acc2);
                    acc3 = zfma(v[3], x3, acc3);
                }
            }
        }
        zd acc = zadd(zadd(acc0, acc1), zadd(acc2, acc3));
#pragma unroll
        for (int off = C; off < 64; off <<= 1) {
            acc.x += __shfl_xor(acc.x, off, 64);
            acc.y += __shfl_xor(acc.y, off, 64);
        }
        if (g == 0 && on) {
            const int64_t e = row * nrhs + cc;
            y[e] = zfma(b, x[e], zmul(a, acc));
        }
    }
}

__global__ __launch_bounds__(256) void k_ionic_resolvent(int64_t total, int nrhs, const zd *__restrict__ E,
                                                         zd alpha, double shift, zd *__restrict__ z) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const zd w = E[e / nrhs];
    // 1 + i alpha (w + shift)
    const double wr = w.x + shift, wi = w.y;
    const zd den = zmake(1. - alpha.y * wr - alpha.x * wi, alpha.x * wr - alpha.y * wi);
    z[e] = zdiv(z[e], den);
}

int launch_apply(const hex_ionic *p, zd a, zd b, const zd *x, zd *y, hipStream_t st) {
    const int c = p->nrhs >= 64 ? 64 : p->nrhs > 32 ? 64 : p->nrhs > 16 ? 32 : p->nrhs > 8 ? 16
                : p->nrhs > 4 ? 8 : p->nrhs > 2 ? 4 : p->nrhs > 1 ? 2 : 1;
    const dim3 grid((unsigned)((p->ndof + 3) / 4)), block(256);
    switch (c) {
    case 1: hipLaunchKernelGGL(k_ionic_apply<1>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    case 2: hipLaunchKernelGGL(k_ionic_apply<2>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    case 4: hipLaunchKernelGGL(k_ionic_apply<4>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    case 8: hipLaunchKernelGGL(k_ionic_apply<8>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    case 16: hipLaunchKernelGGL(k_ionic_apply<16>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    case 32: hipLaunchKernelGGL(k_ionic_apply<32>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    default: hipLaunchKernelGGL(k_ionic_apply<64>, grid, block, 0, st, p->ndof, p->nrhs, p->ptr, p->col, p->values, a, b, x, y); break;
    }
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

}  // namespace

extern "C" int hex_ionic_destroy(hex_ionic *p) {
    if (!p) return HEX_OK;
    (void)hipSetDevice(p->device);
    if (p->blas) rocblas_destroy_handle(p->blas);
    for (void *q : {(void *)p->ptr, (void *)p->col, (void *)p->data, (void *)p->values, (void *)p->S,
                    (void *)p->Sinv, (void *)p->E, (void *)p->work})
        if (q) (void)hipFree(q);
    delete p;
    return HEX_OK;
}

extern "C" int hex_ionic_create(int device, int64_t ndof, int nrhs, int nnz, const int32_t *ptr, const int32_t *col,
                                int nk, const double *data, hex_ionic **out) {
    if (!out || !ptr || !col || !data || ndof < 1 || nrhs < 1 || nnz < 1 || nk < 0 || nk > kIonMaxCoef) {
        hex::set_error("ionic: bad arguments (1 <= nrhs, 0 <= field coefficients <= 16)");
        return HEX_EINVAL;
    }
    if (ptr[0] != 0 || ptr[ndof] != nnz) {
        hex::set_error("ionic: row pointer does not span the nonzeros");
        return HEX_EINVAL;
    }
    for (int64_t r = 0; r < ndof; ++r)
        if (ptr[r + 1] < ptr[r]) {
            hex::set_error("ionic: row pointer is not monotone");
            return HEX_EINVAL;
        }
    for (int k = 0; k < nnz; ++k)
        if (col[k] < 0 || col[k] >= ndof) {
            hex::set_error("ionic: column index out of range");
            return HEX_EINVAL;
        }
    HEX_HIP(hipSetDevice(device));
    hex_ionic *p = new hex_ionic;
    p->device = device;
    p->ndof = ndof;
    p->nrhs = nrhs;
    p->nnz = nnz;
    p->nk = nk;
    auto fail = [&](int code) {
        hex_ionic_destroy(p);
        return code;
    };
    int rc;
    if (hipMalloc(&p->ptr, sizeof(int32_t) * (ndof + 1)) != hipSuccess ||
        hipMalloc(&p->col, sizeof(int32_t) * nnz) != hipSuccess) {
        hex::set_error("ionic: device allocation failed");
        return fail(HEX_EDEVICE);
    }
    if ((rc = ion_alloc((size_t)(nk + 1) * nnz, &p->data)) || (rc = ion_alloc((size_t)nnz, &p->values)) ||
        (rc = ion_alloc((size_t)ndof * nrhs, &p->work)))
        return fail(rc);
    if (hipMemcpy(p->ptr, ptr, sizeof(int32_t) * (ndof + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->data, data, sizeof(zd) * (nk + 1) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) {
        hex::set_error("ionic: table upload failed");
        return fail(HEX_EDEVICE);
    }
    if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
        hex::set_error("ionic: rocblas_create_handle failed");
        return fail(HEX_EDEVICE);
    }
    rocblas_set_pointer_mode(p->blas, rocblas_pointer_mode_host);
    *out = p;
    return HEX_OK;
}

extern "C" int hex_ionic_set_eigen(hex_ionic *p, int n, int na, const double *S, const double *Sinv, const double *E) {
    if (!p || !S || !Sinv || !E || n < 1 || na < 1 || (int64_t)n * na != p->ndof) {
        hex::set_error("ionic: eigenbasis shape must be na x n with na * n = ndof");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(p->device));
    const size_t nn = (size_t)n * n * na;
    for (zd *q : {p->S, p->Sinv, p->E})
        if (q) (void)hipFree(q);
    p->S = p->Sinv = p->E = nullptr;
    int rc;
    if ((rc = ion_alloc(nn, &p->S)) || (rc = ion_alloc(nn, &p->Sinv)) || (rc = ion_alloc((size_t)n * na, &p->E)))
        return rc;
    HEX_HIP(hipMemcpy(p->S, S, sizeof(zd) * nn, hipMemcpyHostToDevice));
    HEX_HIP(hipMemcpy(p->Sinv, Sinv, sizeof(zd) * nn, hipMemcpyHostToDevice));
    HEX_HIP(hipMemcpy(p->E, E, sizeof(zd) * n * (size_t)na, hipMemcpyHostToDevice));
    p->n = n;
    p->na = na;
    return HEX_OK;
}

extern "C" int hex_ionic_apply(hex_ionic *p, const double *coef, const double a[2], const double b[2],
                               const double *x, double *y, void *stream) {
    if (!p || !coef || !a || !b || !x || !y || x == y) {
        hex::set_error("ionic apply: null argument or output aliasing the input");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    if (p->key.size() != (size_t)p->nk || std::memcmp(p->key.data(), coef, sizeof(double) * p->nk) != 0) {
        IonCoef c{};
        for (int m = 0; m < p->nk; ++m) c.c[m] = coef[m];
        hipLaunchKernelGGL(k_ionic_combine, dim3((p->nnz + 255) / 256), dim3(256), 0, st, p->nnz, p->nk, p->data, c,
                           p->values);
        HEX_HIP(hipGetLastError());
        p->key.assign(coef, coef + p->nk);
    }
    return launch_apply(p, zmake(a[0], a[1]), zmake(b[0], b[1]), (const zd *)x, (zd *)y, st);
}

extern "C" int hex_ionic_precondition(hex_ionic *p, double alpha_re, double alpha_im, double shift, const double *x,
                                      double *y, void *stream) {
    if (!p || !x || !y || !p->S) {
        hex::set_error("ionic precondition: no eigenbasis set");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    rocblas_set_stream(p->blas, st);
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    const int n = p->n, m = p->nrhs;
    const rocblas_stride sx = (rocblas_stride)n * m, ss = (rocblas_stride)n * n;
    // work = X' Sinv^T per state (nrhs x n, leading dimension nrhs)
    rocblas_status s1 = rocblas_zgemm_strided_batched(
        p->blas, rocblas_operation_none, rocblas_operation_none, m, n, n, &one, (const rocblas_double_complex *)x, m,
        sx, (const rocblas_double_complex *)p->Sinv, n, ss, &zero, (rocblas_double_complex *)p->work, m, sx, p->na);
    if (s1 != rocblas_status_success) {
        hex::set_error(std::string("ionic precondition: ") + rocblas_status_to_string(s1));
        return HEX_EDEVICE;
    }
    const int64_t total = p->ndof * m;
    hipLaunchKernelGGL(k_ionic_resolvent, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total, m, p->E,
                       zmake(alpha_re, alpha_im), shift, p->work);
    HEX_HIP(hipGetLastError());
    rocblas_status s2 = rocblas_zgemm_strided_batched(
        p->blas, rocblas_operation_none, rocblas_operation_none, m, n, n, &one,
        (const rocblas_double_complex *)p->work, m, sx, (const rocblas_double_complex *)p->S, n, ss, &zero,
        (rocblas_double_complex *)y, m, sx, p->na);
    if (s2 != rocblas_status_success) {
        hex::set_error(std::string("ionic precondition: ") + rocblas_status_to_string(s2));
        return HEX_EDEVICE;
    }
    return HEX_OK;
}

extern "C" int64_t hex_ionic_bytes(const hex_ionic *p) {
    if (!p) return 0;
    return 16 * ((int64_t)(p->nk + 2) * p->nnz + p->ndof * p->nrhs + 2 * (int64_t)p->n * p->n * p->na +
                 (int64_t)p->n * p->na) + 4 * ((int64_t)p->nnz + p->ndof + 1);
}
