// Separable ionic-resolvent preconditioner of the CF4-Pade linear solves.
//
// For every channel c with one-electron labels (l1, l2) it applies
//   Out_c = S1 [ (I1 Psi_c I2^T) / (1 + i alpha (E1_i + E2_j + shift)) ] S2^T
// where S_l diagonalises the one-electron radial Hamiltonian
// T + diag(-2 cut/r + l(l+1)/2r^2) (S_l^{-1} = I_l, eigenvalues E_l), i.e. the
// exact resolvent of H1 (+) H2 (python/implicit.py:13-39).  Psi_c is the F-ordered
// n x n slab psi(:, :, c) seen as a column-major matrix (i rows, j columns).
//
// Dense transforms are plain complex128 GEMMs and go to rocBLAS (MFMA FP64):
// left products for runs of consecutive channels sharing l1 are one wide GEMM
// (the slabs are adjacent in memory), right products are one pointer-batched
// GEMM over all channels.  The pointwise resolvent is a HIP kernel.
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include <rocblas/rocblas.h>

#include <vector>

struct hex_separable {
    int device = 0;
    int n = 0, nc = 0, nl = 0;
    rocblas_handle blas = nullptr;
    zd *S = nullptr, *Sinv = nullptr;   // [nl][n*n] column-major
    zd *E = nullptr;                    // [nl][n]
    int32_t *l1 = nullptr, *l2 = nullptr;  // device copies [nc]
    std::vector<int> h_l1, h_l2;
    std::vector<std::pair<int, int>> runs;  // [c0, c1) with equal l1
    zd *work = nullptr;                 // two N-sized buffers
    // pointer arrays for the batched right products, one set per direction
    const zd **pA_in = nullptr, **pB_in = nullptr;
    zd **pC_in = nullptr;
    const zd **pA_out = nullptr, **pB_out = nullptr;
    zd **pC_out = nullptr;
    std::vector<void *> owned;
};

#define HEX_BLAS(call)                                                            \
    do {                                                                          \
        rocblas_status st_ = (call);                                              \
        if (st_ != rocblas_status_success) {                                      \
            hex::set_error(std::string(#call) + ": " + rocblas_status_to_string(st_)); \
            return HEX_EDEVICE;                                                   \
        }                                                                         \
    } while (0)

// Z[c][i + n j] /= 1 + i alpha (E[l1 c][i] + E[l2 c][j] + shift)
__global__ void k_resolvent(int n, int nc, const zd *__restrict__ E, const int32_t *__restrict__ l1,
                            const int32_t *__restrict__ l2, zd alpha, double shift, zd *__restrict__ Z) {
    const int64_t nn = (int64_t)n * n;
    const int64_t total = nn * nc;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(x / nn);
        const int64_t rem = x - nn * c;
        const int j = (int)(rem / n);
        const int i = (int)(rem - (int64_t)n * j);
        const zd e1 = E[(int64_t)l1[c] * n + i], e2 = E[(int64_t)l2[c] * n + j];
        // reference order: values[b][j] + values[a][i], then + shift
        const zd en = zmake(e2.x + e1.x + shift, e2.y + e1.y);
        const zd ia = zmake(-alpha.y, alpha.x);  // i alpha
        const zd den = zadd(zmake(1.0, 0.0), zmul(ia, en));
        Z[x] = zdiv(Z[x], den);
    }
}

template <typename T>
static int dev_alloc(hex_separable *p, size_t count, T **out) {
    void *q = nullptr;
    HEX_HIP(hipMalloc(&q, count * sizeof(T)));
    p->owned.push_back(q);
    *out = static_cast<T *>(q);
    return HEX_OK;
}

extern "C" int hex_separable_destroy(hex_separable *p) {
    if (!p) return HEX_OK;
    (void)hipSetDevice(p->device);
    if (p->blas) rocblas_destroy_handle(p->blas);
    for (void *q : p->owned) (void)hipFree(q);
    delete p;
    return HEX_OK;
}

extern "C" int hex_separable_create(int device, int n, int nc, int nl, const double *S, const double *Sinv,
                                    const double *E, const int32_t *l1, const int32_t *l2, hex_separable **out) {
    if (!out || n < 1 || nc < 1 || nl < 1 || !S || !Sinv || !E || !l1 || !l2) {
        hex::set_error("separable: bad argument");
        return HEX_EINVAL;
    }
    *out = nullptr;
    for (int c = 0; c < nc; ++c)
        if (l1[c] < 0 || l1[c] >= nl || l2[c] < 0 || l2[c] >= nl) {
            hex::set_error("separable: channel label outside the eigenbasis table");
            return HEX_EINVAL;
        }
    HEX_HIP(hipSetDevice(device));
    auto *p = new hex_separable();
    p->device = device;
    p->n = n;
    p->nc = nc;
    p->nl = nl;
    p->h_l1.assign(l1, l1 + nc);
    p->h_l2.assign(l2, l2 + nc);
    for (int c = 0; c < nc;) {
        int d = c + 1;
        while (d < nc && l1[d] == l1[c]) ++d;
        p->runs.push_back({c, d});
        c = d;
    }
    const size_t nn = (size_t)n * n, N = nn * nc;
    int rc = HEX_OK;
    auto fail = [&](int code) {
        hex_separable_destroy(p);
        return code;
    };
    if ((rc = dev_alloc(p, nn * nl, &p->S)) || (rc = dev_alloc(p, nn * nl, &p->Sinv)) ||
        (rc = dev_alloc(p, (size_t)n * nl, &p->E)) || (rc = dev_alloc(p, (size_t)nc, &p->l1)) ||
        (rc = dev_alloc(p, (size_t)nc, &p->l2)) || (rc = dev_alloc(p, 2 * N, &p->work)))
        return fail(rc);
    if ((rc = dev_alloc(p, (size_t)nc, &p->pA_in)) || (rc = dev_alloc(p, (size_t)nc, &p->pB_in)) ||
        (rc = dev_alloc(p, (size_t)nc, &p->pC_in)) || (rc = dev_alloc(p, (size_t)nc, &p->pA_out)) ||
        (rc = dev_alloc(p, (size_t)nc, &p->pB_out)) || (rc = dev_alloc(p, (size_t)nc, &p->pC_out)))
        return fail(rc);
    if (hipMemcpy(p->S, S, sizeof(zd) * nn * nl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->Sinv, Sinv, sizeof(zd) * nn * nl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->E, E, sizeof(zd) * n * nl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->l1, l1, sizeof(int32_t) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->l2, l2, sizeof(int32_t) * nc, hipMemcpyHostToDevice) != hipSuccess) {
        hex::set_error("separable: table upload failed");
        return fail(HEX_EDEVICE);
    }
    // Right-product pointer arrays (fixed): first pass Z_c = W_c I2^T with
    // W = work[0], Z = work[1]; second pass Z_c = V_c S2^T with V = work[0].
    std::vector<const zd *> a(nc), b(nc);
    std::vector<zd *> cptr(nc);
    zd *W = p->work, *Z = p->work + N;
    for (int c = 0; c < nc; ++c) {
        a[c] = W + nn * c;
        b[c] = p->Sinv + nn * l2[c];
        cptr[c] = Z + nn * c;
    }
    if (hipMemcpy(p->pA_in, a.data(), sizeof(void *) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->pB_in, b.data(), sizeof(void *) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->pC_in, cptr.data(), sizeof(void *) * nc, hipMemcpyHostToDevice) != hipSuccess) {
        hex::set_error("separable: pointer upload failed");
        return fail(HEX_EDEVICE);
    }
    for (int c = 0; c < nc; ++c) {
        a[c] = W + nn * c;
        b[c] = p->S + nn * l2[c];
        cptr[c] = Z + nn * c;
    }
    if (hipMemcpy(p->pA_out, a.data(), sizeof(void *) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->pB_out, b.data(), sizeof(void *) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->pC_out, cptr.data(), sizeof(void *) * nc, hipMemcpyHostToDevice) != hipSuccess) {
        hex::set_error("separable: pointer upload failed");
        return fail(HEX_EDEVICE);
    }
    if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
        hex::set_error("separable: rocblas_create_handle failed");
        return fail(HEX_EDEVICE);
    }
    rocblas_set_pointer_mode(p->blas, rocblas_pointer_mode_host);
    *out = p;
    return HEX_OK;
}

extern "C" int64_t hex_separable_bytes(const hex_separable *p) {
    if (!p) return 0;
    const int64_t nn = (int64_t)p->n * p->n;
    return 16 * (2 * nn * p->nl + (int64_t)p->n * p->nl + 2 * nn * p->nc);
}

static int left_product(hex_separable *p, const zd *M, const zd *in, zd *out) {
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    const int64_t nn = (int64_t)p->n * p->n;
    for (auto run : p->runs) {
        const int c0 = run.first, k = run.second - run.first;
        const zd *A = M + nn * p->h_l1[c0];
        HEX_BLAS(rocblas_zgemm(p->blas, rocblas_operation_none, rocblas_operation_none, p->n, p->n * k, p->n, &one,
                               reinterpret_cast<const rocblas_double_complex *>(A), p->n,
                               reinterpret_cast<const rocblas_double_complex *>(in + nn * c0), p->n, &zero,
                               reinterpret_cast<rocblas_double_complex *>(out + nn * c0), p->n));
    }
    return HEX_OK;
}

extern "C" int hex_separable_apply(hex_separable *p, double alpha_re, double alpha_im, double shift, const double *x,
                                   double *out, void *stream) {
    if (!p || !x || !out) {
        hex::set_error("separable apply: null argument");
        return HEX_EINVAL;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    HEX_BLAS(rocblas_set_stream(p->blas, st));
    const int64_t nn = (int64_t)p->n * p->n, N = nn * p->nc;
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    zd *W = p->work, *Z = p->work + N;
    const zd *in = reinterpret_cast<const zd *>(x);
    zd *res = reinterpret_cast<zd *>(out);
    // W_c = I1 Psi_c
    int rc = left_product(p, p->Sinv, in, W);
    if (rc) return rc;
    // Z_c = W_c I2^T
    HEX_BLAS(rocblas_zgemm_batched(p->blas, rocblas_operation_none, rocblas_operation_transpose, p->n, p->n, p->n,
                                   &one, reinterpret_cast<const rocblas_double_complex *const *>(p->pA_in), p->n,
                                   reinterpret_cast<const rocblas_double_complex *const *>(p->pB_in), p->n, &zero,
                                   reinterpret_cast<rocblas_double_complex *const *>(p->pC_in), p->n, p->nc));
    // pointwise resolvent
    const int blocks = (int)std::min<int64_t>((N + 255) / 256, 8192);
    hipLaunchKernelGGL(k_resolvent, dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->E, p->l1, p->l2,
                       zmake(alpha_re, alpha_im), shift, Z);
    HEX_HIP(hipGetLastError());
    // V_c = S1 Z_c  (into W)
    rc = left_product(p, p->S, Z, W);
    if (rc) return rc;
    // Z_c = V_c S2^T, then out = Z
    HEX_BLAS(rocblas_zgemm_batched(p->blas, rocblas_operation_none, rocblas_operation_transpose, p->n, p->n, p->n,
                                   &one, reinterpret_cast<const rocblas_double_complex *const *>(p->pA_out), p->n,
                                   reinterpret_cast<const rocblas_double_complex *const *>(p->pB_out), p->n, &zero,
                                   reinterpret_cast<rocblas_double_complex *const *>(p->pC_out), p->n, p->nc));
    HEX_HIP(hipMemcpyAsync(res, Z, sizeof(zd) * N, hipMemcpyDeviceToDevice, st));
    return HEX_OK;
}
