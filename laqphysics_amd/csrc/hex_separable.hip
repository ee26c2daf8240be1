// Separable ionic-resolvent preconditioner of the CF4-Pade linear solves.
//
// For every channel c with one-electron labels (l1, l2) it applies
//   Out_c = S1 [ (I1 Psi_c I2^T) / (1 + i alpha (E1_i + E2_j + shift)) ] S2^T
// where S_l diagonalises the one-electron radial Hamiltonian
// T + diag(-2 cut/r + l(l+1)/2r^2) (S_l^{-1} = I_l, eigenvalues E_l), i.e. the
// exact resolvent of H1 (+) H2 (python/implicit.py:13-39).  Psi_c is the F-ordered
// n x n slab psi(:, :, c) seen as a column-major matrix (i rows, j columns).
//
// All four dense transforms are wide complex128 GEMMs (rocBLAS, FP64 MFMA):
//  * left products act on runs of consecutive channels sharing l1, whose slabs
//    are adjacent in memory, so a run is one n x (n k) right-hand side;
//  * right products act on an "l2-stacked" copy in which the slabs of all
//    channels sharing l2 are stacked vertically (a (k n) x n column-major
//    block per l2), so one GEMM per l2 value applies S_l2^{-T} or S_l2^T.
// Two permutation passes move between the layouts; the pointwise resolvent
// runs in the stacked layout.
//
// The complex64 form (python/implicit.py:27-39 with precision 'complex64',
// the reference's production setting, which its GMRES pairs with the flexible
// variant) narrows the input to complex64, runs the four transforms as
// complex64 GEMMs (FP32 matrix cores) on complex64 copies of S and S^-1 and
// the permutations on complex64 slabs, and widens the result; the resolvent's
// denominator is formed in double and the quotient rounded to complex64.
#define ROCBLAS_BETA_FEATURES_API   // rocblas_gemm_ex_get_solutions: per-shape kernel choice
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <vector>

// complex64 element (the layout of rocblas_float_complex)
struct cf {
    float x, y;
};

struct hex_separable {
    int device = 0;
    int n = 0, nc = 0, nl = 0;
    bool single = false;                // complex64 transforms
    rocblas_handle blas = nullptr;
    zd *S = nullptr, *Sinv = nullptr;   // [nl][n*n] column-major (complex128 form)
    cf *S32 = nullptr, *Sinv32 = nullptr;   // the same in complex64 (complex64 form)
    zd *E = nullptr;                    // [nl][n]
    std::vector<int> h_l1, h_l2;
    std::vector<std::pair<int, int>> runs;   // [c0, c1) with equal l1 (channel layout)
    // stacked layout: per l2 value g, channels chan[goff[g] .. goff[g+1]) stacked
    std::vector<int> goff, gchan;
    int32_t *slot = nullptr;            // device [nc]: base offset of channel c's slab in the stacked buffer
    int32_t *sld = nullptr;             // device [nc]: leading dimension (k_g n) of its block
    int32_t *l1 = nullptr, *l2 = nullptr;
    void *work = nullptr;               // two N-sized buffers of the transforms' element type
    std::vector<void *> owned;
    // rocBLAS solution index per left run / right block (0: library default),
    // chosen by timing every solution the library offers for that shape
    std::vector<int> left_sol, right_sol;
    // The GEMMs of one product are independent (disjoint column or row
    // blocks).  They are dealt round-robin over `lanes` streams -- the
    // caller's and kSepLanes - 1 own ones, forked from and joined back to the
    // caller's stream by events -- so that one GEMM's last partial wave of
    // tiles overlaps the next GEMM's first.  `lanes` (1 or kSepLanes) is chosen
    // at creation by timing a whole application both ways.
    int lanes = 1;
    std::vector<hipStream_t> aux;
    std::vector<rocblas_handle> aux_blas;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

static constexpr int kSepLanes = 2;

static void tune(hex_separable *p);
static int tune_lanes(hex_separable *p);

#define HEX_BLAS(call)                                                                 \
    do {                                                                               \
        rocblas_status st_ = (call);                                                   \
        if (st_ != rocblas_status_success) {                                           \
            hex::set_error(std::string(#call) + ": " + rocblas_status_to_string(st_)); \
            return HEX_EDEVICE;                                                        \
        }                                                                              \
    } while (0)

// channel layout <-> stacked layout.  Element (i, j) of channel c sits at
// c n^2 + j n + i (channel) and at slot[c] + j * sld[c] + i (stacked, where
// slot[c] already includes the channel's row offset q n within its block).
template <bool TO_STACKED, typename T>
__global__ void k_permute(int n, int nc, const int32_t *__restrict__ slot, const int32_t *__restrict__ sld,
                          const T *__restrict__ src, T *__restrict__ dst) {
    const int64_t nn = (int64_t)n * n;
    const int64_t total = nn * nc;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(x / nn);
        const int64_t rem = x - nn * c;
        const int j = (int)(rem / n);
        const int i = (int)(rem - (int64_t)n * j);
        const int64_t y = (int64_t)slot[c] + (int64_t)j * sld[c] + i;
        if (TO_STACKED)
            dst[y] = src[x];
        else
            dst[x] = src[y];
    }
}

__device__ __forceinline__ zd widen(zd v) { return v; }
__device__ __forceinline__ zd widen(cf v) { return zmake(v.x, v.y); }
__device__ __forceinline__ void narrow_to(zd &d, zd v) { d = v; }
__device__ __forceinline__ void narrow_to(cf &d, zd v) {
    d.x = (float)v.x;
    d.y = (float)v.y;
}

// complex128 <-> complex64 copies of a state
__global__ void k_narrow(int64_t total, const zd *__restrict__ src, cf *__restrict__ dst) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x)
        narrow_to(dst[x], src[x]);
}

__global__ void k_widen(int64_t total, const cf *__restrict__ src, zd *__restrict__ dst) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x)
        dst[x] = widen(src[x]);
}

// stacked Z /= 1 + i alpha (E[l1 c][i] + E[l2 c][j] + shift), the quotient formed in double
template <typename T>
__global__ void k_resolvent_stacked(int n, int nc, const zd *__restrict__ E, const int32_t *__restrict__ l1,
                                    const int32_t *__restrict__ l2, const int32_t *__restrict__ slot,
                                    const int32_t *__restrict__ sld, zd alpha, double shift, T *__restrict__ Z) {
    const int64_t nn = (int64_t)n * n;
    const int64_t total = nn * nc;
    const zd ia = zmake(-alpha.y, alpha.x);  // i alpha
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(x / nn);
        const int64_t rem = x - nn * c;
        const int j = (int)(rem / n);
        const int i = (int)(rem - (int64_t)n * j);
        const zd e1 = E[(int64_t)l1[c] * n + i], e2 = E[(int64_t)l2[c] * n + j];
        // reference order: values[b][j] + values[a][i], then + shift
        const zd en = zmake(e2.x + e1.x + shift, e2.y + e1.y);
        const zd den = zadd(zmake(1.0, 0.0), zmul(ia, en));
        const int64_t y = (int64_t)slot[c] + (int64_t)j * sld[c] + i;
        narrow_to(Z[y], zdiv(widen(Z[y]), den));
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
    for (rocblas_handle h : p->aux_blas) rocblas_destroy_handle(h);
    for (hipStream_t a : p->aux) (void)hipStreamDestroy(a);
    for (hipEvent_t ev : p->join) (void)hipEventDestroy(ev);
    if (p->fork) (void)hipEventDestroy(p->fork);
    for (void *q : p->owned) (void)hipFree(q);
    delete p;
    return HEX_OK;
}

static int create_impl(int device, int n, int nc, int nl, const double *S, const double *Sinv, const double *E,
                       const int32_t *l1, const int32_t *l2, bool single, hex_separable **out) {
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
    const int64_t nn = (int64_t)n * n, N = nn * nc;
    if ((int64_t)nc * n > INT32_MAX || N > INT32_MAX * 4LL) {
        hex::set_error("separable: problem too large for 32-bit stacked offsets");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(device));
    auto *p = new hex_separable();
    p->device = device;
    p->n = n;
    p->nc = nc;
    p->nl = nl;
    p->single = single;
    p->h_l1.assign(l1, l1 + nc);
    p->h_l2.assign(l2, l2 + nc);
    for (int c = 0; c < nc;) {
        int d = c + 1;
        while (d < nc && l1[d] == l1[c]) ++d;
        p->runs.push_back({c, d});
        c = d;
    }
    // stacked layout: blocks ordered by l2, channels in their original order
    std::vector<int32_t> slot(nc), sld(nc);
    p->goff.push_back(0);
    int64_t base = 0;
    bool overflow = false;
    for (int g = 0; g < nl; ++g) {
        std::vector<int> members;
        for (int c = 0; c < nc; ++c)
            if (l2[c] == g) members.push_back(c);
        const int k = (int)members.size();
        for (int q = 0; q < k; ++q) {
            const int64_t s = base + (int64_t)q * n;
            if (s > INT32_MAX) overflow = true;
            slot[members[q]] = (int32_t)s;
            sld[members[q]] = k * n;
            p->gchan.push_back(members[q]);
        }
        base += (int64_t)k * nn;
        p->goff.push_back((int)p->gchan.size());
    }
    if (overflow) {
        hex_separable_destroy(p);
        hex::set_error("separable: problem too large for 32-bit stacked offsets");
        return HEX_EINVAL;
    }
    int rc = HEX_OK;
    auto fail = [&](int code) {
        hex_separable_destroy(p);
        return code;
    };
    if ((rc = dev_alloc(p, (size_t)n * nl, &p->E)) || (rc = dev_alloc(p, (size_t)nc, &p->l1)) ||
        (rc = dev_alloc(p, (size_t)nc, &p->l2)) || (rc = dev_alloc(p, (size_t)nc, &p->slot)) ||
        (rc = dev_alloc(p, (size_t)nc, &p->sld)))
        return fail(rc);
    bool uploaded = true;
    if (single) {
        // the complex64 form keeps only complex64 copies of the eigenbases
        cf *w = nullptr;
        if ((rc = dev_alloc(p, nn * nl, &p->S32)) || (rc = dev_alloc(p, nn * nl, &p->Sinv32)) ||
            (rc = dev_alloc(p, 2 * (size_t)N, &w)))
            return fail(rc);
        p->work = w;
        std::vector<cf> s32(nn * nl), i32(nn * nl);
        const zd *sz = reinterpret_cast<const zd *>(S), *iz = reinterpret_cast<const zd *>(Sinv);
        for (int64_t q = 0; q < nn * nl; ++q) {
            s32[q] = cf{(float)sz[q].x, (float)sz[q].y};
            i32[q] = cf{(float)iz[q].x, (float)iz[q].y};
        }
        uploaded = hipMemcpy(p->S32, s32.data(), sizeof(cf) * nn * nl, hipMemcpyHostToDevice) == hipSuccess &&
                   hipMemcpy(p->Sinv32, i32.data(), sizeof(cf) * nn * nl, hipMemcpyHostToDevice) == hipSuccess;
    } else {
        zd *w = nullptr;
        if ((rc = dev_alloc(p, nn * nl, &p->S)) || (rc = dev_alloc(p, nn * nl, &p->Sinv)) ||
            (rc = dev_alloc(p, 2 * (size_t)N, &w)))
            return fail(rc);
        p->work = w;
        uploaded = hipMemcpy(p->S, S, sizeof(zd) * nn * nl, hipMemcpyHostToDevice) == hipSuccess &&
                   hipMemcpy(p->Sinv, Sinv, sizeof(zd) * nn * nl, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!uploaded || hipMemcpy(p->E, E, sizeof(zd) * n * nl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->l1, l1, sizeof(int32_t) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->l2, l2, sizeof(int32_t) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->slot, slot.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->sld, sld.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice) != hipSuccess) {
        hex::set_error("separable: table upload failed");
        return fail(HEX_EDEVICE);
    }
    if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
        hex::set_error("separable: rocblas_create_handle failed");
        return fail(HEX_EDEVICE);
    }
    rocblas_set_pointer_mode(p->blas, rocblas_pointer_mode_host);
    if (hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess) {
        hex::set_error("separable: event creation failed");
        return fail(HEX_EDEVICE);
    }
    for (int a = 1; a < kSepLanes; ++a) {
        hipStream_t sa = nullptr;
        rocblas_handle ha = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&sa, hipStreamNonBlocking) != hipSuccess) {
            hex::set_error("separable: stream creation failed");
            return fail(HEX_EDEVICE);
        }
        p->aux.push_back(sa);
        if (rocblas_create_handle(&ha) != rocblas_status_success) {
            hex::set_error("separable: rocblas_create_handle failed");
            return fail(HEX_EDEVICE);
        }
        p->aux_blas.push_back(ha);
        rocblas_set_pointer_mode(ha, rocblas_pointer_mode_host);
        rocblas_set_stream(ha, sa);
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            hex::set_error("separable: event creation failed");
            return fail(HEX_EDEVICE);
        }
        p->join.push_back(ev);
    }
    // tune on zeroed work buffers (timing only; the results are discarded)
    if (hipMemset(p->work, 0, (single ? sizeof(cf) : sizeof(zd)) * 2 * N) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        hex::set_error("separable: work buffer initialisation failed");
        return fail(HEX_EDEVICE);
    }
    tune(p);
    if (hipDeviceSynchronize() != hipSuccess) {
        hex::set_error("separable: GEMM tuning failed");
        return fail(HEX_EDEVICE);
    }
    if ((rc = tune_lanes(p))) return fail(rc);
    *out = p;
    return HEX_OK;
}

extern "C" int hex_separable_create(int device, int n, int nc, int nl, const double *S, const double *Sinv,
                                    const double *E, const int32_t *l1, const int32_t *l2, hex_separable **out) {
    return create_impl(device, n, nc, nl, S, Sinv, E, l1, l2, false, out);
}

extern "C" int hex_separable_create_c64(int device, int n, int nc, int nl, const double *S, const double *Sinv,
                                        const double *E, const int32_t *l1, const int32_t *l2, hex_separable **out) {
    return create_impl(device, n, nc, nl, S, Sinv, E, l1, l2, true, out);
}

extern "C" int hex_separable_single(const hex_separable *p) { return p ? (int)p->single : -1; }

extern "C" int64_t hex_separable_bytes(const hex_separable *p) {
    if (!p) return 0;
    const int64_t nn = (int64_t)p->n * p->n;
    const int64_t elem = p->single ? 8 : 16;   // bytes of one transform element
    return elem * (2 * nn * p->nl + 2 * nn * p->nc) + 16 * (int64_t)p->n * p->nl;
}

// complex128 or complex64 GEMM C = op(A) op(B) with rocBLAS solution `sol` (0: the library's choice)
template <typename T> struct GemmType;
template <> struct GemmType<zd> {
    static constexpr rocblas_datatype type = rocblas_datatype_f64_c;
    using scalar = rocblas_double_complex;
};
template <> struct GemmType<cf> {
    static constexpr rocblas_datatype type = rocblas_datatype_f32_c;
    using scalar = rocblas_float_complex;
};

template <typename T>
static rocblas_status gemm_sol(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                               const T *A, int lda, const T *B, int ldb, T *C, int ldc, int sol) {
    using G = GemmType<T>;
    const typename G::scalar one{1.0, 0.0}, zero{0.0, 0.0};
    return rocblas_gemm_ex(h, ta, tb, m, n, k, &one, A, G::type, lda, B, G::type, ldb, &zero, C, G::type, ldc, C,
                           G::type, ldc, G::type, sol ? rocblas_gemm_algo_solution_index : rocblas_gemm_algo_standard,
                           sol, rocblas_gemm_flags_none);
}

// Fastest rocBLAS solution for one GEMM shape, timed on the given operands
// (3 launches each after one warm-up); 0 (library default) when the library
// offers no list or nothing beats the default by more than 2 %.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"   // the solution list is a beta rocBLAS API
template <typename T>
static int tune_gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, const T *A,
                     int lda, const T *B, int ldb, T *C, int ldc) {
    using G = GemmType<T>;
    const typename G::scalar one{1.0, 0.0}, zero{0.0, 0.0};
    rocblas_int size = 0;
    if (rocblas_gemm_ex_get_solutions(h, ta, tb, m, n, k, &one, A, G::type, lda, B, G::type, ldb, &zero, C, G::type,
                                      ldc, C, G::type, ldc, G::type, rocblas_gemm_algo_solution_index,
                                      rocblas_gemm_flags_none, nullptr, &size) != rocblas_status_success ||
        size <= 0)
        return 0;
    std::vector<rocblas_int> sols(size);
    if (rocblas_gemm_ex_get_solutions(h, ta, tb, m, n, k, &one, A, G::type, lda, B, G::type, ldb, &zero, C, G::type,
                                      ldc, C, G::type, ldc, G::type, rocblas_gemm_algo_solution_index,
                                      rocblas_gemm_flags_none, sols.data(), &size) != rocblas_status_success)
        return 0;
    hipStream_t st = nullptr;
    rocblas_get_stream(h, &st);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 0;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 0;
    }
    auto timed = [&](int sol) -> float {
        if (gemm_sol(h, ta, tb, m, n, k, A, lda, B, ldb, C, ldc, sol) != rocblas_status_success) return 1e30f;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 3; ++r) gemm_sol(h, ta, tb, m, n, k, A, lda, B, ldb, C, ldc, sol);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float ms = 1e30f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    };
    const float base = timed(0);
    float best = base;
    int pick = 0;
    for (rocblas_int sol : sols) {
        if (sol == 0) continue;
        const float t = timed(sol);
        if (t < best) {
            best = t;
            pick = sol;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best < 0.98f * base ? pick : 0;
}
#pragma clang diagnostic pop

// GEMM q of a product goes to lane q % lanes; fork_lanes / join_lanes bracket
// the product on the caller's stream st.
static int fork_lanes(hex_separable *p, hipStream_t st) {
    if (p->lanes < 2) return HEX_OK;
    HEX_HIP(hipEventRecord(p->fork, st));
    for (int a = 0; a + 1 < p->lanes; ++a) HEX_HIP(hipStreamWaitEvent(p->aux[a], p->fork, 0));
    return HEX_OK;
}

static int join_lanes(hex_separable *p, hipStream_t st) {
    for (int a = 0; a + 1 < p->lanes; ++a) {
        HEX_HIP(hipEventRecord(p->join[a], p->aux[a]));
        HEX_HIP(hipStreamWaitEvent(st, p->join[a], 0));
    }
    return HEX_OK;
}

static rocblas_handle lane_handle(hex_separable *p, int q) {
    const int a = q % p->lanes;
    return a == 0 ? p->blas : p->aux_blas[a - 1];
}

// out(run) = M[l1] . in(run) for every run of channels sharing l1 (channel layout)
template <typename T>
static int left_product(hex_separable *p, const T *M, const T *in, T *out, hipStream_t st) {
    const int64_t nn = (int64_t)p->n * p->n;
    int rc = fork_lanes(p, st);
    if (rc) return rc;
    for (size_t r = 0; r < p->runs.size(); ++r) {
        const int c0 = p->runs[r].first, k = p->runs[r].second - p->runs[r].first;
        HEX_BLAS(gemm_sol(lane_handle(p, (int)r), rocblas_operation_none, rocblas_operation_none, p->n, p->n * k,
                          p->n, M + nn * p->h_l1[c0], p->n, in + nn * c0, p->n, out + nn * c0, p->n,
                          p->left_sol[r]));
    }
    return join_lanes(p, st);
}

// out(block g) = in(block g) . M[g]^T for every l2 block (stacked layout)
template <typename T>
static int right_product(hex_separable *p, const T *M, const T *in, T *out, hipStream_t st) {
    const int64_t nn = (int64_t)p->n * p->n;
    int rc = fork_lanes(p, st);
    if (rc) return rc;
    int64_t base = 0;
    int q = 0;
    for (int g = 0; g < p->nl; ++g) {
        const int k = p->goff[g + 1] - p->goff[g];
        if (k == 0) continue;
        const int rows = k * p->n;
        HEX_BLAS(gemm_sol(lane_handle(p, q++), rocblas_operation_none, rocblas_operation_transpose, rows, p->n, p->n,
                          in + base, rows, M + nn * g, p->n, out + base, rows, p->right_sol[g]));
        base += (int64_t)k * nn;
    }
    return join_lanes(p, st);
}

// Choose the solution of every left run and right block once, on the
// preconditioner's own matrices and work buffers.
template <typename T>
static void tune_as(hex_separable *p, const T *S) {
    const int64_t nn = (int64_t)p->n * p->n, N = nn * p->nc;
    T *W = static_cast<T *>(p->work), *X = W + N;
    p->left_sol.assign(p->runs.size(), 0);
    p->right_sol.assign(p->nl, 0);
    for (size_t r = 0; r < p->runs.size(); ++r) {
        const int c0 = p->runs[r].first, k = p->runs[r].second - p->runs[r].first;
        p->left_sol[r] = tune_gemm(p->blas, rocblas_operation_none, rocblas_operation_none, p->n, p->n * k, p->n,
                                   S + nn * p->h_l1[c0], p->n, X + nn * c0, p->n, W + nn * c0, p->n);
    }
    int64_t base = 0;
    for (int g = 0; g < p->nl; ++g) {
        const int k = p->goff[g + 1] - p->goff[g];
        if (k == 0) continue;
        const int rows = k * p->n;
        p->right_sol[g] = tune_gemm(p->blas, rocblas_operation_none, rocblas_operation_transpose, rows, p->n, p->n,
                                    X + base, rows, S + nn * g, p->n, W + base, rows);
        base += (int64_t)k * nn;
    }
}

static void tune(hex_separable *p) {
    if (p->single)
        tune_as<cf>(p, p->S32);
    else
        tune_as<zd>(p, p->S);
}

extern "C" int hex_separable_apply(hex_separable *p, double alpha_re, double alpha_im, double shift, const double *x,
                                   double *out, void *stream) {
    if (!p || !x || !out) {
        hex::set_error("separable apply: null argument");
        return HEX_EINVAL;
    }
    if (x == out) {
        hex::set_error("separable apply: input and output must not alias");
        return HEX_EINVAL;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    HEX_BLAS(rocblas_set_stream(p->blas, st));
    const int64_t nn = (int64_t)p->n * p->n, N = nn * p->nc;
    const zd *in = reinterpret_cast<const zd *>(x);
    zd *res = reinterpret_cast<zd *>(out);
    const int blocks = (int)std::min<int64_t>((N + 255) / 256, 8192);
    if (p->single) {
        cf *W = static_cast<cf *>(p->work), *X = W + N;
        hipLaunchKernelGGL(k_narrow, dim3(blocks), dim3(256), 0, st, N, in, X);          // X = Psi (complex64)
        HEX_HIP(hipGetLastError());
        int rc = left_product(p, p->Sinv32, X, W, st);                                  // W = I1 Psi
        if (rc) return rc;
        hipLaunchKernelGGL((k_permute<true, cf>), dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->slot, p->sld, W, X);
        HEX_HIP(hipGetLastError());
        if ((rc = right_product(p, p->Sinv32, X, W, st))) return rc;                   // W = X I2^T (stacked)
        hipLaunchKernelGGL(k_resolvent_stacked<cf>, dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->E, p->l1, p->l2,
                           p->slot, p->sld, zmake(alpha_re, alpha_im), shift, W);
        HEX_HIP(hipGetLastError());
        if ((rc = right_product(p, p->S32, W, X, st))) return rc;                      // X = W S2^T (stacked)
        hipLaunchKernelGGL((k_permute<false, cf>), dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->slot, p->sld, X, W);
        HEX_HIP(hipGetLastError());
        if ((rc = left_product(p, p->S32, W, X, st))) return rc;                       // X = S1 W
        hipLaunchKernelGGL(k_widen, dim3(blocks), dim3(256), 0, st, N, X, res);          // out = X (complex128)
        HEX_HIP(hipGetLastError());
        return HEX_OK;
    }
    zd *W = static_cast<zd *>(p->work), *X = W + N;
    int rc = left_product(p, p->Sinv, in, W, st);                          // W = I1 Psi
    if (rc) return rc;
    hipLaunchKernelGGL((k_permute<true, zd>), dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->slot, p->sld, W, X);
    HEX_HIP(hipGetLastError());
    if ((rc = right_product(p, p->Sinv, X, W, st))) return rc;             // W = X I2^T   (stacked)
    hipLaunchKernelGGL(k_resolvent_stacked<zd>, dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->E, p->l1, p->l2,
                       p->slot, p->sld, zmake(alpha_re, alpha_im), shift, W);
    HEX_HIP(hipGetLastError());
    if ((rc = right_product(p, p->S, W, X, st))) return rc;                // X = W S2^T   (stacked)
    hipLaunchKernelGGL((k_permute<false, zd>), dim3(blocks), dim3(256), 0, st, p->n, p->nc, p->slot, p->sld, X, W);
    HEX_HIP(hipGetLastError());
    return left_product(p, p->S, W, res, st);                             // out = S1 W
}

// One stream or kSepLanes for the GEMMs of a product: whole applications on
// zeroed operands (the GEMM kernels' run time does not depend on the values)
// are timed both ways, and the lanes are kept only if they are more than 2 %
// faster.
static int tune_lanes(hex_separable *p) {
    p->lanes = 1;
    if (kSepLanes < 2) return HEX_OK;
    const int64_t N = (int64_t)p->n * p->n * p->nc;
    void *buf = nullptr;
    HEX_HIP(hipMalloc(&buf, sizeof(zd) * 2 * N));
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms[2] = {1e30f, 1e30f};
    int rc = HEX_OK;
    if (hipMemset(buf, 0, sizeof(zd) * 2 * N) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        hex::set_error("separable: lane tuning set-up failed");
        rc = HEX_EDEVICE;
    }
    const double *x = static_cast<const double *>(buf);
    double *y = static_cast<double *>(buf) + 2 * N;
    for (int v = 0; v < 2 && rc == HEX_OK; ++v) {
        p->lanes = v == 0 ? 1 : kSepLanes;
        rc = hex_separable_apply(p, 0.1, 0.0, 0.0, x, y, st);
        if (rc == HEX_OK && hipEventRecord(e0, st) != hipSuccess) rc = HEX_EDEVICE;
        for (int r = 0; r < 3 && rc == HEX_OK; ++r) rc = hex_separable_apply(p, 0.1, 0.0, 0.0, x, y, st);
        if (rc == HEX_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms[v], e0, e1) != hipSuccess))
            rc = HEX_EDEVICE;
    }
    p->lanes = rc == HEX_OK && ms[1] < 0.98f * ms[0] ? kSepLanes : 1;
    if (rc == HEX_EDEVICE) hex::set_error("separable: lane tuning failed");
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    return rc;
}

extern "C" int hex_separable_lanes(const hex_separable *p) { return p ? p->lanes : 0; }

extern "C" int hex_separable_set_lanes(hex_separable *p, int lanes) {
    if (!p || (lanes != 1 && lanes != kSepLanes)) {
        hex::set_error("separable set_lanes: lanes must be 1 or " + std::to_string(kSepLanes));
        return HEX_EINVAL;
    }
    p->lanes = lanes;
    return HEX_OK;
}
