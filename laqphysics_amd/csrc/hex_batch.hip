// Batched propagation: B wavefunctions on one operator, one captured graph per
// Krylov attempt.
//
// Every member runs the reference's midpoint Arnoldi step
// (python/torch_backend.py:140-167: CGS2, happy breakdown below 1e-13, checks at
// j >= 3 when (j+1) % 4 == 0 or at maxdim, c = expm(-i dt H_m)[:, 0] beta,
// err = |dt h_{m+1,m} c_m|, and on failure two half steps with tol/2, depth <= 8)
// with its own sin^2 pulses, evaluated on the device at the midpoint of the
// current (sub)interval.  The recursive halving is a per-member stack of
// pending intervals (t, h, tol, depth), popped depth first exactly like the
// reference's recursion.  One replay of the recorded graph is one Krylov
// attempt for every live member; converged members leave the column loop
// early and members whose step count is reached stop, so members of uneven
// length share the launches.  No host synchronisation inside an attempt.
#include "hex_internal.h"
#include "hex_expm.h"
#include "../../include/laqphysics_hip.h"

#include <algorithm>
#include <array>
#include <utility>
#include <vector>

static constexpr int kBThreads = 256;
static constexpr int kBBlocks = 512;     // reduction blocks per member
static constexpr int kBCols = 32;        // basis vectors per reduction
static constexpr int kBChunk = 16;
static constexpr int kBatchMaxDim = 16;  // Krylov dimension held in LDS by the check kernel

struct hex_batch {
    hex_engine *e = nullptr;
    int B = 0, maxdim = 0;
    int64_t N = 0;
    double dt = 0, tol = 0;
    int device = 0, by = 4, red_blocks = kBBlocks;
    zd *psi = nullptr, *Q = nullptr, *V = nullptr;
    zd *y1 = nullptr, *y2 = nullptr;      // radial-walk scratch [B][N] each
    zd *dcz = nullptr;                    // field-coupling coefficients [B][nd]
    double *fields = nullptr;    // [B][3] current midpoint A
    double *pulses = nullptr;    // [B][kMaxPulses][7]: omega, cycles, F, start, cep, polarisation, used
    double *t0 = nullptr;        // [B] start time
    double *cur = nullptr;       // [B][3]: t, h, tol of the current attempt
    double *stk = nullptr;       // [B][kStack][3]: pending (t, h, tol)
    int32_t *sdep = nullptr;     // [B][kStack]: pending depth
    int32_t *ctl = nullptr;      // [kCtl][B]
    int32_t *counts = nullptr;   // [B]
    zd *partials = nullptr;      // [B][red_blocks][kBCols]
    zd *z = nullptr;             // [B][2 kBCols]: z1, z2
    double *nrm = nullptr, *beta = nullptr;   // [B]
    zd *Hm = nullptr;            // [B][maxdim+1][maxdim]
    zd *coef = nullptr;          // [B][maxdim]
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> owned;
};

enum { kAlive = 0, kActive = 1, kDone = 2, kDim = 3, kFail = 4, kStep = 5, kSplit = 6, kSp = 7, kAttempts = 8,
       kSubsteps = 9, kCtl = 10 };
static constexpr int kStack = 16;
static constexpr int kMaxPulses = 4;
static constexpr int kMaxDepth = 8;

__device__ __forceinline__ int32_t &ctl_at(int32_t *ctl, int B, int which, int b) { return ctl[which * B + b]; }

// ---------------------------------------------------------------------------
// control kernels
// ---------------------------------------------------------------------------
// A(t) of one sin^2 pulse (python/pulses.py:28-40): int_0^u cos(cep + w s) ds
// = u cos(cep + w u/2) sinc(w u / 2 pi), sin^2 = 1/2 - cos(b u)/2.
__device__ double pulse_vector(const double *p, double t, double cep) {
    const double omega = p[0], cycles = p[1], F = p[2], start = p[3];
    const double T = 2.0 * M_PI * cycles / omega;
    double u = t - start;
    u = u < 0.0 ? 0.0 : (u > T ? T : u);
    const double b = 2.0 * M_PI / T;
    auto integ = [&](double w) {
        const double x = w * u / 2.0;
        const double sinc = x == 0.0 ? 1.0 : sin(x) / x;
        return u * cos(cep + x) * sinc;
    };
    double value = -F * (0.5 * integ(omega) - 0.25 * integ(omega + b) - 0.25 * integ(omega - b));
    const double rc = rint(cycles);
    if (rc >= 2.0 && fabs(cycles - rc) < 1e-12 && t >= start + T) value = 0.0;
    return value;
}

// python/run3d.py:41-56: z adds A; sigma+/- add A/sqrt2 to x and
// +-A(cep - pi/2)/sqrt2 to y
__device__ void member_field(const double *pulses, double t, double *A) {
    A[0] = A[1] = A[2] = 0.0;
    const double root2 = 1.4142135623730951;
    for (int k = 0; k < kMaxPulses; ++k) {
        const double *p = pulses + 7 * k;
        if (p[6] == 0.0) continue;
        const int pol = (int)p[5];
        const double v = pulse_vector(p, t, p[4]);
        if (pol == 0) {
            A[2] += v;
        } else {
            A[0] += v / root2;
            A[1] += pulse_vector(p, t, p[4] - M_PI / 2.0) / root2 * (pol == 1 ? 1.0 : -1.0);
        }
    }
}

// Start an attempt: if no interval is pending, queue the next macro step (or
// retire the member); the top interval fixes h, tol and the midpoint field.
__global__ void kb_begin(int B, double dt, double tol, int32_t *ctl, const int32_t *counts, const double *t0,
                         const double *pulses, double *stk, int32_t *sdep, double *cur, double *fields) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    ctl_at(ctl, B, kDone, b) = 0;
    ctl_at(ctl, B, kSplit, b) = 0;
    ctl_at(ctl, B, kActive, b) = 0;
    if (!ctl_at(ctl, B, kAlive, b)) return;
    int sp = ctl_at(ctl, B, kSp, b);
    double *sb = stk + (int64_t)b * kStack * 3;
    if (sp == 0) {
        const int s = ctl_at(ctl, B, kStep, b);
        if (s >= counts[b]) {
            ctl_at(ctl, B, kAlive, b) = 0;
            return;
        }
        sb[0] = t0[b] + s * dt;
        sb[1] = dt;
        sb[2] = tol;
        sdep[b * kStack] = 0;
        sp = 1;
        ctl_at(ctl, B, kSp, b) = 1;
    }
    const double *top = sb + 3 * (sp - 1);
    cur[3 * b] = top[0];
    cur[3 * b + 1] = top[1];
    cur[3 * b + 2] = top[2];
    member_field(pulses + (int64_t)b * kMaxPulses * 7, top[0] + top[1] / 2.0, fields + 3 * b);
    ctl_at(ctl, B, kActive, b) = 1;
    ctl_at(ctl, B, kAttempts, b) += 1;
}

// Finish an attempt: accepted intervals are popped (a macro step completes when
// nothing is pending); a failed interval is replaced by its two halves with
// tol/2, first half on top, exactly the reference's recursion order.
__global__ void kb_end(int B, int32_t *ctl, double *stk, int32_t *sdep) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || !ctl_at(ctl, B, kAlive, b)) return;
    int sp = ctl_at(ctl, B, kSp, b);
    double *sb = stk + (int64_t)b * kStack * 3;
    int32_t *db = sdep + b * kStack;
    if (ctl_at(ctl, B, kDone, b)) {
        --sp;
        ctl_at(ctl, B, kSubsteps, b) += 1;
        if (sp == 0) ctl_at(ctl, B, kStep, b) += 1;
    } else if (ctl_at(ctl, B, kSplit, b)) {
        const double t = sb[3 * (sp - 1)], h = sb[3 * (sp - 1) + 1], tl = sb[3 * (sp - 1) + 2];
        const int d = db[sp - 1];
        if (d >= kMaxDepth || sp + 1 > kStack) {
            ctl_at(ctl, B, kFail, b) += 1;
            ctl_at(ctl, B, kAlive, b) = 0;
            return;
        }
        sb[3 * (sp - 1)] = t + h / 2.0;     // second half (below)
        sb[3 * (sp - 1) + 1] = h / 2.0;
        sb[3 * (sp - 1) + 2] = tl / 2.0;
        db[sp - 1] = d + 1;
        sb[3 * sp] = t;                     // first half (top)
        sb[3 * sp + 1] = h / 2.0;
        sb[3 * sp + 2] = tl / 2.0;
        db[sp] = d + 1;
        ++sp;
    }
    ctl_at(ctl, B, kSp, b) = sp;
}

// ---------------------------------------------------------------------------
// batched reductions (member = blockIdx.y); deterministic two-stage order
// ---------------------------------------------------------------------------
template <int KC>
__device__ __forceinline__ void bred_store(zd (&acc)[KC], int count, zd *partials, int col0) {
    __shared__ double red[kBThreads / 64][KC][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < KC; ++m)
        if (m < count) {
            double re = wave_sum(acc[m].x), im = wave_sum(acc[m].y);
            if (lane == 0) {
                red[wave][m][0] = re;
                red[wave][m][1] = im;
            }
        }
    __syncthreads();
    if ((int)threadIdx.x < count) {
        double re = 0.0, im = 0.0;
        for (int w = 0; w < kBThreads / 64; ++w) {
            re += red[w][threadIdx.x][0];
            im += red[w][threadIdx.x][1];
        }
        partials[(int64_t)blockIdx.x * kBCols + col0 + threadIdx.x] = zmake(re, im);
    }
}

// partial z_b[m] = Q_b[m]^H V_b for m in [c0, c0+C); grid (blocks, B).  C is a
// template argument so that the C loads of an element are issued together.
template <int C>
__global__ __launch_bounds__(kBThreads) void kb_dots(const zd *__restrict__ Q, int64_t qstride, int64_t N, int c0,
                                                     const zd *__restrict__ V, const int32_t *active, zd *partials,
                                                     int pblocks) {
    const int b = blockIdx.y;
    if (!active[b]) return;
    const zd *Qb = Q + qstride * b + (int64_t)c0 * N;
    const zd *Vb = V + N * b;
    zd acc[C];
#pragma unroll
    for (int m = 0; m < C; ++m) acc[m] = zmake(0.0, 0.0);
    for (int64_t x = blockIdx.x * (int64_t)kBThreads + threadIdx.x; x < N; x += (int64_t)gridDim.x * kBThreads) {
        const zd v = Vb[x];
        zd q[C];
#pragma unroll
        for (int m = 0; m < C; ++m) q[m] = Qb[m * N + x];
#pragma unroll
        for (int m = 0; m < C; ++m) acc[m] = zcfma(q[m], v, acc[m]);
    }
    bred_store<C>(acc, C, partials + (int64_t)b * pblocks * kBCols, c0);
}

// V_b -= sum_m Q_b[m] z_b[m] for m < full*kBChunk + R, in m order; optional
// partial ||V_b||^2.  The basis is walked in chunks of kBChunk vectors whose
// loads are issued together; R, the last partial chunk, is a template argument.
template <int R, bool NORM>
__global__ __launch_bounds__(kBThreads) void kb_update(const zd *__restrict__ Q, int64_t qstride, int64_t N,
                                                       int full, const zd *__restrict__ z, int zstride,
                                                       zd *__restrict__ V, const int32_t *active, zd *partials,
                                                       int pblocks) {
    const int b = blockIdx.y;
    if (!active[b]) return;
    __shared__ zd zs[kBCols];
    const int count = full * kBChunk + R;
    if ((int)threadIdx.x < count) zs[threadIdx.x] = z[(int64_t)b * zstride + threadIdx.x];
    __syncthreads();
    const zd *Qb = Q + qstride * b;
    zd *Vb = V + N * b;
    double nrm = 0.0;
    for (int64_t x = blockIdx.x * (int64_t)kBThreads + threadIdx.x; x < N; x += (int64_t)gridDim.x * kBThreads) {
        zd v = Vb[x];
        for (int c = 0; c < full; ++c) {
            const zd *Qc = Qb + (int64_t)c * kBChunk * N;
            zd q[kBChunk];
#pragma unroll
            for (int m = 0; m < kBChunk; ++m) q[m] = Qc[m * N + x];
#pragma unroll
            for (int m = 0; m < kBChunk; ++m) v = zsub(v, zmul(q[m], zs[c * kBChunk + m]));
        }
        if constexpr (R > 0) {
            const zd *Qc = Qb + (int64_t)full * kBChunk * N;
            zd q[R];
#pragma unroll
            for (int m = 0; m < R; ++m) q[m] = Qc[m * N + x];
#pragma unroll
            for (int m = 0; m < R; ++m) v = zsub(v, zmul(q[m], zs[full * kBChunk + m]));
        }
        if (count > 0) st_stream(Vb + x, v);
        if (NORM) nrm += zabs2(v);
    }
    if (NORM) {
        zd one[1] = {zmake(nrm, 0.0)};
        bred_store<1>(one, 1, partials + (int64_t)b * pblocks * kBCols, 0);
    }
}

// V_b -= sum_m Q_b[m] z_b[m] and partial dots Q_b[m]^H V_b (m < C).  C is a
// template argument so that all C loads of an element are issued together
// (a runtime count serialises them behind branches); the arithmetic order
// is that of the single-member k_update_dots.
template <int C>
__global__ __launch_bounds__(kBThreads) void kb_update_dots(const zd *__restrict__ Q, int64_t qstride, int64_t N,
                                                            const zd *__restrict__ z, int zstride,
                                                            zd *__restrict__ V, const int32_t *active,
                                                            zd *partials, int pblocks) {
    const int b = blockIdx.y;
    if (!active[b]) return;
    __shared__ zd zs[C];
    if ((int)threadIdx.x < C) zs[threadIdx.x] = z[(int64_t)b * zstride + threadIdx.x];
    __syncthreads();
    const zd *Qb = Q + qstride * b;
    zd *Vb = V + N * b;
    zd acc[C];
#pragma unroll
    for (int m = 0; m < C; ++m) acc[m] = zmake(0.0, 0.0);
    for (int64_t x = blockIdx.x * (int64_t)kBThreads + threadIdx.x; x < N; x += (int64_t)gridDim.x * kBThreads) {
        zd v = Vb[x];
        zd q[C];
#pragma unroll
        for (int m = 0; m < C; ++m) q[m] = Qb[m * N + x];
#pragma unroll
        for (int m = 0; m < C; ++m) v = zsub(v, zmul(q[m], zs[m]));
#pragma unroll
        for (int m = 0; m < C; ++m) acc[m] = zcfma(q[m], v, acc[m]);
        st_stream(Vb + x, v);
    }
    bred_store<C>(acc, C, partials + (int64_t)b * pblocks * kBCols, 0);
}

typedef void (*BatchUpdateDotsKernel)(const zd *, int64_t, int64_t, const zd *, int, zd *, const int32_t *, zd *,
                                      int);

template <int... I>
static std::array<BatchUpdateDotsKernel, sizeof...(I)> batch_update_dots_table(std::integer_sequence<int, I...>) {
    return {{&kb_update_dots<I + 1>...}};
}

// kb_update_dots<count> for count in [1, kBChunk]
static BatchUpdateDotsKernel batch_update_dots_kernel(int count) {
    static const auto table = batch_update_dots_table(std::make_integer_sequence<int, kBChunk>{});
    return table[count - 1];
}

typedef void (*BatchDotsKernel)(const zd *, int64_t, int64_t, int, const zd *, const int32_t *, zd *, int);

template <int... I>
static std::array<BatchDotsKernel, sizeof...(I)> batch_dots_table(std::integer_sequence<int, I...>) {
    return {{&kb_dots<I + 1>...}};
}

// kb_dots<count> for count in [1, kBChunk]
static BatchDotsKernel batch_dots_kernel(int count) {
    static const auto table = batch_dots_table(std::make_integer_sequence<int, kBChunk>{});
    return table[count - 1];
}

typedef void (*BatchUpdateKernel)(const zd *, int64_t, int64_t, int, const zd *, int, zd *, const int32_t *, zd *,
                                  int);

template <bool NORM, int... I>
static std::array<BatchUpdateKernel, sizeof...(I)> batch_update_table(std::integer_sequence<int, I...>) {
    return {{&kb_update<I, NORM>...}};
}

// kb_update<count % kBChunk, NORM>, launched with full = count / kBChunk
static BatchUpdateKernel batch_update_kernel(int count, bool norm) {
    static const auto plain = batch_update_table<false>(std::make_integer_sequence<int, kBChunk>{});
    static const auto with_norm = batch_update_table<true>(std::make_integer_sequence<int, kBChunk>{});
    return norm ? with_norm[count % kBChunk] : plain[count % kBChunk];
}

// out_b[col] = sum over the member's blocks (mode 1: sqrt of the real sum into
// a double); grid (count, B), fixed-order strided sums and block reduction.
__global__ __launch_bounds__(kBThreads) void kb_finish(const zd *partials, int pblocks, int col0, int count,
                                                       int mode, const int32_t *active, zd *out, int ostride,
                                                       double *dout) {
    const int col = blockIdx.x, b = blockIdx.y;
    if (!active[b]) return;
    const zd *pb = partials + (int64_t)b * pblocks * kBCols;
    double re = 0.0, im = 0.0;
    for (int k = threadIdx.x; k < pblocks; k += kBThreads) {
        const zd p = pb[(int64_t)k * kBCols + col0 + col];
        re += p.x;
        im += p.y;
    }
    __shared__ double red[kBThreads / 64][2];
    re = wave_sum(re);
    im = wave_sum(im);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[wave][0] = re;
        red[wave][1] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < kBThreads / 64; ++w) {
            x += red[w][0];
            y += red[w][1];
        }
        if (mode == 1)
            dout[b] = sqrt(x);
        else
            out[(int64_t)b * ostride + col0 + col] = zmake(x, y);
    }
    (void)count;
}

// y_b = x_b / s_b  (members with active mask; zero norm -> member converged trivially)
__global__ void kb_scale(const zd *__restrict__ X, int64_t xstride, const double *s, zd *__restrict__ Y,
                         int64_t ystride, int64_t N, const int32_t *active) {
    const int b = blockIdx.y;
    if (!active[b] || s[b] == 0.0) return;
    const double d = s[b];
    const zd *x = X + xstride * b;
    zd *y = Y + ystride * b;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const zd a = x[i];
        st_stream(y + i, zmake(a.x / d, a.y / d));
    }
}

// zero initial norm: the step maps 0 -> 0 (python/torch_backend.py:143)
__global__ void kb_zero_norm(int B, const double *beta, int32_t *ctl) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || !ctl_at(ctl, B, kActive, b)) return;
    if (beta[b] == 0.0) {
        ctl_at(ctl, B, kActive, b) = 0;
        ctl_at(ctl, B, kDone, b) = 1;
        ctl_at(ctl, B, kDim, b) = 0;
    }
}

// Hessenberg column, happy breakdown / convergence check with a device expm.
__global__ void kb_check(int B, int j, int maxdim, const double *cur, const zd *z, const double *nrm,
                         const double *beta, zd *Hm, zd *coef, int32_t *ctl) {
    const int b = blockIdx.x;
    if (!ctl_at(ctl, B, kActive, b)) return;
    const double dt = cur[3 * b + 1], tol = cur[3 * b + 2];
    __shared__ zd work[8 * kBatchMaxDim * kBatchMaxDim];
    zd *H = Hm + (int64_t)b * (maxdim + 1) * maxdim;   // row-major [maxdim+1][maxdim]
    const zd *zb = z + (int64_t)b * 2 * kBCols;
    for (int r = threadIdx.x; r <= j; r += blockDim.x) H[r * maxdim + j] = zadd(zb[r], zb[kBCols + r]);
    if (threadIdx.x == 0) H[(j + 1) * maxdim + j] = zmake(nrm[b], 0.0);
    __syncthreads();
    const double h = nrm[b];
    const bool happy = h < 1e-13;
    const int m = j + 1;
    if (!(happy || (j >= 3 && ((j + 1) % 4 == 0 || j + 1 == maxdim)))) return;   // uniform over the block
    zd *A = work, *E = work + kBatchMaxDim * kBatchMaxDim, *W = work + 2 * kBatchMaxDim * kBatchMaxDim;
    // -i dt H_m; entries below the subdiagonal are never written (Hessenberg)
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int r = e / m, c = e - r * m;
        const zd x = r <= c + 1 ? H[r * maxdim + c] : zmake(0.0, 0.0);
        A[e] = zmake(dt * x.y, -dt * x.x);
    }
    __syncthreads();
    hexw::expm(m, A, E, W);   // the whole wave; same result as the serial hexm::expm
    if (threadIdx.x != 0) return;
    const double bt = beta[b];
    const zd last = zscale(E[(m - 1) * m], bt);
    const double err = fabs(dt * h) * sqrt(zabs2(last));
    if (happy || err < tol) {
        for (int r = 0; r < m; ++r) coef[(int64_t)b * maxdim + r] = zscale(E[r * m], bt);
        ctl_at(ctl, B, kDim, b) = m;
        ctl_at(ctl, B, kDone, b) = 1;
        ctl_at(ctl, B, kActive, b) = 0;
    } else if (j + 1 == maxdim) {
        ctl_at(ctl, B, kSplit, b) = 1;   // not converged: halve the interval
        ctl_at(ctl, B, kActive, b) = 0;
    }
}

// psi_b = sum_{m < dim_b} Q_b[m] coef_b[m] for members that converged this step
__global__ void kb_combine(const zd *__restrict__ Q, int64_t qstride, int64_t N, int maxdim, const zd *coef,
                           const int32_t *ctl, int B, zd *__restrict__ psi) {
    const int b = blockIdx.y;
    if (!ctl_at(const_cast<int32_t *>(ctl), B, kDone, b)) return;
    const int m = ctl_at(const_cast<int32_t *>(ctl), B, kDim, b);
    if (m == 0) return;
    const zd *Qb = Q + qstride * b;
    zd *pb = psi + N * b;
    const zd *cb = coef + (int64_t)b * maxdim;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        zd acc = zmake(0.0, 0.0);
        for (int k = 0; k < m; ++k) acc = zfma(Qb[k * N + i], cb[k], acc);
        st_stream(pb + i, acc);
    }
}

// ---------------------------------------------------------------------------
static int record_step(hex_batch *p) {
    hipStream_t st = p->stream;
    const int B = p->B, md = p->maxdim;
    const int64_t N = p->N, qstride = (int64_t)(md + 1) * N;
    const int g = p->red_blocks;
    int32_t *active = p->ctl + kActive * B;
    const int sgrid = (int)std::min<int64_t>((N + 255) / 256, 2048);
    hipLaunchKernelGGL(kb_begin, dim3((B + 63) / 64), dim3(64), 0, st, B, p->dt, p->tol, p->ctl, p->counts, p->t0,
                       p->pulses, p->stk, p->sdep, p->cur, p->fields);
    // beta_b = ||psi_b||, Q_b[0] = psi_b / beta_b
    hipLaunchKernelGGL(batch_update_kernel(0, true), dim3(g, B), dim3(kBThreads), 0, st, p->Q, qstride, N, 0,
                       p->z, 2 * kBCols, p->psi, active, p->partials, g);
    hipLaunchKernelGGL(kb_finish, dim3(1, B), dim3(kBThreads), 0, st, p->partials, g, 0, 1, 1, active, p->z,
                       2 * kBCols, p->beta);
    hipLaunchKernelGGL(kb_zero_norm, dim3((B + 63) / 64), dim3(64), 0, st, B, p->beta, p->ctl);
    hipLaunchKernelGGL(kb_scale, dim3(sgrid, B), dim3(256), 0, st, p->psi, N, p->beta, p->Q, qstride, N, active);
    const BatchArgs bt{p->fields, active, qstride, N};
    for (int j = 0; j < md; ++j) {
        const int k = j + 1;
        // V_b = H(A_b) Q_b[j]  (velocity gauge)
        int rc = hex_launch_apply2(p->e->v, p->by, nullptr, 1, 0, zmake(1.0, 0.0), zmake(0.0, 0.0), bt, B,
                                   p->Q + (int64_t)j * N, p->V, p->y1, p->y2, p->dcz, st, kTermAll);
        if (rc) return rc;
        // CGS2 pass 1: z1 = Q^H V
        for (int c0 = 0; c0 < k; c0 += kBChunk) {
            const int cnt = std::min(kBChunk, k - c0);
            hipLaunchKernelGGL(batch_dots_kernel(cnt), dim3(g, B), dim3(kBThreads), 0, st, p->Q, qstride, N, c0,
                               p->V, active, p->partials, g);
        }
        hipLaunchKernelGGL(kb_finish, dim3(k, B), dim3(kBThreads), 0, st, p->partials, g, 0, k, 0, active, p->z,
                           2 * kBCols, p->nrm);
        // pass 2: V -= Q z1, z2 = Q^H V (fused when k <= chunk)
        if (k <= kBChunk) {
            hipLaunchKernelGGL(batch_update_dots_kernel(k), dim3(g, B), dim3(kBThreads), 0, st, p->Q, qstride, N,
                               p->z, 2 * kBCols, p->V, active, p->partials, g);
        } else {
            hipLaunchKernelGGL(batch_update_kernel(k, false), dim3(g, B), dim3(kBThreads), 0, st, p->Q, qstride,
                               N, k / kBChunk, p->z, 2 * kBCols, p->V, active, p->partials, g);
            for (int c0 = 0; c0 < k; c0 += kBChunk) {
                const int cnt = std::min(kBChunk, k - c0);
                hipLaunchKernelGGL(batch_dots_kernel(cnt), dim3(g, B), dim3(kBThreads), 0, st, p->Q, qstride, N,
                                   c0, p->V, active, p->partials, g);
            }
        }
        hipLaunchKernelGGL(kb_finish, dim3(k, B), dim3(kBThreads), 0, st, p->partials, g, 0, k, 0, active,
                           p->z + kBCols, 2 * kBCols, p->nrm);
        // V -= Q z2, ||V||
        hipLaunchKernelGGL(batch_update_kernel(k, true), dim3(g, B), dim3(kBThreads), 0, st, p->Q, qstride, N,
                           k / kBChunk, p->z + kBCols, 2 * kBCols, p->V, active, p->partials, g);
        hipLaunchKernelGGL(kb_finish, dim3(1, B), dim3(kBThreads), 0, st, p->partials, g, 0, 1, 1, active, p->z,
                           2 * kBCols, p->nrm);
        hipLaunchKernelGGL(kb_check, dim3(B), dim3(64), 0, st, B, j, md, p->cur, p->z, p->nrm, p->beta, p->Hm,
                           p->coef, p->ctl);
        if (j + 1 < md)
            hipLaunchKernelGGL(kb_scale, dim3(sgrid, B), dim3(256), 0, st, p->V, N, p->nrm, p->Q + (int64_t)k * N,
                               qstride, N, active);
    }
    hipLaunchKernelGGL(kb_combine, dim3(sgrid, B), dim3(256), 0, st, p->Q, qstride, N, md, p->coef, p->ctl, B,
                       p->psi);
    hipLaunchKernelGGL(kb_end, dim3((B + 63) / 64), dim3(64), 0, st, B, p->ctl, p->stk, p->sdep);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
template <typename T>
static int balloc(hex_batch *p, size_t count, T **out) {
    void *q = nullptr;
    HEX_HIP(hipMalloc(&q, count * sizeof(T)));
    p->owned.push_back(q);
    *out = static_cast<T *>(q);
    return HEX_OK;
}

extern "C" int hex_batch_destroy(hex_batch *p) {
    if (!p) return HEX_OK;
    (void)hipSetDevice(p->device);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    for (void *q : p->owned) (void)hipFree(q);
    delete p;
    return HEX_OK;
}

static int build_graph(hex_batch *p) {
    if (p->exec) {
        HEX_HIP(hipGraphExecDestroy(p->exec));
        p->exec = nullptr;
    }
    if (p->graph) {
        HEX_HIP(hipGraphDestroy(p->graph));
        p->graph = nullptr;
    }
    HEX_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    int rc = record_step(p);
    hipGraph_t g = nullptr;
    hipError_t end = hipStreamEndCapture(p->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (end != hipSuccess) {
        hex::set_error(std::string("batch graph capture: ") + hipGetErrorString(end));
        return HEX_EDEVICE;
    }
    p->graph = g;
    HEX_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return HEX_OK;
}

extern "C" int hex_batch_create(hex_engine *e, int B, int maxdim, double dt, double tol, int npulses,
                                const double *pulses, const double *t0, const int32_t *counts, double *psi,
                                hex_batch **out) {
    if (!out || !e || B < 1 || maxdim < 4 || maxdim > kBatchMaxDim || maxdim + 1 > kBCols || npulses < 0 ||
        npulses > kMaxPulses || (npulses && !pulses) || !t0 || !counts || !psi || !(dt > 0) || !(tol > 0)) {
        hex::set_error("batch: need B >= 1, 4 <= maxdim <= 16, 0 <= pulses per member <= 4, dt > 0, tol > 0");
        return HEX_EINVAL;
    }
    for (int b = 0; b < B; ++b)
        if (counts[b] < 0) {
            hex::set_error("batch: member step counts must be >= 0");
            return HEX_EINVAL;
        }
    std::vector<double> pk((size_t)B * kMaxPulses * 7, 0.0);
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < npulses; ++k) {
            const double *src = pulses + ((size_t)b * npulses + k) * 6;
            double *dst = pk.data() + ((size_t)b * kMaxPulses + k) * 7;
            if (!(src[0] > 0) || !(src[1] > 1) || src[2] < 0 || src[3] < 0 || src[5] < 0 || src[5] > 2) {
                hex::set_error("batch: pulse needs omega > 0, cycles > 1, field >= 0, start >= 0, polarisation 0..2");
                return HEX_EINVAL;
            }
            for (int q = 0; q < 6; ++q) dst[q] = src[q];
            dst[6] = src[2] != 0.0 ? 1.0 : 0.0;
        }
    *out = nullptr;
    HEX_HIP(hipSetDevice(e->device));
    auto *p = new hex_batch();
    p->e = e;
    p->device = e->device;
    p->B = B;
    p->maxdim = maxdim;
    p->dt = dt;
    p->tol = tol;
    p->by = e->apply_by;
    p->N = (int64_t)e->v.n * e->v.n * e->v.nc;
    p->psi = reinterpret_cast<zd *>(psi);
    const int64_t N = p->N;
    p->red_blocks = (int)std::max<int64_t>(1, std::min<int64_t>((N + 4 * kBThreads - 1) / (4 * kBThreads), kBBlocks));
    int rc = HEX_OK;
    auto fail = [&](int code) {
        hex_batch_destroy(p);
        return code;
    };
    if ((rc = balloc(p, (size_t)B * (maxdim + 1) * N, &p->Q)) || (rc = balloc(p, (size_t)B * N, &p->V)) ||
        (rc = balloc(p, (size_t)B * N, &p->y1)) || (rc = balloc(p, (size_t)B * N, &p->y2)) ||
        (rc = balloc(p, (size_t)B * std::max(e->v.nd, 1), &p->dcz)) ||
        (rc = balloc(p, (size_t)B * 3, &p->fields)) || (rc = balloc(p, pk.size(), &p->pulses)) ||
        (rc = balloc(p, (size_t)B, &p->t0)) || (rc = balloc(p, (size_t)B * 3, &p->cur)) ||
        (rc = balloc(p, (size_t)B * kStack * 3, &p->stk)) || (rc = balloc(p, (size_t)B * kStack, &p->sdep)) ||
        (rc = balloc(p, (size_t)kCtl * B, &p->ctl)) || (rc = balloc(p, (size_t)B, &p->counts)) ||
        (rc = balloc(p, (size_t)B * p->red_blocks * kBCols, &p->partials)) ||
        (rc = balloc(p, (size_t)B * 2 * kBCols, &p->z)) || (rc = balloc(p, (size_t)B, &p->nrm)) ||
        (rc = balloc(p, (size_t)B, &p->beta)) || (rc = balloc(p, (size_t)B * (maxdim + 1) * maxdim, &p->Hm)) ||
        (rc = balloc(p, (size_t)B * maxdim, &p->coef)))
        return fail(rc);
    std::vector<int32_t> ctl((size_t)kCtl * B, 0);
    for (int b = 0; b < B; ++b) ctl[kAlive * B + b] = counts[b] > 0;
    if (hipMemcpy(p->pulses, pk.data(), sizeof(double) * pk.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->t0, t0, sizeof(double) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->counts, counts, sizeof(int32_t) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->ctl, ctl.data(), sizeof(int32_t) * kCtl * B, hipMemcpyHostToDevice) != hipSuccess) {
        hex::set_error("batch: upload failed");
        return fail(HEX_EDEVICE);
    }
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        hex::set_error("batch: stream creation failed");
        return fail(HEX_EDEVICE);
    }
    if ((rc = build_graph(p))) return fail(rc);
    *out = p;
    return HEX_OK;
}

extern "C" int hex_batch_set_launch(hex_batch *p, int rows_per_block) {
    if (!p || (rows_per_block != 0 && rows_per_block != 1 && rows_per_block != 2 && rows_per_block != 4 &&
               rows_per_block != 8 && rows_per_block != kLaunchMfma)) {
        hex::set_error("batch set_launch: launch shape must be 0 (staged), 1, 2, 4, 8 rows per block or 16 (matrix cores)");
        return HEX_EINVAL;
    }
    p->by = rows_per_block;
    return build_graph(p);
}

// Replay the attempt graph `steps` times on `stream` (NULL: the batch's own
// stream), ordered after the work already queued on it.
extern "C" int hex_batch_run(hex_batch *p, int steps, void *stream) {
    if (!p || steps < 0) {
        hex::set_error("batch run: bad argument");
        return HEX_EINVAL;
    }
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : p->stream;
    for (int s = 0; s < steps; ++s) HEX_HIP(hipGraphLaunch(p->exec, st));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// Per-member control state (host int32[10][B]): alive, active, converged in the
// last attempt, its Krylov dimension, failures, completed steps, split in the
// last attempt, pending intervals, attempts, accepted (sub)steps.  Synchronises.
extern "C" int hex_batch_status(hex_batch *p, int32_t *out, void *stream) {
    if (!p || !out) {
        hex::set_error("batch status: bad argument");
        return HEX_EINVAL;
    }
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : p->stream;
    HEX_HIP(hipMemcpyAsync(out, p->ctl, sizeof(int32_t) * kCtl * p->B, hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    return HEX_OK;
}

extern "C" int64_t hex_batch_bytes(const hex_batch *p) {
    if (!p) return 0;
    const int64_t N = p->N;
    return 16 * ((int64_t)p->B * (p->maxdim + 4) * N);
}
