// Krylov building blocks: CGS2, norms, basis combinations.
//
// Every reduction is two-stage and deterministic: a fixed grid of blocks
// walks the vectors with a grid-stride loop, each block reduces its lanes in a
// fixed butterfly order into partials[block][column], and one finishing block
// sums the partials in block order.  Results therefore do not depend on timing.
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include "hex_expm.h"

#include <algorithm>
#include <array>
#include <mutex>
#include <utility>
#include <vector>

static constexpr int kRedThreads = 256;
static constexpr int kRedBlocks = 1024;  // must match the scratch size in hex_engine.hip
static constexpr int kRedCols = 64;
static constexpr int kChunk = 16;        // basis vectors held in registers by fused kernels

struct hex_engine_scratch {
    zd *partials;
    zd *zbuf;
    zd *hbuf_host;
};
hex_engine_scratch hex_scratch_of(hex_engine *e);

struct Coeffs {
    zd c[kRedCols];
};

static int grid_for(int64_t N) {
    int64_t g = (N + kRedThreads * 4 - 1) / (kRedThreads * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, kRedBlocks));
}

// Reduce `count` complex accumulators over the block; thread 0..count-1 write
// partials[blockIdx.x * kRedCols + col0 + t].
template <int KC>
__device__ __forceinline__ void block_reduce_store(zd (&acc)[KC], int count, zd *partials, int col0) {
    __shared__ double red[kRedThreads / 64][KC][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < KC; ++m) {
        if (m < count) {
            double re = wave_sum(acc[m].x), im = wave_sum(acc[m].y);
            if (lane == 0) {
                red[wave][m][0] = re;
                red[wave][m][1] = im;
            }
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < count) {
        double re = 0.0, im = 0.0;
        for (int w = 0; w < kRedThreads / 64; ++w) {
            re += red[w][threadIdx.x][0];
            im += red[w][threadIdx.x][1];
        }
        partials[(int64_t)blockIdx.x * kRedCols + col0 + threadIdx.x] = zmake(re, im);
    }
}

// partial z[m] = sum_x conj(Q_m[x]) v[x] for m in [c0, c0+C).  C is a template
// argument so that the C loads of an element are issued together.
template <int C>
__global__ __launch_bounds__(kRedThreads) void k_dots(const zd *__restrict__ Q, int64_t ldq, int c0,
                                                      const zd *__restrict__ v, int64_t N, zd *partials) {
    const zd *Qc = Q + (int64_t)c0 * ldq;
    zd acc[C];
#pragma unroll
    for (int m = 0; m < C; ++m) acc[m] = zmake(0.0, 0.0);
    for (int64_t x = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; x < N; x += (int64_t)gridDim.x * kRedThreads) {
        const zd vx = v[x];
        zd q[C];
#pragma unroll
        for (int m = 0; m < C; ++m) q[m] = Qc[m * ldq + x];
#pragma unroll
        for (int m = 0; m < C; ++m) acc[m] = zcfma(q[m], vx, acc[m]);
    }
    block_reduce_store<C>(acc, C, partials, c0);
}

// v -= sum_m Q_m z[m] for m < full*kChunk + R, in m order; optionally partial
// ||v||^2 into partials[.][0].  The basis is walked in chunks of kChunk
// vectors whose loads are issued together; R, the last partial chunk, is a
// template argument for the same reason.
template <int R, bool NORM>
__global__ __launch_bounds__(kRedThreads) void k_update(const zd *__restrict__ Q, int64_t ldq, int full,
                                                        const zd *__restrict__ z, zd *__restrict__ v, int64_t N,
                                                        zd *partials) {
    __shared__ zd zs[kRedCols];
    const int count = full * kChunk + R;
    if ((int)threadIdx.x < count) zs[threadIdx.x] = z[threadIdx.x];
    __syncthreads();
    double nrm = 0.0;
    for (int64_t x = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; x < N; x += (int64_t)gridDim.x * kRedThreads) {
        zd vx = v[x];
        for (int c = 0; c < full; ++c) {
            const zd *Qc = Q + (int64_t)c * kChunk * ldq;
            zd q[kChunk];
#pragma unroll
            for (int m = 0; m < kChunk; ++m) q[m] = Qc[m * ldq + x];
#pragma unroll
            for (int m = 0; m < kChunk; ++m) vx = zsub(vx, zmul(q[m], zs[c * kChunk + m]));
        }
        if constexpr (R > 0) {
            const zd *Qc = Q + (int64_t)full * kChunk * ldq;
            zd q[R];
#pragma unroll
            for (int m = 0; m < R; ++m) q[m] = Qc[m * ldq + x];
#pragma unroll
            for (int m = 0; m < R; ++m) vx = zsub(vx, zmul(q[m], zs[full * kChunk + m]));
        }
        if (count > 0) st_stream(v + x, vx);
        if (NORM) nrm += zabs2(vx);
    }
    if (NORM) {
        zd one[1] = {zmake(nrm, 0.0)};
        block_reduce_store<1>(one, 1, partials, 0);
    }
}

// v -= sum_m Q_m z[m] and partial dots Q_m^H v_new into partials[.][0..C),
// optionally partial ||v_new||^2 into partials[.][C].  The count C is a
// template argument: with a runtime count every load of Q sits behind its own
// branch and is issued only after the previous one, which halves the
// bandwidth of this kernel; with C fixed all C loads of an element are in
// flight together.  The arithmetic (update chain in m order, then the dots)
// is the same for every C.
template <int C, bool NORM>
__global__ __launch_bounds__(kRedThreads) void k_update_dots(const zd *__restrict__ Q, int64_t ldq,
                                                             const zd *__restrict__ z, zd *__restrict__ v,
                                                             int64_t N, zd *partials) {
    __shared__ zd zs[C];
    if ((int)threadIdx.x < C) zs[threadIdx.x] = z[threadIdx.x];
    __syncthreads();
    zd acc[C];
#pragma unroll
    for (int m = 0; m < C; ++m) acc[m] = zmake(0.0, 0.0);
    double nrm = 0.0;
    for (int64_t x = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; x < N; x += (int64_t)gridDim.x * kRedThreads) {
        zd vx = v[x];
        zd q[C];
#pragma unroll
        for (int m = 0; m < C; ++m) q[m] = Q[m * ldq + x];
#pragma unroll
        for (int m = 0; m < C; ++m) vx = zsub(vx, zmul(q[m], zs[m]));
#pragma unroll
        for (int m = 0; m < C; ++m) acc[m] = zcfma(q[m], vx, acc[m]);
        st_stream(v + x, vx);
        if (NORM) nrm += zabs2(vx);
    }
    block_reduce_store<C>(acc, C, partials, 0);
    if (NORM) {
        __syncthreads();
        zd one[1] = {zmake(nrm, 0.0)};
        block_reduce_store<1>(one, 1, partials, C);
    }
}

typedef void (*UpdateDotsKernel)(const zd *, int64_t, const zd *, zd *, int64_t, zd *);

template <bool NORM, int... I>
static std::array<UpdateDotsKernel, sizeof...(I)> update_dots_table(std::integer_sequence<int, I...>) {
    return {{&k_update_dots<I + 1, NORM>...}};
}

// k_update_dots<count, NORM> for count in [1, kChunk]
static UpdateDotsKernel update_dots_kernel(int count, bool norm) {
    static const auto plain = update_dots_table<false>(std::make_integer_sequence<int, kChunk>{});
    static const auto with_norm = update_dots_table<true>(std::make_integer_sequence<int, kChunk>{});
    return norm ? with_norm[count - 1] : plain[count - 1];
}

typedef void (*DotsKernel)(const zd *, int64_t, int, const zd *, int64_t, zd *);

template <int... I>
static std::array<DotsKernel, sizeof...(I)> dots_table(std::integer_sequence<int, I...>) {
    return {{&k_dots<I + 1>...}};
}

// k_dots<count> for count in [1, kChunk]
static DotsKernel dots_kernel(int count) {
    static const auto table = dots_table(std::make_integer_sequence<int, kChunk>{});
    return table[count - 1];
}

typedef void (*UpdateKernel)(const zd *, int64_t, int, const zd *, zd *, int64_t, zd *);

template <bool NORM, int... I>
static std::array<UpdateKernel, sizeof...(I)> update_table(std::integer_sequence<int, I...>) {
    return {{&k_update<I, NORM>...}};
}

// k_update<count % kChunk, NORM>, launched with full = count / kChunk
static UpdateKernel update_kernel(int count, bool norm) {
    static const auto plain = update_table<false>(std::make_integer_sequence<int, kChunk>{});
    static const auto with_norm = update_table<true>(std::make_integer_sequence<int, kChunk>{});
    return norm ? with_norm[count % kChunk] : plain[count % kChunk];
}

// out[col] = sum_b partials[b][col]: one block per column, threads stride the
// partials in a fixed pattern, then a fixed-order block reduction.
// mode 0: out = sum; mode 1: out = (sqrt(sum.re), 0)
__global__ __launch_bounds__(kRedThreads) void k_finish(const zd *partials, int blocks, int count, int mode,
                                                        zd *out) {
    const int col = blockIdx.x;
    double re = 0.0, im = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kRedThreads) {
        const zd p = partials[(int64_t)b * kRedCols + col];
        re += p.x;
        im += p.y;
    }
    __shared__ double red[kRedThreads / 64][2];
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
        for (int w = 0; w < kRedThreads / 64; ++w) {
            x += red[w][0];
            y += red[w][1];
        }
        out[col] = mode == 1 ? zmake(sqrt(x), 0.0) : zmake(x, y);
    }
    (void)count;
}

__global__ void k_scale(const zd *__restrict__ x, double s, zd *__restrict__ y, int64_t N) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const zd a = x[i];
        st_stream(y + i, zmake(a.x / s, a.y / s));
    }
}

__global__ void k_combine(const zd *__restrict__ Q, int64_t ldq, int count, Coeffs c, zd *__restrict__ y,
                          int64_t N) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        zd acc = zmake(0.0, 0.0);
        for (int m = 0; m < count; ++m) acc = zfma(Q[m * ldq + i], c.c[m], acc);
        st_stream(y + i, acc);
    }
}

// out = a x + b y  (complex a, b); out may alias x or y
__global__ void k_axpby(int64_t N, zd a, const zd *x, zd b, const zd *y, zd *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        st_stream(out + i, zadd(zmul(a, x[i]), zmul(b, y[i])));
}

extern "C" int hex_axpby(int64_t N, const double a[2], const double *x, const double b[2], const double *y,
                         double *out, void *stream) {
    if (N < 1 || !a || !b || !x || !y || !out) {
        hex::set_error("axpby: bad argument");
        return HEX_EINVAL;
    }
    const int g = (int)std::min<int64_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, static_cast<hipStream_t>(stream), N, zmake(a[0], a[1]),
                       reinterpret_cast<const zd *>(x), zmake(b[0], b[1]), reinterpret_cast<const zd *>(y),
                       reinterpret_cast<zd *>(out));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// ---------------------------------------------------------------------------
static int check_basis(const void *Q, int64_t ldq, int k, int64_t N, const char *who) {
    if (k < 0 || k > kRedCols - 1 || N < 1 || (k > 0 && (!Q || ldq < N))) {
        hex::set_error(std::string(who) + ": basis must hold 0..63 vectors with ldq >= N");
        return HEX_EINVAL;
    }
    return HEX_OK;
}

// z (device, count) = Q^H v over [0, count) into dst
static int dots_to(hex_engine_scratch s, const zd *Q, int64_t ldq, int count, const zd *v, int64_t N, zd *dst,
                   hipStream_t st) {
    const int g = grid_for(N);
    for (int c0 = 0; c0 < count; c0 += kChunk) {
        const int cnt = std::min(kChunk, count - c0);
        hipLaunchKernelGGL(dots_kernel(cnt), dim3(g), dim3(kRedThreads), 0, st, Q, ldq, c0, v, N, s.partials);
    }
    hipLaunchKernelGGL(k_finish, dim3(count), dim3(kRedThreads), 0, st, s.partials, g, count, 0, dst);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// v -= Q z (z device); if want_dots (count <= kChunk) also dst_dots = Q^H v_new;
// if want_norm dst_norm = ||v_new||.
static int update_to(hex_engine_scratch s, const zd *Q, int64_t ldq, int count, const zd *z, zd *v, int64_t N,
                     bool want_dots, zd *dst_dots, bool want_norm, zd *dst_norm, hipStream_t st) {
    const int g = grid_for(N);
    if (want_dots && count >= 1 && count <= kChunk) {
        hipLaunchKernelGGL(update_dots_kernel(count, want_norm), dim3(g), dim3(kRedThreads), 0, st, Q, ldq, z, v, N,
                           s.partials);
        hipLaunchKernelGGL(k_finish, dim3(count), dim3(kRedThreads), 0, st, s.partials, g, count, 0, dst_dots);
        if (want_norm)
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, st, s.partials + count, g, 1, 1, dst_norm);
    } else {
        hipLaunchKernelGGL(update_kernel(count, want_norm), dim3(g), dim3(kRedThreads), 0, st, Q, ldq,
                           count / kChunk, z, v, N, s.partials);
        if (want_norm)
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, st, s.partials, g, 1, 1, dst_norm);
        if (want_dots) {
            int rc = dots_to(s, Q, ldq, count, v, N, dst_dots, st);
            if (rc) return rc;
        }
    }
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// Both CGS2 passes of v against Q[0 .. k), leaving z1 in zbuf[0, 64), z2 in
// zbuf[64, 128) and ||v_new|| right after z2 at zbuf[64 + k].
static int cgs2_passes(hex_engine_scratch s, const zd *Q, int64_t ldq, int k, zd *v, int64_t N, hipStream_t st) {
    zd *z1 = s.zbuf, *z2 = s.zbuf + kRedCols, *nrm = s.zbuf + kRedCols + k;
    if (k == 0) return update_to(s, Q, ldq, 0, z1, v, N, false, nullptr, true, nrm, st);
    int rc = dots_to(s, Q, ldq, k, v, N, z1, st);
    if (!rc) rc = update_to(s, Q, ldq, k, z1, v, N, true, z2, false, nullptr, st);
    if (!rc) rc = update_to(s, Q, ldq, k, z2, v, N, false, nullptr, true, nrm, st);
    return rc;
}

// h[m] = (0 + z1[m]) + z2[m] for m < k and h[k] = ||v||, as the host forms
// the Hessenberg column from the same values.
__global__ void k_cgs2_commit(const zd *__restrict__ zbuf, int k, zd *__restrict__ h) {
    const int m = threadIdx.x;
    if (m < k) h[m] = zadd(zbuf[m], zbuf[kRedCols + m]);
    else if (m == k) h[k] = zbuf[kRedCols + k];
}

// y = x / s->x, or y = x when s->x < tiny (a breakdown column: the host
// discards it, and no infinity enters the basis)
__global__ void k_scale_by(const zd *__restrict__ x, const zd *__restrict__ s, double tiny, zd *__restrict__ y,
                           int64_t N) {
    double d = s->x;
    if (!(d >= tiny)) d = 1.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const zd a = x[i];
        st_stream(y + i, zmake(a.x / d, a.y / d));
    }
}

extern "C" int hex_cgs2_commit(hex_engine *e, const double *Qd, int64_t ldq, int k, double *vd, int64_t N,
                               double *hdev, double *qnext, void *stream) {
    if (!e || !vd || !hdev || !qnext) {
        hex::set_error("cgs2_commit: null argument");
        return HEX_EINVAL;
    }
    int rc = check_basis(Qd, ldq, k, N, "cgs2_commit");
    if (rc) return rc;
    hex_engine_scratch s = hex_scratch_of(e);
    hipStream_t st = static_cast<hipStream_t>(stream);
    zd *v = reinterpret_cast<zd *>(vd);
    if ((rc = cgs2_passes(s, reinterpret_cast<const zd *>(Qd), ldq, k, v, N, st))) return rc;
    zd *h = reinterpret_cast<zd *>(hdev);
    hipLaunchKernelGGL(k_cgs2_commit, dim3(1), dim3(kRedCols), 0, st, s.zbuf, k, h);
    const int g = (int)std::min<int64_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_scale_by, dim3(g), dim3(256), 0, st, v, h + k, 1e-14, reinterpret_cast<zd *>(qnext), N);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// hex_cgs2_commit without an engine and without allocation.  Its scratch --
// the per-block partial sums (kRedBlocks x kRedCols) and the two coefficient
// vectors plus the norm (2 kRedCols) -- is this file-scope device array, held
// by every GPU the library runs on until the process ends.  The partials are
// written by every block before the finishing pass reads them and the
// coefficient vectors before they are used, so nothing needs clearing between
// calls; calls that overlap on one GPU share the array and race.
static constexpr int64_t kNoAllocScratch = (int64_t)kRedBlocks * kRedCols + 2 * kRedCols;   // 1 MiB + 2 KiB
static __device__ __align__(256) zd scratch[kNoAllocScratch];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_lock;
static zd *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (the host-side symbol is
// only a stand-in): resolved once per device with that device current.
static int scratch_on_current_device(zd **out) {
    int dev = 0;
    HEX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        hex::set_error("cgs2_commit_noalloc: device index beyond the library's table of 64");
        return HEX_EINVAL;
    }
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        HEX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<zd *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return HEX_OK;
}

extern "C" int hex_cgs2_commit_noalloc(const double *Qd, int64_t ldq, int k, double *vd, int64_t N, double *hdev,
                                       double *qnext, void *stream) {
    if (!vd || !hdev || !qnext) {
        hex::set_error("cgs2_commit_noalloc: null argument");
        return HEX_EINVAL;
    }
    if (k > kRedCols - 1) {
        hex::set_error("cgs2_commit_noalloc: at most 63 basis vectors fit its fixed scratch");
        return HEX_EINVAL;
    }
    int rc = check_basis(Qd, ldq, k, N, "cgs2_commit_noalloc");
    if (rc) return rc;
    zd *base = nullptr;
    if ((rc = scratch_on_current_device(&base))) return rc;
    hex_engine_scratch s{base, base + (int64_t)kRedBlocks * kRedCols, nullptr};
    hipStream_t st = static_cast<hipStream_t>(stream);
    zd *v = reinterpret_cast<zd *>(vd);
    if ((rc = cgs2_passes(s, reinterpret_cast<const zd *>(Qd), ldq, k, v, N, st))) return rc;
    zd *h = reinterpret_cast<zd *>(hdev);
    hipLaunchKernelGGL(k_cgs2_commit, dim3(1), dim3(kRedCols), 0, st, s.zbuf, k, h);
    const int g = (int)std::min<int64_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_scale_by, dim3(g), dim3(256), 0, st, v, h + k, 1e-14, reinterpret_cast<zd *>(qnext), N);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

extern "C" int hex_cgs2(hex_engine *e, const double *Qd, int64_t ldq, int k, double *vd, int64_t N, double *hcol,
                        void *stream) {
    if (!e || !vd || !hcol) {
        hex::set_error("cgs2: null argument");
        return HEX_EINVAL;
    }
    int rc = check_basis(Qd, ldq, k, N, "cgs2");
    if (rc) return rc;
    hex_engine_scratch s = hex_scratch_of(e);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if ((rc = cgs2_passes(s, reinterpret_cast<const zd *>(Qd), ldq, k, reinterpret_cast<zd *>(vd), N, st)))
        return rc;
    HEX_HIP(hipMemcpyAsync(s.hbuf_host, s.zbuf, sizeof(zd) * 2 * kRedCols, hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    zd *h = reinterpret_cast<zd *>(hcol);
    for (int m = 0; m < k; ++m) {
        // H[:k, j] += z1, then += z2 (propagate.py:17-25 order)
        zd a = zadd(h[m], s.hbuf_host[m]);
        h[m] = zadd(a, s.hbuf_host[kRedCols + m]);
    }
    h[k] = s.hbuf_host[kRedCols + k];
    return HEX_OK;
}

extern "C" int hex_dots(hex_engine *e, const double *Qd, int64_t ldq, int k, const double *vd, int64_t N, double *z,
                        void *stream) {
    if (!e || !vd || !z) {
        hex::set_error("dots: null argument");
        return HEX_EINVAL;
    }
    int rc = check_basis(Qd, ldq, k, N, "dots");
    if (rc) return rc;
    hex_engine_scratch s = hex_scratch_of(e);
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = dots_to(s, reinterpret_cast<const zd *>(Qd), ldq, k, reinterpret_cast<const zd *>(vd), N, s.zbuf, st);
    if (rc) return rc;
    HEX_HIP(hipMemcpyAsync(s.hbuf_host, s.zbuf, sizeof(zd) * k, hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    std::copy(s.hbuf_host, s.hbuf_host + k, reinterpret_cast<zd *>(z));
    return HEX_OK;
}

extern "C" int hex_dots_device(hex_engine *e, const double *Qd, int64_t ldq, int k, const double *vd, int64_t N,
                               double *zd_out, void *stream) {
    if (!e || !vd || !zd_out || k < 1) {
        hex::set_error("dots_device: null argument or empty basis");
        return HEX_EINVAL;
    }
    int rc = check_basis(Qd, ldq, k, N, "dots_device");
    if (rc) return rc;
    return dots_to(hex_scratch_of(e), reinterpret_cast<const zd *>(Qd), ldq, k, reinterpret_cast<const zd *>(vd), N,
                   reinterpret_cast<zd *>(zd_out), static_cast<hipStream_t>(stream));
}

extern "C" int hex_norm(hex_engine *e, const double *xd, int64_t N, double *norm, void *stream) {
    if (!e || !xd || !norm || N < 1) {
        hex::set_error("norm: bad argument");
        return HEX_EINVAL;
    }
    hex_engine_scratch s = hex_scratch_of(e);
    hipStream_t st = static_cast<hipStream_t>(stream);
    // k_update with count 0 only reads v.
    int rc = update_to(s, nullptr, N, 0, s.zbuf, reinterpret_cast<zd *>(const_cast<double *>(xd)), N, false,
                       nullptr, true, s.zbuf + kRedCols, st);
    if (rc) return rc;
    HEX_HIP(hipMemcpyAsync(s.hbuf_host, s.zbuf + kRedCols, sizeof(zd), hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    *norm = s.hbuf_host[0].x;
    return HEX_OK;
}

extern "C" int hex_scale(const double *x, double s, double *y, int64_t N, void *stream) {
    if (!x || !y || N < 1) {
        hex::set_error("scale: bad argument");
        return HEX_EINVAL;
    }
    const int g = (int)std::min<int64_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_scale, dim3(g), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const zd *>(x), s, reinterpret_cast<zd *>(y), N);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

extern "C" int hex_combine(hex_engine *e, const double *Qd, int64_t ldq, int k, const double *c, double *y,
                           int64_t N, void *stream) {
    if (!e || !c || !y || k < 1) {
        hex::set_error("combine: bad argument");
        return HEX_EINVAL;
    }
    int rc = check_basis(Qd, ldq, k, N, "combine");
    if (rc) return rc;
    Coeffs cf{};
    std::copy(reinterpret_cast<const zd *>(c), reinterpret_cast<const zd *>(c) + k, cf.c);
    const int g = (int)std::min<int64_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_combine, dim3(g), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const zd *>(Qd), ldq, k, cf, reinterpret_cast<zd *>(y), N);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// Host evaluation of the small-matrix exponential used by the device Krylov
// control (same code path as the kernel); A, E row-major complex m x m.
extern "C" int hex_expm_host(int m, const double *A, double *E) {
    if (m < 1 || m > kExpmMax || !A || !E) {
        hex::set_error("expm: need 1 <= m <= 64");
        return HEX_EINVAL;
    }
    std::vector<zd> work(6 * (size_t)m * m);
    hexm::expm(m, reinterpret_cast<const zd *>(A), reinterpret_cast<zd *>(E), work.data());
    return HEX_OK;
}

// Many small exponentials at once, one wave each: E_b = expm(A_b) for b <
// count, all of order m (row-major, contiguous), work 6 m^2 per matrix.
__global__ __launch_bounds__(64) void k_expm_batched(int m, const zd *__restrict__ A, zd *__restrict__ E,
                                                     zd *__restrict__ work) {
    const int64_t b = blockIdx.x, mm = (int64_t)m * m;
    hexw::expm(m, A + b * mm, E + b * mm, work + b * 6 * mm);
}

extern "C" int hex_expm_batched(int count, int m, const double *A, double *E, double *work, void *stream) {
    if (count < 1 || m < 1 || m > kExpmMax || !A || !E || !work) {
        hex::set_error("expm_batched: need count >= 1, 1 <= m <= 64 and device A, E, work");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(k_expm_batched, dim3(count), dim3(64), 0, static_cast<hipStream_t>(stream), m,
                       reinterpret_cast<const zd *>(A), reinterpret_cast<zd *>(E), reinterpret_cast<zd *>(work));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// Krylov control on the device.  The Hessenberg columns written by
// hex_cgs2_commit (column jj: h[0..jj], then ||v|| at jj + 1, 64 entries per
// column) are gathered into the caller's H (row-major, (maxdim + 1) x maxdim)
// and the reference's per-iteration tests run in one wave; only a small result
// record goes back to the host: result[0] = (status, m), result[1] =
// (estimate, h_{m, m-1}) of the last test made, result[2 ..] = the coefficients.
static constexpr int kCheckLds = 22;   // matrices of order <= 22 fit the check kernels' LDS

// Arnoldi (python/propagate.py:26-35 / python/torch_backend.py:146-158): after
// column jj, breakdown if h_{jj+1,jj} < 1e-13; at jj >= 3 with (jj + 1) % 4 == 0
// or jj + 1 == maxdim, c = expm(-i dt H_m)[:, 0] beta and err = |dt h c_{m-1}|;
// status 1 (with c) when happy or err < tol.  Columns first .. last.
// fixed > 0: the forced dimension of a replayed step -- the only test besides
// a breakdown is at jj + 1 == fixed, and it always accepts.
// expm of an order above kCheckLds in global scratch (A, E, then 6 m^2 of work)
__device__ __noinline__ void expm_global(int m, zd *gwork) {
    hexw::expm(m, gwork, gwork + m * m, gwork + 2 * m * m);
}

__global__ __launch_bounds__(64) void k_arnoldi_check(const zd *__restrict__ hcols, int first, int last, int maxdim,
                                                      int fixed, double dt, double beta, double tol,
                                                      zd *__restrict__ H, zd *__restrict__ gwork,
                                                      zd *__restrict__ result) {
    __shared__ zd lwork[8 * kCheckLds * kCheckLds];
    const int lane = threadIdx.x;
    zd status = zmake(0.0, 0.0), estimate = zmake(-1.0, 0.0);
    for (int jj = first; jj <= last; ++jj) {
        const zd *col = hcols + (int64_t)jj * 64;
        for (int r = lane; r <= jj; r += 64) H[r * maxdim + jj] = col[r];
        const double h = col[jj + 1].x;
        if (lane == 0) H[(jj + 1) * maxdim + jj] = zmake(h, 0.0);
        __syncthreads();
        const bool happy = h < 1e-13;
        const bool check = fixed > 0 ? jj + 1 == fixed : jj >= 3 && ((jj + 1) % 4 == 0 || jj + 1 == maxdim);
        if (!(happy || check)) continue;
        const int m = jj + 1;
        const bool in_lds = m <= kCheckLds;
        zd *work = in_lds ? lwork : gwork;
        zd *A = work, *E = work + m * m;
        for (int e = lane; e < m * m; e += 64) {
            const int r = e / m, c = e - r * m;
            const zd x = r <= c + 1 ? H[r * maxdim + c] : zmake(0.0, 0.0);
            A[e] = zmake(dt * x.y, -dt * x.x);   // -i dt H
        }
        __syncthreads();
        // the LDS call is inlined (LDS instructions); orders above kCheckLds call out
        if (in_lds)
            hexw::expm(m, lwork, lwork + m * m, lwork + 2 * m * m);
        else
            expm_global(m, gwork);
        const double err = fabs(dt * h) * beta * sqrt(zabs2(E[(m - 1) * m]));
        estimate = zmake(err, h);
        if (happy || fixed > 0 || err < tol) {
            for (int r = lane; r < m; r += 64) result[2 + r] = zscale(E[r * m], beta);
            status = zmake(1.0, (double)m);
            break;
        }
        __syncthreads();
    }
    if (lane == 0) {
        result[0] = status;
        result[1] = estimate;
    }
}

// GMRES (python/implicit.py:55-72): after column jj, at a breakdown
// (h_{jj+1,jj} < 1e-14), every `cadence`-th column or the restart length, the
// least-squares problem min |beta e_1 - H_m y| of the (m + 1) x m Hessenberg is
// solved by Givens rotations and back substitution; its residual is the
// estimate.  status 1 (with y) when the reference would form a candidate:
// estimate < tol |b|, a breakdown, or m = restart.  Columns first .. last.
__global__ __launch_bounds__(64) void k_gmres_check(const zd *__restrict__ hcols, int first, int last, int restart,
                                                    int cadence, double beta, double tol_b, zd *__restrict__ H,
                                                    zd *__restrict__ result) {
    __shared__ zd R[64 * 64];   // the rotated Hessenberg of one check
    __shared__ zd g[65];
    __shared__ zd y[64];
    const int lane = threadIdx.x;
    zd status = zmake(0.0, 0.0), estimate = zmake(-1.0, 0.0);
    for (int jj = first; jj <= last; ++jj) {
        const zd *col = hcols + (int64_t)jj * 64;
        for (int r = lane; r <= jj; r += 64) H[r * restart + jj] = col[r];
        const double h = col[jj + 1].x;
        if (lane == 0) H[(jj + 1) * restart + jj] = zmake(h, 0.0);
        __syncthreads();
        if (!(h < 1e-14 || (jj + 1) % cadence == 0 || jj + 1 == restart)) continue;
        const int m = jj + 1;
        for (int e = lane; e < (m + 1) * m; e += 64) {
            const int r = e / m, c = e - r * m;
            R[r * 64 + c] = r <= c + 1 ? H[r * restart + c] : zmake(0.0, 0.0);
        }
        for (int r = lane; r <= m; r += 64) g[r] = r == 0 ? zmake(beta, 0.0) : zmake(0.0, 0.0);
        __syncthreads();
        // column k: rotate rows k, k+1 so that R[k+1][k] = 0 (lane c updates column c)
        for (int k = 0; k < m; ++k) {
            const zd a = R[k * 64 + k], b = R[(k + 1) * 64 + k];
            const double na = sqrt(zabs2(a)), nb = sqrt(zabs2(b));
            const double rho = hypot(na, nb);
            // [c  s; -conj(s) c] with c real: c = |a|/rho, s = (a/|a|) conj(b)/rho
            zd cs, sn;
            if (nb == 0.0) {
                cs = zmake(1.0, 0.0);
                sn = zmake(0.0, 0.0);
            } else if (na == 0.0) {
                cs = zmake(0.0, 0.0);
                sn = zmake(1.0, 0.0);
            } else {
                const zd phase = zscale(a, 1.0 / na);
                cs = zmake(na / rho, 0.0);
                sn = zscale(zmul(phase, zconj(b)), 1.0 / rho);
            }
            __syncthreads();
            for (int c = k + lane; c < m; c += 64) {
                const zd u = R[k * 64 + c], v = R[(k + 1) * 64 + c];
                R[k * 64 + c] = zadd(zmul(cs, u), zmul(sn, v));
                R[(k + 1) * 64 + c] = zsub(zmul(cs, v), zmul(zconj(sn), u));
            }
            if (lane == 0) {
                const zd u = g[k], v = g[k + 1];
                g[k] = zadd(zmul(cs, u), zmul(sn, v));
                g[k + 1] = zsub(zmul(cs, v), zmul(zconj(sn), u));
            }
            __syncthreads();
        }
        const double est = sqrt(zabs2(g[m]));
        estimate = zmake(est, h);
        if (est < tol_b || h < 1e-14 || m == restart) {
            if (lane == 0) {
                for (int k = m - 1; k >= 0; --k) {
                    zd acc = g[k];
                    for (int c = k + 1; c < m; ++c) acc = zsub(acc, zmul(R[k * 64 + c], y[c]));
                    y[k] = zdiv(acc, R[k * 64 + k]);
                }
            }
            __syncthreads();
            for (int r = lane; r < m; r += 64) result[2 + r] = y[r];
            status = zmake(1.0, (double)m);
            break;
        }
    }
    if (lane == 0) {
        result[0] = status;
        result[1] = estimate;
    }
}

extern "C" int hex_arnoldi_check(const double *hcols, int first, int last, int maxdim, double dt, double beta,
                                 double tol, double *H, double *work, double *result, void *stream) {
    if (!hcols || !H || !result || first < 0 || last < first || last >= maxdim || maxdim > kRedCols - 2 ||
        (maxdim > kCheckLds && !work)) {
        hex::set_error("arnoldi_check: bad argument (columns first <= last < maxdim <= 62; work for maxdim > 22)");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(k_arnoldi_check, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const zd *>(hcols), first, last, maxdim, 0, dt, beta, tol,
                       reinterpret_cast<zd *>(H), reinterpret_cast<zd *>(work), reinterpret_cast<zd *>(result));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

extern "C" int hex_arnoldi_check_fixed(const double *hcols, int first, int last, int maxdim, int fixed, double dt,
                                       double beta, double *H, double *work, double *result, void *stream) {
    if (!hcols || !H || !result || first < 0 || last < first || last >= maxdim || maxdim > kRedCols - 2 ||
        fixed < 1 || fixed > maxdim || (maxdim > kCheckLds && !work)) {
        hex::set_error("arnoldi_check_fixed: bad argument (columns first <= last < maxdim <= 62, "
                       "1 <= fixed <= maxdim; work for maxdim > 22)");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(k_arnoldi_check, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const zd *>(hcols), first, last, maxdim, fixed, dt, beta, 0.0,
                       reinterpret_cast<zd *>(H), reinterpret_cast<zd *>(work), reinterpret_cast<zd *>(result));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

extern "C" int hex_gmres_check(const double *hcols, int first, int last, int restart, int cadence, double beta,
                               double tol_b, double *H, double *result, void *stream) {
    if (!hcols || !H || !result || first < 0 || last < first || last >= restart || restart > kRedCols - 2 ||
        cadence < 1) {
        hex::set_error("gmres_check: bad argument (columns first <= last < restart <= 62)");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(k_gmres_check, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const zd *>(hcols), first, last, restart, cadence, beta, tol_b,
                       reinterpret_cast<zd *>(H), reinterpret_cast<zd *>(result));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}
