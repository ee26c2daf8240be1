// He+ channel yield of the adjoint objective and its exact gradient on the
// device: the ionic duals by backward CF4-Pade steps of dense factor inverses,
// their projection onto the surface frames, the Volkov amplitudes with their
// phase integrals, and the reverse sweep of all of it.
//
// Value (python/streaming_surface.py:39-64 and :198-214 restated for a
// sampled control; python/implicit.py:111-123 for the dual CF4 steps):
//   chi_{i-1} = P_i chi_i, i = N .. 1, chi_N = the dual final states,
//   P_i = F_3 F_2 F_1 F_0,  F_f = 2 M_f^-1 - 1,  M_f = 1 + c_f H_f,
//   c_f = i (-dt) / root_f,  H_f = s h0 + alpha_f + beta_f pz, with
//   (alpha, beta) from the stage fields a1, a2 of step i (two CF4 pairs);
//   q_n(j, s)  = sum_{E: ej = j} ew_E sum_r conj(chi_n[ea, r, eu]) F_n[ec][s][r],
//   F_n = F0_n + A_n Fd_n;
//   amp(l, e, d) = dt sum_n w_n i sqrt2 e^{i phi_n(e, d)} sum_o harm[o, d]
//                  sum_s tests[o, e, s] q_n(l nout + o, s),
//   phi_n = E_e t_n + kz[e, d] intA_n + intA2_n / 2 (trapezoid integrals of A, A^2),
//   Y = sum wy[e, d] |amp(l, e, d)|^2.
// Gradient (Wirtinger, bar x = dY/d conj(x)): bar amp = wy amp; bar q and the
// phase derivatives per sample; bar F_n = sum_E conj(ew) chi_n q-bar and
// bar chi_n = sum_E ew conj(q-bar) F_n; the dual costate mu_0 = bar chi_0,
// mu_n = bar chi_n + P_n^H mu_{n-1}, and per factor
//   dY/da = -4 Re sum_f c_f (dalpha_f/da y_f^H z_f + dbeta_f/da y_f^H pz z_f),
// z_f = M_f^-1 x_{f-1} (forward), y_f = M_f^-H w_f (reverse).
//
// Every reduction runs in a fixed order, so results do not depend on timing.
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int kYThreads = 256;
constexpr double kSqrt3 = 1.7320508075688772;

struct Dev {
    std::vector<void *> owned;
    int64_t bytes = 0;
    template <typename T> int alloc(size_t count, T **out) {
        void *p = nullptr;
        if (count == 0) count = 1;
        if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) {
            hex::set_error("yield: device allocation failed");
            return HEX_EDEVICE;
        }
        owned.push_back(p);
        bytes += (int64_t)(sizeof(T) * count);
        *out = static_cast<T *>(p);
        return HEX_OK;
    }
    template <typename T> int upload(const T *host, size_t count, T **out) {
        int rc = alloc(count, out);
        if (rc) return rc;
        if (hipMemcpy(*out, host, sizeof(T) * count, hipMemcpyHostToDevice) != hipSuccess) {
            hex::set_error("yield: upload failed");
            return HEX_EDEVICE;
        }
        return HEX_OK;
    }
};

// block-wide sum in a fixed butterfly order; every thread gets the result
__device__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    return s;
}

__device__ __forceinline__ zd zexpi(double phi) {
    double s, c;
    sincos(phi, &s, &c);
    return zmake(c, s);
}

struct Factor {
    zd c;              // i (-dt) / root
    double alpha, beta, s;
    double da1, da2;   // d alpha / d a1, d alpha / d a2
    double db1, db2;   // d beta / d a1, d beta / d a2
};

// factor f (0..3) of a step with stage fields a1, a2: f < 2 the pair
// A2 h1 + A1 h2, f >= 2 the pair A1 h1 + A2 h2; roots 3 + i sqrt3 (even f),
// 3 - i sqrt3 (odd f); h_k = h0 + a_k^2 / 2 + a_k pz
__host__ __device__ inline Factor factor_of(int f, double a1, double a2, double dt) {
    const double A1 = (3.0 - 2.0 * kSqrt3) / 12.0, A2 = (3.0 + 2.0 * kSqrt3) / 12.0;
    const double w1 = f < 2 ? A2 : A1, w2 = f < 2 ? A1 : A2;
    const zd root = zmake(3.0, (f & 1) ? -kSqrt3 : kSqrt3);
    Factor F;
    F.c = zdiv(zmake(0.0, -dt), root);
    F.s = w1 + w2;
    F.alpha = w1 * a1 * a1 / 2 + w2 * a2 * a2 / 2;
    F.beta = w1 * a1 + w2 * a2;
    F.da1 = w1 * a1;
    F.da2 = w2 * a2;
    F.db1 = w1;
    F.db2 = w2;
    return F;
}

// M_f of every (step i = 1..N, factor f), column-major M x M, batch i-1 major
__global__ void k_dual_matrices(int M, int N, double dt, const double *__restrict__ a, const zd *__restrict__ h0,
                                const zd *__restrict__ pz, zd *__restrict__ mats) {
    const int64_t per = (int64_t)M * M, total = per * 4 * N;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = x / per, e = x - b * per;
        const int i = (int)(b / 4) + 1, f = (int)(b % 4);
        const int col = (int)(e / M), row = (int)(e - (int64_t)col * M);
        const Factor F = factor_of(f, a[2 * i], a[2 * i + 1], dt);
        // h0, pz row-major
        const zd h = h0[(int64_t)row * M + col], p = pz[(int64_t)row * M + col];
        zd H = zmake(F.s * h.x + F.beta * p.x, F.s * h.y + F.beta * p.y);
        if (row == col) H.x += F.alpha;
        zd v = zmul(F.c, H);
        if (row == col) v.x += 1.0;
        mats[x] = v;
    }
}

// z = Minv x (Minv column-major), x_out = 2 z - x; M x nu blocks row-major.
// The sweep runs these one after another, so one product is latency-bound:
// each 256-lane block takes kDualRows rows, lane t the row t % kDualRows and
// the columns c = t / kDualRows (mod kDualSlices) (a wave reads four 256-byte
// column segments per step), and the kDualSlices partial sums of a row are
// added in LDS in slice order.
constexpr int kDualRows = 16, kDualSlices = 16;

__global__ __launch_bounds__(kDualRows *kDualSlices) void k_dual_forward(int M, int nu, const zd *__restrict__ minv,
                                                                        const zd *__restrict__ x,
                                                                        zd *__restrict__ z, zd *__restrict__ xout) {
    __shared__ zd part[2][kDualSlices][kDualRows];
    const int t = threadIdx.x, rr = t % kDualRows, sl = t / kDualRows;
    const int r = blockIdx.x * kDualRows + rr;
    zd acc0 = zmake(0.0, 0.0), acc1 = zmake(0.0, 0.0);
    if (r < M) {
        for (int c = sl; c < M; c += kDualSlices) {
            const zd m = minv[r + (int64_t)c * M];
            acc0 = zfma(m, x[c * nu], acc0);
            if (nu > 1) acc1 = zfma(m, x[c * nu + 1], acc1);
        }
    }
    part[0][sl][rr] = acc0;
    part[1][sl][rr] = acc1;
    __syncthreads();
    if (t < 2 * kDualRows) {
        const int u = t / kDualRows, row = blockIdx.x * kDualRows + t % kDualRows;
        if (u < nu && row < M) {
            zd s = part[u][0][t % kDualRows];
            for (int k = 1; k < kDualSlices; ++k) s = zadd(s, part[u][k][t % kDualRows]);
            z[row * nu + u] = s;
            xout[row * nu + u] = zsub(zscale(s, 2.0), x[row * nu + u]);
        }
    }
}

// y = Minv^H w; w_out = 2 y - w (+ add, when given); one block per column c
__global__ __launch_bounds__(64) void k_dual_reverse(int M, int nu, const zd *__restrict__ minv,
                                                     const zd *__restrict__ w, const zd *__restrict__ add,
                                                     zd *__restrict__ y, zd *__restrict__ wout) {
    const int c = blockIdx.x, lane = threadIdx.x;
    const zd *col = minv + (int64_t)c * M;
    double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;
    for (int r = lane; r < M; r += 64) {
        const zd m = col[r];
        const zd t0 = zcmul(m, w[r * nu]);   // conj(m) w
        re0 += t0.x;
        im0 += t0.y;
        if (nu > 1) {
            const zd t1 = zcmul(m, w[r * nu + 1]);
            re1 += t1.x;
            im1 += t1.y;
        }
    }
    re0 = wave_sum(re0);
    im0 = wave_sum(im0);
    re1 = wave_sum(re1);
    im1 = wave_sum(im1);
    if (lane == 0) {
        for (int u = 0; u < nu && u < 2; ++u) {
            const zd yy = u == 0 ? zmake(re0, im0) : zmake(re1, im1);
            y[c * nu + u] = yy;
            zd o = zsub(zscale(yy, 2.0), w[c * nu + u]);
            if (add) o = zadd(o, add[c * nu + u]);
            wout[c * nu + u] = o;
        }
    }
}

// v_b = pz z_b for every stored forward vector (pz row-major)
__global__ void k_pz_apply(int M, int nu, int64_t count, const zd *__restrict__ pz, const zd *__restrict__ z,
                           zd *__restrict__ v) {
    const int64_t total = count * M;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = x / M;
        const int r = (int)(x - b * M);
        const zd *zb = z + b * M * nu;
        const zd *pr = pz + (int64_t)r * M;
        for (int u = 0; u < nu; ++u) {
            zd acc = zmake(0.0, 0.0);
            for (int c = 0; c < M; ++c) acc = zfma(pr[c], zb[c * nu + u], acc);
            v[b * M * nu + r * nu + u] = acc;
        }
    }
}

// per stored factor b: (y_b^H z_b, y_b^H v_b)
__global__ void k_factor_dots(int len, const zd *__restrict__ y, const zd *__restrict__ z, const zd *__restrict__ v,
                              zd *__restrict__ dots) {
    __shared__ double red[kYThreads / 64];
    const int64_t b = blockIdx.x;
    const zd *yb = y + b * len, *zb = z + b * len, *vb = v + b * len;
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
        const zd p = zcmul(yb[k], zb[k]), q = zcmul(yb[k], vb[k]);
        a0 += p.x;
        a1 += p.y;
        b0 += q.x;
        b1 += q.y;
    }
    a0 = block_sum(a0, red);
    a1 = block_sum(a1, red);
    b0 = block_sum(b0, red);
    b1 = block_sum(b1, red);
    if (threadIdx.x == 0) {
        dots[2 * b] = zmake(a0, a1);
        dots[2 * b + 1] = zmake(b0, b1);
    }
}

// dY/da1, dY/da2 of every step from the factor dots
__global__ void k_stage_gradients(int N, double dt, const double *__restrict__ a, const zd *__restrict__ dots,
                                  double *__restrict__ da) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // 0 .. N
    if (i > N) return;
    double g1 = 0.0, g2 = 0.0;
    if (i >= 1) {
        for (int f = 0; f < 4; ++f) {
            const Factor F = factor_of(f, a[2 * i], a[2 * i + 1], dt);
            const int64_t b = (int64_t)(i - 1) * 4 + f;
            const zd yz = dots[2 * b], yv = dots[2 * b + 1];
            const zd t1 = zmul(F.c, zmake(F.da1 * yz.x + F.db1 * yv.x, F.da1 * yz.y + F.db1 * yv.y));
            const zd t2 = zmul(F.c, zmake(F.da2 * yz.x + F.db2 * yv.x, F.da2 * yz.y + F.db2 * yv.y));
            g1 += -4.0 * t1.x;
            g2 += -4.0 * t2.x;
        }
    }
    da[2 * i] = g1;
    da[2 * i + 1] = g2;
}

struct Tables {
    int nent, nlab, nout, ns, nc, nr, na, nu, nE, nD, S;
    const int32_t *ej, *ec, *ea, *eu;
    const zd *ew;
    const zd *tests, *harm;         // [nout][nE][ns], [nout][nD]
    const double *energy, *kz, *wy; // [nE], [nE][nD], [nE][nD]
    const double *times;            // [S]
};

// q_n(j, s), one block per sample
__global__ void k_project(Tables t, const zd *__restrict__ chis, const zd *__restrict__ F0, const zd *__restrict__ Fd,
                          const double *__restrict__ A, zd *__restrict__ q) {
    const int n = blockIdx.x;
    const int J = t.nlab * t.nout;
    const int64_t fstride = (int64_t)t.nc * t.ns * t.nr;
    const zd *f0 = F0 + n * fstride, *fd = Fd + n * fstride;
    const zd *chi = chis + (int64_t)n * t.na * t.nr * t.nu;
    const double An = A[n];
    for (int k = threadIdx.x; k < J * t.ns; k += blockDim.x) {
        const int j = k / t.ns, s = k - j * t.ns;
        zd acc = zmake(0.0, 0.0);
        for (int E = 0; E < t.nent; ++E) {
            if (t.ej[E] != j) continue;
            const int64_t fo = ((int64_t)t.ec[E] * t.ns + s) * t.nr;
            zd part = zmake(0.0, 0.0);
            for (int r = 0; r < t.nr; ++r) {
                const zd F = zmake(f0[fo + r].x + An * fd[fo + r].x, f0[fo + r].y + An * fd[fo + r].y);
                part = zcfma(chi[((int64_t)t.ea[E] * t.nr + r) * t.nu + t.eu[E]], F, part);
            }
            acc = zfma(t.ew[E], part, acc);
        }
        q[(int64_t)n * J * t.ns + k] = acc;
    }
}

// intA_n, intA2_n (trapezoid from t_0), one thread in sample order
__global__ void k_phase_integrals(int S, double dt, const double *__restrict__ A, double *__restrict__ ints) {
    if (threadIdx.x || blockIdx.x) return;
    double i1 = 0.0, i2 = 0.0;
    ints[0] = 0.0;
    ints[1] = 0.0;
    for (int n = 1; n < S; ++n) {
        i1 += dt * (A[n] + A[n - 1]) / 2;
        i2 += dt * (A[n] * A[n] + A[n - 1] * A[n - 1]) / 2;
        ints[2 * n] = i1;
        ints[2 * n + 1] = i2;
    }
}

// sum_o harm[o, d] sum_s tests[o, e, s] q_n(l, o, s)
__device__ zd radial_sum(const Tables &t, const zd *qn, int l, int e, int d) {
    zd acc = zmake(0.0, 0.0);
    for (int o = 0; o < t.nout; ++o) {
        zd b = zmake(0.0, 0.0);
        const zd *qo = qn + ((int64_t)l * t.nout + o) * t.ns;
        const zd *to = t.tests + ((int64_t)o * t.nE + e) * t.ns;
        for (int s = 0; s < t.ns; ++s) b = zfma(to[s], qo[s], b);
        acc = zfma(t.harm[o * t.nD + d], b, acc);
    }
    return acc;
}

// amp(l, e, d): one block per (l, e, d), samples strided over the threads
__global__ void k_amplitude(Tables t, double dt, const zd *__restrict__ q, const double *__restrict__ ints,
                            zd *__restrict__ amp) {
    __shared__ double red[kYThreads / 64];
    const int x = blockIdx.x;
    const int d = x % t.nD, e = (x / t.nD) % t.nE, l = x / (t.nD * t.nE);
    const int64_t qs = (int64_t)t.nlab * t.nout * t.ns;
    double re = 0.0, im = 0.0;
    for (int n = threadIdx.x; n < t.S; n += blockDim.x) {
        const double w = (n == 0 || n == t.S - 1) ? 0.5 : 1.0;
        const double phi = t.energy[e] * t.times[n] + t.kz[e * t.nD + d] * ints[2 * n] + ints[2 * n + 1] / 2;
        const zd v = zmul(zexpi(phi), radial_sum(t, q + n * qs, l, e, d));
        re += w * v.x;
        im += w * v.y;
    }
    re = block_sum(re, red);
    im = block_sum(im, red);
    // i sqrt2 dt (re + i im)
    if (threadIdx.x == 0) amp[x] = zmake(-M_SQRT2 * dt * im, M_SQRT2 * dt * re);
}

__global__ void k_yield_value(Tables t, const zd *__restrict__ amp, double *__restrict__ Y) {
    __shared__ double red[kYThreads / 64];
    const int total = t.nlab * t.nE * t.nD;
    double s = 0.0;
    for (int x = threadIdx.x; x < total; x += blockDim.x) {
        const int ed = x % (t.nE * t.nD);
        s += t.wy[ed] * zabs2(amp[x]);
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) *Y = s;
}

// per sample: bar q_n and dY/d intA_n, dY/d intA2_n
__global__ void k_amplitude_reverse(Tables t, double dt, const zd *__restrict__ q, const double *__restrict__ ints,
                                    const zd *__restrict__ amp, zd *__restrict__ qbar, double *__restrict__ dints) {
    extern __shared__ zd ph[];   // [nE nD]: dt w_n i sqrt2 e^{i phi}
    __shared__ double red[kYThreads / 64];
    const int n = blockIdx.x;
    const int ED = t.nE * t.nD;
    const double w = (n == 0 || n == t.S - 1) ? 0.5 : 1.0;
    const int64_t qs = (int64_t)t.nlab * t.nout * t.ns;
    const zd *qn = q + n * qs;
    double g1 = 0.0, g2 = 0.0;
    for (int x = threadIdx.x; x < ED; x += blockDim.x) {
        const int e = x / t.nD, d = x - e * t.nD;
        const double phi = t.energy[e] * t.times[n] + t.kz[x] * ints[2 * n] + ints[2 * n + 1] / 2;
        const zd k = zmul(zmake(0.0, M_SQRT2 * dt * w), zexpi(phi));
        ph[x] = k;
        // dY/dphi = 2 Re sum_l conj(bar amp) i k radial
        double g = 0.0;
        for (int l = 0; l < t.nlab; ++l) {
            const zd bar = zscale(amp[(int64_t)l * ED + x], t.wy[x]);
            const zd term = zmul(zmake(0.0, 1.0), zmul(k, radial_sum(t, qn, l, e, d)));
            g += 2.0 * (bar.x * term.x + bar.y * term.y);
        }
        g1 += t.kz[x] * g;
        g2 += 0.5 * g;
    }
    g1 = block_sum(g1, red);
    g2 = block_sum(g2, red);
    if (threadIdx.x == 0) {
        dints[2 * n] = g1;
        dints[2 * n + 1] = g2;
    }
    __syncthreads();
    // bar q(l, o, s) = sum_{e, d} conj(tests[o, e, s] k harm[o, d]) bar amp(l, e, d)
    for (int x = threadIdx.x; x < (int)qs; x += blockDim.x) {
        const int s = x % t.ns, o = (x / t.ns) % t.nout, l = x / (t.ns * t.nout);
        zd acc = zmake(0.0, 0.0);
        for (int e = 0; e < t.nE; ++e) {
            zd inner = zmake(0.0, 0.0);
            for (int d = 0; d < t.nD; ++d) {
                const int ed = e * t.nD + d;
                const zd bar = zscale(amp[(int64_t)l * ED + ed], t.wy[ed]);
                inner = zcfma(zmul(ph[ed], t.harm[o * t.nD + d]), bar, inner);
            }
            acc = zcfma(t.tests[((int64_t)o * t.nE + e) * t.ns + s], inner, acc);
        }
        qbar[n * qs + x] = acc;
    }
}

// dY/dA_m from the phase integrals (suffix sums of their derivatives)
__global__ void k_phase_reverse(int S, double dt, const double *__restrict__ A, const double *__restrict__ dints,
                                double *__restrict__ dA) {
    if (threadIdx.x || blockIdx.x) return;
    double r1 = 0.0, r2 = 0.0;   // sums over n >= m + 1
    for (int m = S - 1; m >= 0; --m) {
        const double n1 = r1, n2 = r2;   // n >= m + 1
        if (m >= 1) {
            r1 += dints[2 * m];
            r2 += dints[2 * m + 1];
        }
        // intA_n holds A_m with weight dt/2 for each of the pairs (m-1, m), (m, m+1) inside [0, n]
        const double s1 = (m >= 1 ? r1 : 0.0) + n1;
        const double s2 = (m >= 1 ? r2 : 0.0) + n2;
        dA[m] = dt / 2 * s1 + dt / 2 * 2.0 * A[m] * s2;
    }
}

// bar F_n (as G0, A_n bar F_n as Gd), bar chi_n and the frames' dY/dA_n
__global__ void k_project_reverse(Tables t, const zd *__restrict__ chis, const zd *__restrict__ F0,
                                  const zd *__restrict__ Fd, const double *__restrict__ A, const zd *__restrict__ qbar,
                                  zd *__restrict__ G0, zd *__restrict__ Gd, zd *__restrict__ chibar,
                                  double *__restrict__ dA) {
    __shared__ double red[kYThreads / 64];
    const int n = blockIdx.x;
    const int64_t fstride = (int64_t)t.nc * t.ns * t.nr;
    const int64_t qs = (int64_t)t.nlab * t.nout * t.ns;
    const zd *f0 = F0 + n * fstride, *fd = Fd + n * fstride;
    const zd *chi = chis + (int64_t)n * t.na * t.nr * t.nu;
    const zd *qb = qbar + n * qs;
    const double An = A[n];
    double g = 0.0;
    for (int64_t x = threadIdx.x; x < fstride; x += blockDim.x) {
        const int r = (int)(x % t.nr), s = (int)((x / t.nr) % t.ns), c = (int)(x / ((int64_t)t.nr * t.ns));
        zd acc = zmake(0.0, 0.0);
        for (int E = 0; E < t.nent; ++E) {
            if (t.ec[E] != c) continue;
            const zd cq = zmul(chi[((int64_t)t.ea[E] * t.nr + r) * t.nu + t.eu[E]], qb[(int64_t)t.ej[E] * t.ns + s]);
            acc = zcfma(t.ew[E], cq, acc);   // conj(ew) chi qbar
        }
        G0[n * fstride + x] = acc;
        Gd[n * fstride + x] = zscale(acc, An);
        g += 2.0 * (acc.x * fd[x].x + acc.y * fd[x].y);   // 2 Re conj(bar F) Fd
    }
    g = block_sum(g, red);
    if (threadIdx.x == 0) dA[n] += g;
    const int len = t.na * t.nr * t.nu;
    for (int x = threadIdx.x; x < len; x += blockDim.x) {
        const int u = x % t.nu, r = (x / t.nu) % t.nr, a = x / (t.nu * t.nr);
        zd acc = zmake(0.0, 0.0);
        for (int E = 0; E < t.nent; ++E) {
            if (t.ea[E] != a || t.eu[E] != u) continue;
            const int64_t fo = ((int64_t)t.ec[E] * t.ns) * t.nr + r;
            zd part = zmake(0.0, 0.0);
            for (int s = 0; s < t.ns; ++s) {
                const zd F = zmake(f0[fo + (int64_t)s * t.nr].x + An * fd[fo + (int64_t)s * t.nr].x,
                                   f0[fo + (int64_t)s * t.nr].y + An * fd[fo + (int64_t)s * t.nr].y);
                part = zcfma(qb[(int64_t)t.ej[E] * t.ns + s], F, part);   // conj(qbar) F
            }
            acc = zfma(t.ew[E], part, acc);
        }
        chibar[(int64_t)n * len + x] = acc;
    }
}

}  // namespace

struct hex_yield {
    int device = 0;
    Dev mem;
    rocblas_handle blas = nullptr;
    Tables t{};
    int M = 0, N = 0;
    double dt = 0.0;
    zd *h0 = nullptr, *pz = nullptr, *chiN = nullptr;
    zd *minv = nullptr;        // [4N][M][M] column-major inverses
    rocblas_int *ipiv = nullptr, *info = nullptr;
    zd *chis = nullptr;        // [S][M][nu]
    zd *zs = nullptr, *vs = nullptr, *ys = nullptr;   // [4N][M][nu]
    zd *q = nullptr, *qbar = nullptr;                 // [S][nlab nout][ns]
    zd *amp = nullptr;                                // [nlab][nE][nD]
    zd *chibar = nullptr, *mu = nullptr, *dots = nullptr;
    double *a = nullptr, *A = nullptr, *ints = nullptr, *dints = nullptr, *dA = nullptr, *da = nullptr, *Y = nullptr;
    bool have_value = false;
};

extern "C" int hex_yield_create(int device, const hex_yield_desc *d, hex_yield **out) {
    if (!d || !out || d->M < 1 || d->nu < 1 || d->nu > 2 || d->na * d->nr != d->M || d->N < 1 || d->nent < 1 ||
        d->nlab < 1 || d->nout < 1 || d->ns < 1 || d->nc < 1 || d->nE < 1 || d->nD < 1 || !d->h0 || !d->pz ||
        !d->chiN || !d->ej || !d->ec || !d->ea || !d->eu || !d->ew || !d->tests || !d->harm || !d->energy ||
        !d->kz || !d->wy || !d->times) {
        hex::set_error("yield_create: bad descriptor (nu must be 1 or 2, na nr = M)");
        return HEX_EINVAL;
    }
    for (int E = 0; E < d->nent; ++E)
        if (d->ej[E] < 0 || d->ej[E] >= d->nlab * d->nout || d->ec[E] < 0 || d->ec[E] >= d->nc || d->ea[E] < 0 ||
            d->ea[E] >= d->na || d->eu[E] < 0 || d->eu[E] >= d->nu) {
            hex::set_error("yield_create: projection entry out of range");
            return HEX_EINVAL;
        }
    HEX_HIP(hipSetDevice(device));
    auto *p = new hex_yield;
    p->device = device;
    p->M = d->M;
    p->N = d->N;
    p->dt = d->dt;
    const int S = d->N + 1, M = d->M, nu = d->nu;
    Tables &t = p->t;
    t.nent = d->nent;
    t.nlab = d->nlab;
    t.nout = d->nout;
    t.ns = d->ns;
    t.nc = d->nc;
    t.nr = d->nr;
    t.na = d->na;
    t.nu = nu;
    t.nE = d->nE;
    t.nD = d->nD;
    t.S = S;
    int32_t *ej, *ec, *ea, *eu;
    zd *ew, *tests, *harm;
    double *energy, *kz, *wy, *times;
    const size_t vec = (size_t)M * nu, fac = (size_t)4 * d->N;
    int rc = 0;
    Dev &m = p->mem;
    if ((rc = m.upload(reinterpret_cast<const zd *>(d->h0), (size_t)M * M, &p->h0)) ||
        (rc = m.upload(reinterpret_cast<const zd *>(d->pz), (size_t)M * M, &p->pz)) ||
        (rc = m.upload(reinterpret_cast<const zd *>(d->chiN), vec, &p->chiN)) ||
        (rc = m.upload(d->ej, (size_t)d->nent, &ej)) || (rc = m.upload(d->ec, (size_t)d->nent, &ec)) ||
        (rc = m.upload(d->ea, (size_t)d->nent, &ea)) || (rc = m.upload(d->eu, (size_t)d->nent, &eu)) ||
        (rc = m.upload(reinterpret_cast<const zd *>(d->ew), (size_t)d->nent, &ew)) ||
        (rc = m.upload(reinterpret_cast<const zd *>(d->tests), (size_t)d->nout * d->nE * d->ns, &tests)) ||
        (rc = m.upload(reinterpret_cast<const zd *>(d->harm), (size_t)d->nout * d->nD, &harm)) ||
        (rc = m.upload(d->energy, (size_t)d->nE, &energy)) ||
        (rc = m.upload(d->kz, (size_t)d->nE * d->nD, &kz)) || (rc = m.upload(d->wy, (size_t)d->nE * d->nD, &wy)) ||
        (rc = m.upload(d->times, (size_t)S, &times)) || (rc = m.alloc(fac * M * M, &p->minv)) ||
        (rc = m.alloc(fac * M, &p->ipiv)) || (rc = m.alloc(fac, &p->info)) ||
        (rc = m.alloc((size_t)S * vec, &p->chis)) || (rc = m.alloc(fac * vec, &p->zs)) ||
        (rc = m.alloc(fac * vec, &p->vs)) || (rc = m.alloc(fac * vec, &p->ys)) ||
        (rc = m.alloc((size_t)S * d->nlab * d->nout * d->ns, &p->q)) ||
        (rc = m.alloc((size_t)S * d->nlab * d->nout * d->ns, &p->qbar)) ||
        (rc = m.alloc((size_t)d->nlab * d->nE * d->nD, &p->amp)) || (rc = m.alloc((size_t)S * vec, &p->chibar)) ||
        (rc = m.alloc(2 * vec, &p->mu)) || (rc = m.alloc(2 * fac, &p->dots)) || (rc = m.alloc((size_t)2 * S, &p->a)) ||
        (rc = m.alloc((size_t)S, &p->A)) || (rc = m.alloc((size_t)2 * S, &p->ints)) ||
        (rc = m.alloc((size_t)2 * S, &p->dints)) || (rc = m.alloc((size_t)S, &p->dA)) ||
        (rc = m.alloc((size_t)2 * S, &p->da)) || (rc = m.alloc(1, &p->Y))) {
        hex_yield_destroy(p);
        return rc;
    }
    t.ej = ej;
    t.ec = ec;
    t.ea = ea;
    t.eu = eu;
    t.ew = ew;
    t.tests = tests;
    t.harm = harm;
    t.energy = energy;
    t.kz = kz;
    t.wy = wy;
    t.times = times;
    if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
        hex::set_error("yield: rocblas_create_handle failed");
        hex_yield_destroy(p);
        return HEX_EDEVICE;
    }
    *out = p;
    return HEX_OK;
}

extern "C" int hex_yield_destroy(hex_yield *p) {
    if (!p) return HEX_OK;
    (void)hipSetDevice(p->device);
    if (p->blas) rocblas_destroy_handle(p->blas);
    for (void *x : p->mem.owned) (void)hipFree(x);
    delete p;
    return HEX_OK;
}

extern "C" int64_t hex_yield_bytes(const hex_yield *p) { return p ? p->mem.bytes : -1; }

extern "C" int hex_yield_value(hex_yield *p, const double *A, const double *a, const double *F0, const double *Fd,
                               double *Y, void *stream) {
    if (!p || !A || !a || !F0 || !Fd || !Y) {
        hex::set_error("yield_value: null argument");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(p->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int M = p->M, N = p->N, S = N + 1, nu = p->t.nu;
    const int64_t vec = (int64_t)M * nu, per = (int64_t)M * M;
    HEX_HIP(hipMemcpyAsync(p->A, A, sizeof(double) * S, hipMemcpyHostToDevice, st));
    HEX_HIP(hipMemcpyAsync(p->a, a, sizeof(double) * 2 * S, hipMemcpyHostToDevice, st));
    // the dual factors and their inverses (LU with partial pivoting)
    const int64_t total = per * 4 * N;
    const int mblocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_dual_matrices, dim3(mblocks), dim3(256), 0, st, M, N, p->dt, p->a, p->h0, p->pz, p->minv);
    HEX_HIP(hipGetLastError());
    rocblas_set_stream(p->blas, st);
    if (rocsolver_zgetrf_strided_batched(p->blas, M, M, reinterpret_cast<rocblas_double_complex *>(p->minv), M, per,
                                         p->ipiv, M, p->info, 4 * N) != rocblas_status_success ||
        rocsolver_zgetri_strided_batched(p->blas, M, reinterpret_cast<rocblas_double_complex *>(p->minv), M, per,
                                         p->ipiv, M, p->info, 4 * N) != rocblas_status_success) {
        hex::set_error("yield: rocsolver factorisation failed");
        return HEX_EDEVICE;
    }
    std::vector<rocblas_int> info(4 * (size_t)N);
    HEX_HIP(hipMemcpyAsync(info.data(), p->info, sizeof(rocblas_int) * info.size(), hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    for (rocblas_int v : info)
        if (v != 0) {
            hex::set_error("yield: a dual CF4 factor is singular");
            return HEX_EDEVICE;
        }
    // backward sweep of the duals: chi_N, then chi_{i-1} = P_i chi_i
    HEX_HIP(hipMemcpyAsync(p->chis + (int64_t)N * vec, p->chiN, sizeof(zd) * vec, hipMemcpyDeviceToDevice, st));
    const int rb = (M + kDualRows - 1) / kDualRows;
    for (int i = N; i >= 1; --i) {
        const zd *x = p->chis + (int64_t)i * vec;
        for (int f = 0; f < 4; ++f) {
            const int64_t b = (int64_t)(i - 1) * 4 + f;
            zd *xout = f == 3 ? p->chis + (int64_t)(i - 1) * vec : p->mu + (f & 1) * vec;
            hipLaunchKernelGGL(k_dual_forward, dim3(rb), dim3(kDualRows * kDualSlices), 0, st, M, nu,
                               p->minv + b * per, x, p->zs + b * vec, xout);
            x = xout;
        }
    }
    HEX_HIP(hipGetLastError());
    // projection, phase integrals, amplitudes, value
    hipLaunchKernelGGL(k_project, dim3(S), dim3(64), 0, st, p->t, p->chis, reinterpret_cast<const zd *>(F0),
                       reinterpret_cast<const zd *>(Fd), p->A, p->q);
    hipLaunchKernelGGL(k_phase_integrals, dim3(1), dim3(64), 0, st, S, p->dt, p->A, p->ints);
    hipLaunchKernelGGL(k_amplitude, dim3(p->t.nlab * p->t.nE * p->t.nD), dim3(kYThreads), 0, st, p->t, p->dt, p->q,
                       p->ints, p->amp);
    hipLaunchKernelGGL(k_yield_value, dim3(1), dim3(kYThreads), 0, st, p->t, p->amp, p->Y);
    HEX_HIP(hipGetLastError());
    HEX_HIP(hipMemcpyAsync(Y, p->Y, sizeof(double), hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    p->have_value = true;
    return HEX_OK;
}

extern "C" int hex_yield_gradient(hex_yield *p, const double *F0, const double *Fd, double *dA, double *da,
                                  double *G0, double *Gd, void *stream) {
    if (!p || !F0 || !Fd || !dA || !da || !G0 || !Gd) {
        hex::set_error("yield_gradient: null argument");
        return HEX_EINVAL;
    }
    if (!p->have_value) {
        hex::set_error("yield_gradient: hex_yield_value must run first on the same control and frames");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(p->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int M = p->M, N = p->N, S = N + 1, nu = p->t.nu;
    const int64_t vec = (int64_t)M * nu, per = (int64_t)M * M, fac = 4 * (int64_t)N;
    const Tables &t = p->t;
    // Volkov reverse: bar q and the phase-integral derivatives, then dY/dA_m
    const size_t ph = sizeof(zd) * t.nE * t.nD;
    hipLaunchKernelGGL(k_amplitude_reverse, dim3(S), dim3(kYThreads), ph, st, t, p->dt, p->q, p->ints, p->amp,
                       p->qbar, p->dints);
    hipLaunchKernelGGL(k_phase_reverse, dim3(1), dim3(64), 0, st, S, p->dt, p->A, p->dints, p->dA);
    // projection reverse: frames' adjoint sources, bar chi, frames' dY/dA_n
    hipLaunchKernelGGL(k_project_reverse, dim3(S), dim3(kYThreads), 0, st, t, p->chis,
                       reinterpret_cast<const zd *>(F0), reinterpret_cast<const zd *>(Fd), p->A, p->qbar,
                       reinterpret_cast<zd *>(G0), reinterpret_cast<zd *>(Gd), p->chibar, p->dA);
    HEX_HIP(hipGetLastError());
    // dual costate: mu_0 = bar chi_0, mu_i = bar chi_i + P_i^H mu_{i-1}
    const int64_t pzb = std::min<int64_t>((fac * M + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pz_apply, dim3((unsigned)pzb), dim3(256), 0, st, M, nu, fac, p->pz, p->zs, p->vs);
    const zd *w = p->chibar;   // mu_0
    for (int i = 1; i <= N; ++i) {
        for (int f = 3; f >= 0; --f) {
            const int64_t b = (int64_t)(i - 1) * 4 + f;
            zd *wout = p->mu + ((i * 4 + f) & 1) * vec;
            const zd *add = f == 0 ? p->chibar + (int64_t)i * vec : nullptr;
            hipLaunchKernelGGL(k_dual_reverse, dim3(M), dim3(64), 0, st, M, nu, p->minv + b * per, w, add,
                               p->ys + b * vec, wout);
            w = wout;
        }
    }
    HEX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_factor_dots, dim3((unsigned)fac), dim3(kYThreads), 0, st, (int)vec, p->ys, p->zs, p->vs,
                       p->dots);
    hipLaunchKernelGGL(k_stage_gradients, dim3((S + 63) / 64), dim3(64), 0, st, N, p->dt, p->a, p->dots, p->da);
    HEX_HIP(hipGetLastError());
    HEX_HIP(hipMemcpyAsync(dA, p->dA, sizeof(double) * S, hipMemcpyDeviceToHost, st));
    HEX_HIP(hipMemcpyAsync(da, p->da, sizeof(double) * 2 * S, hipMemcpyDeviceToHost, st));
    HEX_HIP(hipStreamSynchronize(st));
    return HEX_OK;
}
