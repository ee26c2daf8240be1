// Small dense complex matrix exponential, usable on the host and in a kernel.
//
// Scaling and squaring with diagonal Pade approximants of degree 3, 5, 7, 9 or
// 13 chosen from the 1-norm (N. J. Higham, "The scaling and squaring method
// for the matrix exponential revisited", SIAM J. Matrix Anal. Appl. 26 (2005)
// 1179-1193, Algorithm 2.3: theta_m thresholds and Pade coefficients b_k from
// that paper).  Matrices are m x m, m <= kExpmMax, row-major, in caller-owned
// scratch (LDS in a kernel).  Written single-threaded: one lane of a
// workgroup runs it for one small Hessenberg matrix.
#pragma once
#include "hex_common.h"

static constexpr int kExpmMax = 64;

namespace hexm {

__host__ __device__ inline void matmul(int m, const zd *A, const zd *B, zd *C) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            zd acc = zmake(0.0, 0.0);
            for (int k = 0; k < m; ++k) acc = zfma(A[i * m + k], B[k * m + j], acc);
            C[i * m + j] = acc;
        }
}

__host__ __device__ inline double norm1(int m, const zd *A) {
    double best = 0.0;
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += sqrt(zabs2(A[i * m + j]));
        best = s > best ? s : best;
    }
    return best;
}

// Solve P X = Q in place (X overwrites Q), LU with partial pivoting; P destroyed.
__host__ __device__ inline void solve(int m, zd *P, zd *Q) {
    for (int k = 0; k < m; ++k) {
        int piv = k;
        double big = zabs2(P[k * m + k]);
        for (int i = k + 1; i < m; ++i) {
            double v = zabs2(P[i * m + k]);
            if (v > big) {
                big = v;
                piv = i;
            }
        }
        if (piv != k) {
            for (int j = 0; j < m; ++j) {
                zd t = P[k * m + j];
                P[k * m + j] = P[piv * m + j];
                P[piv * m + j] = t;
                t = Q[k * m + j];
                Q[k * m + j] = Q[piv * m + j];
                Q[piv * m + j] = t;
            }
        }
        const zd d = P[k * m + k];
        for (int i = k + 1; i < m; ++i) {
            const zd f = zdiv(P[i * m + k], d);
            if (f.x == 0.0 && f.y == 0.0) continue;
            for (int j = k; j < m; ++j) P[i * m + j] = zsub(P[i * m + j], zmul(f, P[k * m + j]));
            for (int j = 0; j < m; ++j) Q[i * m + j] = zsub(Q[i * m + j], zmul(f, Q[k * m + j]));
        }
    }
    for (int k = m - 1; k >= 0; --k) {
        const zd d = P[k * m + k];
        for (int j = 0; j < m; ++j) {
            zd acc = Q[k * m + j];
            for (int i = k + 1; i < m; ++i) acc = zsub(acc, zmul(P[k * m + i], Q[i * m + j]));
            Q[k * m + j] = zdiv(acc, d);
        }
    }
}

// theta_13 of the scaling step
static constexpr double kTheta13 = 5.371920351148152e0;

// Pade degree from the 1-norm: the first of 3, 5, 7, 9 whose theta_m bounds
// it, else 13.
__host__ __device__ __forceinline__ int pade_degree(double nrm) {
    if (nrm <= 1.495585217958292e-2) return 3;
    if (nrm <= 2.539398330063230e-1) return 5;
    if (nrm <= 9.504178996162932e-1) return 7;
    if (nrm <= 2.097847961257068e0) return 9;
    return 13;
}

// b_k of the degree-3, 5, 7, 9 or 13 approximant.  Every call passes a k known
// at compile time (after unrolling), so the tables fold into immediates and
// no private memory is used in a kernel.
__host__ __device__ __forceinline__ double pade_b(int degree, int k) {
    const double b3[4] = {120., 60., 12., 1.};
    const double b5[6] = {30240., 15120., 3360., 420., 30., 1.};
    const double b7[8] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
    const double b9[10] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
                           2162160., 110880., 3960., 90., 1.};
    const double b13[14] = {64764752532480000., 32382376266240000., 7771770303897600., 1187353796428800.,
                            129060195264000., 10559470521600., 670442572800., 33522128640.,
                            1323241920., 40840800., 960960., 16380., 182., 1.};
    switch (degree) {
    case 3: return k < 4 ? b3[k] : 0.0;
    case 5: return k < 6 ? b5[k] : 0.0;
    case 7: return k < 8 ? b7[k] : 0.0;
    case 9: return k < 10 ? b9[k] : 0.0;
    default: return k < 14 ? b13[k] : 0.0;
    }
}

// E = expm(A).  work must hold 6 * m * m entries.  A is not modified.
__host__ __device__ inline void expm(int m, const zd *A, zd *E, zd *work) {
    const int mm = m * m;
    zd *X = work, *A2 = work + mm, *A4 = work + 2 * mm, *A6 = work + 3 * mm, *U = work + 4 * mm,
       *V = work + 5 * mm;
    const double nrm = norm1(m, A);
    const int degree = hexm::pade_degree(nrm);
    int s = 0;
    double scale = 1.0;
    if (degree == 13 && nrm > hexm::kTheta13) {
        s = (int)ceil(log2(nrm / hexm::kTheta13));
        if (s < 0) s = 0;
        scale = ldexp(1.0, -s);
    }
    for (int i = 0; i < mm; ++i) X[i] = zscale(A[i], scale);
    matmul(m, X, X, A2);
    if (degree == 13) {
        matmul(m, A2, A2, A4);
        matmul(m, A4, A2, A6);
        // U = X [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I]
        // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
        for (int i = 0; i < mm; ++i) {
            E[i] = zadd(zadd(zscale(A6[i], hexm::pade_b(13, 13)), zscale(A4[i], hexm::pade_b(13, 11))), zscale(A2[i], hexm::pade_b(13, 9)));
            U[i] = zadd(zadd(zscale(A6[i], hexm::pade_b(13, 12)), zscale(A4[i], hexm::pade_b(13, 10))), zscale(A2[i], hexm::pade_b(13, 8)));
        }
        matmul(m, A6, E, V);   // V <- A6 (odd inner)
        for (int i = 0; i < mm; ++i) E[i] = V[i];
        matmul(m, A6, U, V);   // V <- A6 (even inner)
        for (int i = 0; i < mm; ++i) {
            const int r = i / m, c = i % m;
            const double id = r == c ? 1.0 : 0.0;
            U[i] = zadd(zadd(zadd(E[i], zscale(A6[i], hexm::pade_b(13, 7))), zadd(zscale(A4[i], hexm::pade_b(13, 5)), zscale(A2[i], hexm::pade_b(13, 3)))),
                        zmake(hexm::pade_b(13, 1) * id, 0.0));
            V[i] = zadd(zadd(zadd(V[i], zscale(A6[i], hexm::pade_b(13, 6))), zadd(zscale(A4[i], hexm::pade_b(13, 4)), zscale(A2[i], hexm::pade_b(13, 2)))),
                        zmake(hexm::pade_b(13, 0) * id, 0.0));
        }
        matmul(m, X, U, E);
        for (int i = 0; i < mm; ++i) U[i] = E[i];
    } else {
        // powers: A2^k accumulated; odd part Uo = sum_k b[2k+1] A2^k, even V = sum_k b[2k] A2^k
        for (int i = 0; i < mm; ++i) {
            const double id = (i / m == i % m) ? 1.0 : 0.0;
            U[i] = zmake(pade_b(degree, 1) * id, 0.0);
            V[i] = zmake(pade_b(degree, 0) * id, 0.0);
            A4[i] = zmake(id, 0.0);  // running power of A2
        }
        for (int k = 1; 2 * k <= degree; ++k) {
            matmul(m, A4, A2, A6);
            for (int i = 0; i < mm; ++i) {
                A4[i] = A6[i];
                V[i] = zadd(V[i], zscale(A4[i], pade_b(degree, 2 * k)));
                if (2 * k + 1 <= degree) U[i] = zadd(U[i], zscale(A4[i], pade_b(degree, 2 * k + 1)));
            }
        }
        matmul(m, X, U, E);
        for (int i = 0; i < mm; ++i) U[i] = E[i];
    }
    // (V - U) R = (V + U)
    for (int i = 0; i < mm; ++i) {
        A6[i] = zsub(V[i], U[i]);
        E[i] = zadd(V[i], U[i]);
    }
    solve(m, A6, E);
    for (int q = 0; q < s; ++q) {
        matmul(m, E, E, A6);
        for (int i = 0; i < mm; ++i) E[i] = A6[i];
    }
}

}  // namespace hexm

// The same algorithm run by one wavefront: entries of every matrix product and
// elementwise update are spread over the 64 lanes (each entry's k-sum keeps
// the serial order), the LU row updates likewise; results are identical to
// hexm::expm.  All 64 lanes of the calling wave must enter; matrices live in
// LDS (caller scratch, 6 m^2 entries of work).  The routines are inlined so
// that a caller's LDS pointers keep their address space (LDS instructions,
// not flat ones).
#if defined(__HIPCC__)
namespace hexw {

// the callers run these with a one-wave workgroup: a workgroup barrier orders the LDS traffic
__device__ __forceinline__ void sync() { __syncthreads(); }

__device__ __forceinline__ void matmul(int m, const zd *A, const zd *B, zd *C) {
    const int lane = threadIdx.x & 63;
    for (int e = lane; e < m * m; e += 64) {
        const int i = e / m, j = e - i * m;
        zd acc = zmake(0.0, 0.0);
        for (int k = 0; k < m; ++k) acc = zfma(A[i * m + k], B[k * m + j], acc);
        C[e] = acc;
    }
    sync();
}

// 1-norm: lane j sums column j in the serial order, then the wave takes the
// maximum (exact, so the same value as hexm::norm1; NaN sums never win).
__device__ __forceinline__ double norm1(int m, const zd *A) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    if (lane < m)
        for (int i = 0; i < m; ++i) s += sqrt(zabs2(A[i * m + lane]));
    if (!(s > 0.0)) s = 0.0;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(s, off, 64);
        s = o > s ? o : s;
    }
    return s;
}

// hexm::solve spread over the wave: the pivot by a wave argmax (the first row
// of largest |P[i][k]|^2, as the serial scan picks it), then every entry
// update of rows below the pivot by its own lane.  Each entry sees the same
// operations as in the serial code.
__device__ __forceinline__ void solve(int m, zd *P, zd *Q) {
    const int lane = threadIdx.x & 63;
    for (int k = 0; k < m; ++k) {
        double v = -1.0;
        int idx = 64;
        if (lane >= k && lane < m) {
            v = zabs2(P[lane * m + k]);
            // the serial scan keeps row k when its own entry is NaN, and never
            // moves to a NaN entry below it
            if (!(v == v)) v = lane == k ? __builtin_huge_val() : -1.0;
            idx = lane;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off, 64);
            const int oi = __shfl_xor(idx, off, 64);
            if (ov > v || (ov == v && oi < idx)) {
                v = ov;
                idx = oi;
            }
        }
        const int piv = idx;
        sync();
        if (piv != k)
            for (int j = lane; j < m; j += 64) {
                zd t = P[k * m + j];
                P[k * m + j] = P[piv * m + j];
                P[piv * m + j] = t;
                t = Q[k * m + j];
                Q[k * m + j] = Q[piv * m + j];
                Q[piv * m + j] = t;
            }
        sync();
        const zd d = P[k * m + k];
        // lane i holds the factor of row i; entry (i, j) of [P | Q] right of the
        // pivot column is updated by one lane, which reads the factor by shuffle
        zd f = zmake(0.0, 0.0);
        if (lane > k && lane < m) f = zdiv(P[lane * m + k], d);
        const int w = 2 * m - k, total = (m - k - 1) * w;
        for (int base = 0; base < total; base += 64) {   // the same trip count on every lane
            const int e = base + lane;
            const int i = k + 1 + (e < total ? e / w : 0), c = e < total ? e % w : 0;
            const double fx = __shfl(f.x, i, 64), fy = __shfl(f.y, i, 64);
            if (e < total && !(fx == 0.0 && fy == 0.0)) {
                const zd fi = zmake(fx, fy);
                if (c < m - k) {
                    const int j = k + c;
                    P[i * m + j] = zsub(P[i * m + j], zmul(fi, P[k * m + j]));
                } else {
                    const int j = c - (m - k);
                    Q[i * m + j] = zsub(Q[i * m + j], zmul(fi, Q[k * m + j]));
                }
            }
        }
        sync();
    }
    for (int k = m - 1; k >= 0; --k) {
        const zd d = P[k * m + k];
        for (int j = lane; j < m; j += 64) {
            zd acc = Q[k * m + j];
#pragma unroll 4
            for (int i = k + 1; i < m; ++i) acc = zsub(acc, zmul(P[k * m + i], Q[i * m + j]));
            Q[k * m + j] = zdiv(acc, d);
        }
        sync();
    }
}

__device__ __forceinline__ void expm(int m, const zd *A, zd *E, zd *work) {
    const int lane = threadIdx.x & 63;
    const int mm = m * m;
    zd *X = work, *A2 = work + mm, *A4 = work + 2 * mm, *A6 = work + 3 * mm, *U = work + 4 * mm,
       *V = work + 5 * mm;
    const double nrm = norm1(m, A);
    const int degree = hexm::pade_degree(nrm);
    int s = 0;
    double scale = 1.0;
    if (degree == 13 && nrm > hexm::kTheta13) {
        s = (int)ceil(log2(nrm / hexm::kTheta13));
        if (s < 0) s = 0;
        scale = ldexp(1.0, -s);
    }
    for (int i = lane; i < mm; i += 64) X[i] = zscale(A[i], scale);
    sync();
    matmul(m, X, X, A2);
    if (degree == 13) {
        matmul(m, A2, A2, A4);
        matmul(m, A4, A2, A6);
        for (int i = lane; i < mm; i += 64) {
            E[i] = zadd(zadd(zscale(A6[i], hexm::pade_b(13, 13)), zscale(A4[i], hexm::pade_b(13, 11))), zscale(A2[i], hexm::pade_b(13, 9)));
            U[i] = zadd(zadd(zscale(A6[i], hexm::pade_b(13, 12)), zscale(A4[i], hexm::pade_b(13, 10))), zscale(A2[i], hexm::pade_b(13, 8)));
        }
        sync();
        matmul(m, A6, E, V);
        for (int i = lane; i < mm; i += 64) E[i] = V[i];
        sync();
        matmul(m, A6, U, V);
        for (int i = lane; i < mm; i += 64) {
            const int r = i / m, c = i % m;
            const double id = r == c ? 1.0 : 0.0;
            U[i] = zadd(zadd(zadd(E[i], zscale(A6[i], hexm::pade_b(13, 7))), zadd(zscale(A4[i], hexm::pade_b(13, 5)), zscale(A2[i], hexm::pade_b(13, 3)))),
                        zmake(hexm::pade_b(13, 1) * id, 0.0));
            V[i] = zadd(zadd(zadd(V[i], zscale(A6[i], hexm::pade_b(13, 6))), zadd(zscale(A4[i], hexm::pade_b(13, 4)), zscale(A2[i], hexm::pade_b(13, 2)))),
                        zmake(hexm::pade_b(13, 0) * id, 0.0));
        }
        sync();
        matmul(m, X, U, E);
        for (int i = lane; i < mm; i += 64) U[i] = E[i];
        sync();
    } else {
        const double b0 = hexm::pade_b(degree, 0), b1 = hexm::pade_b(degree, 1);
        for (int i = lane; i < mm; i += 64) {
            const double id = (i / m == i % m) ? 1.0 : 0.0;
            U[i] = zmake(b1 * id, 0.0);
            V[i] = zmake(b0 * id, 0.0);
            A4[i] = zmake(id, 0.0);
        }
        sync();
#pragma unroll
        for (int k = 1; k <= 4; ++k) {
            if (2 * k > degree) break;
            matmul(m, A4, A2, A6);
            const double be = hexm::pade_b(degree, 2 * k), bo = hexm::pade_b(degree, 2 * k + 1);
            for (int i = lane; i < mm; i += 64) {
                A4[i] = A6[i];
                V[i] = zadd(V[i], zscale(A4[i], be));
                if (2 * k + 1 <= degree) U[i] = zadd(U[i], zscale(A4[i], bo));
            }
            sync();
        }
        matmul(m, X, U, E);
        for (int i = lane; i < mm; i += 64) U[i] = E[i];
        sync();
    }
    for (int i = lane; i < mm; i += 64) {
        A6[i] = zsub(V[i], U[i]);
        E[i] = zadd(V[i], U[i]);
    }
    sync();
    solve(m, A6, E);
    for (int q = 0; q < s; ++q) {
        matmul(m, E, E, A6);
        for (int i = lane; i < mm; i += 64) E[i] = A6[i];
        sync();
    }
}

}  // namespace hexw
#endif
