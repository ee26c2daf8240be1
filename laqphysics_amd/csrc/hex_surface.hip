// tSURFF channel projection and Volkov accumulation on the device.
//
// contract:  q[j][o][s] = sum_{entries (c, a, u, w) of (j, o)} w *
//                         sum_r conj(chi[(a n + r) nu + u]) * frame[c][s][r]
//   (python/surface_contract.py:41-56; every contraction path of the reference
//   -- product channels, coupled fast path, sparse bipolar transform -- is an
//   entry list of this form, built on the host once)
// volkov:    radial[j,e,l] = sum_s q[j,l,s] tests[l,e,s]
//            current[j,e,d] = i sqrt2 sum_l radial[j,e,l] Y[l,d]
//                             * exp(i (E_e t + k_e (u_d . intA) + intA2 / 2))
//            amplitude += dt (previous + current) / 2;  previous = current
//   (python/streaming_surface.py:54-64, 97-109)
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include <algorithm>

__global__ void k_contract(int n, int ns, int nu, int nout, const int32_t *__restrict__ eptr,
                           const int32_t *__restrict__ ec, const int32_t *__restrict__ ea,
                           const int32_t *__restrict__ eu, const zd *__restrict__ ew, const zd *__restrict__ chi,
                           const zd *__restrict__ frame, zd *__restrict__ q) {
    // block = (s, j*nout + o); threads stride over r and entries
    const int s = blockIdx.x;
    const int jo = blockIdx.y;
    zd acc = zmake(0.0, 0.0);
    for (int k = eptr[jo]; k < eptr[jo + 1]; ++k) {
        const int c = ec[k], a = ea[k], u = eu[k];
        const zd w = ew[k];
        zd part = zmake(0.0, 0.0);
        for (int r = threadIdx.x; r < n; r += blockDim.x)
            part = zcfma(chi[((int64_t)a * n + r) * nu + u], frame[((int64_t)c * ns + s) * n + r], part);
        acc = zfma(w, part, acc);
    }
    __shared__ double red[4][2];
    double re = wave_sum(acc.x), im = wave_sum(acc.y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[wave][0] = re;
        red[wave][1] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            x += red[w][0];
            y += red[w][1];
        }
        q[(int64_t)jo * ns + s] = zmake(x, y);
    }
    (void)nout;
}

extern "C" int hex_surface_contract(int n, int ns, int nu, int nlab, int nout, const int32_t *eptr,
                                    const int32_t *ec, const int32_t *ea, const int32_t *eu, const double *ew,
                                    const double *chi, const double *frame, double *q, void *stream) {
    if (n < 1 || ns < 1 || nu < 1 || nlab < 1 || nout < 1 || !eptr || !chi || !frame || !q) {
        hex::set_error("surface_contract: bad argument");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(k_contract, dim3(ns, nlab * nout), dim3(256), 0, static_cast<hipStream_t>(stream), n, ns, nu,
                       nout, eptr, ec, ea, eu, reinterpret_cast<const zd *>(ew), reinterpret_cast<const zd *>(chi),
                       reinterpret_cast<const zd *>(frame), reinterpret_cast<zd *>(q));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

static constexpr int kMaxOuter = 128;

__global__ void k_volkov(int nlab, int nout, int ns, int nE, int nd, const zd *__restrict__ q,
                         const zd *__restrict__ tests, const zd *__restrict__ harm, const double *__restrict__ energy,
                         const double *__restrict__ kk, const double *__restrict__ unit, double time, double A0,
                         double A1, double A2, double intA2, double dt, int first, zd *__restrict__ previous,
                         zd *__restrict__ amplitude) {
    const int e = blockIdx.x, j = blockIdx.y;
    __shared__ zd radial[kMaxOuter];
    for (int l = threadIdx.x; l < nout; l += blockDim.x) {
        zd acc = zmake(0.0, 0.0);
        const zd *ql = q + ((int64_t)j * nout + l) * ns;
        const zd *tl = tests + ((int64_t)l * nE + e) * ns;
        for (int s = 0; s < ns; ++s) acc = zfma(ql[s], tl[s], acc);
        radial[l] = acc;
    }
    __syncthreads();
    const double root2 = 1.4142135623730951;
    for (int d = threadIdx.x; d < nd; d += blockDim.x) {
        zd sum = zmake(0.0, 0.0);
        for (int l = 0; l < nout; ++l) sum = zfma(radial[l], harm[(int64_t)l * nd + d], sum);
        const double *u = unit + 3 * d;
        const double disp = u[0] * A0 + u[1] * A1 + u[2] * A2;
        const double arg = energy[e] * time + kk[e] * disp + intA2 / 2;
        double sn, cs;
        sincos(arg, &sn, &cs);
        // i sqrt2 * sum * exp(i arg)
        const zd isum = zmake(-root2 * sum.y, root2 * sum.x);
        const zd cur = zmul(isum, zmake(cs, sn));
        const int64_t at = ((int64_t)j * nE + e) * nd + d;
        if (!first) {
            const zd p = previous[at];
            const zd a = amplitude[at];
            amplitude[at] = zmake(a.x + dt * (p.x + cur.x) / 2, a.y + dt * (p.y + cur.y) / 2);
        }
        previous[at] = cur;
    }
}

extern "C" int hex_volkov_add(int nlab, int nout, int ns, int nE, int nd, const double *q, const double *tests,
                              const double *harm, const double *energy, const double *k, const double *unit,
                              double time, const double intA[3], double intA2, double dt, int first,
                              double *previous, double *amplitude, void *stream) {
    if (nlab < 1 || nout < 1 || nout > kMaxOuter || ns < 1 || nE < 1 || nd < 1 || !q || !tests || !harm ||
        !energy || !k || !unit || !intA || !previous || !amplitude) {
        hex::set_error("volkov_add: bad argument");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(k_volkov, dim3(nE, nlab), dim3(64), 0, static_cast<hipStream_t>(stream), nlab, nout, ns, nE,
                       nd, reinterpret_cast<const zd *>(q), reinterpret_cast<const zd *>(tests),
                       reinterpret_cast<const zd *>(harm), energy, k, unit, time, intA[0], intA[1], intA[2], intA2,
                       dt, first, reinterpret_cast<zd *>(previous), reinterpret_cast<zd *>(amplitude));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}
