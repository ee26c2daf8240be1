// H.psi in two passes: radial walks, then channel coupling per Coulomb group.
//
//   k_radial   one thread per (i, j, c): out = diag psi_c + T_i psi_c + T_j psi_c
//              and, when a field couples, the two electrons' (T R - R T) walks
//              Y1 = B_i psi_c, Y2 = B_j psi_c into scratch.  Lanes run along i,
//              so the i walk reads neighbouring lanes' rows (L1) and every j
//              walk step is one coalesced 1 KiB row.
//   k_angular  one thread per (i, j) and Coulomb group G (channels linked by
//              any multipole edge): psi_G(i, j) is staged in LDS once, then for
//              every output c in G: out_c += sum_edges vcoef rad_lam psi_b
//              + dipole edges (velocity: i coef [Y_e(s) + ldiff/(2 r_e) psi_s];
//              length: coef r_e psi_s) + |A|^2 psi_c, optionally out = a out + b psi.
// The sum is the one of the reference kernel (src/kernels.f90:130-175) with the
// velocity dipole regrouped as sum_q T_iq (r_q - r_i) psi_s(q) = (B psi_s)_i.
// The radial pass comes in three forms chosen by the launch shape: direct walks
// (k_radial), walks through an LDS-staged tile (k_radial_staged) and FP64
// matrix-core walks (k_radial_mfma).  The matrix-core form may instead apply B
// to each output channel's mixed coupling sources, sum_k i coef_k B_e psi_s(k)
// = B_e (sum_k i coef_k psi_s(k)), adding straight into out, so Y1/Y2 are not
// formed and the coupling pass adds only the pointwise part of those edges.
#include "hex_internal.h"
#include "../../include/laqphysics_hip.h"

template <bool ADJ>
__device__ __forceinline__ zd cz(zd a) { return ADJ ? zconj(a) : a; }

template <bool ADJ, bool BASE, bool DIPOLE, int BY>
__global__ __launch_bounds__(kApplyBX *BY) void k_radial(EngineView e, BatchArgs bt, int64_t ystride,
                                                         const zd *__restrict__ psi, zd *__restrict__ out,
                                                         zd *__restrict__ y1, zd *__restrict__ y2) {
    const int nc = e.nc;
    const int member = blockIdx.z / nc, c = blockIdx.z - member * nc;
    if (bt.active && !bt.active[member]) return;
    const int i = blockIdx.x * kApplyBX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int n = e.n;
    if (i >= n || j >= n) return;
    const int64_t nn = (int64_t)n * n;
    const int64_t at = i + (int64_t)n * j;
    const zd *pc = psi + bt.in_stride * member + nn * c;
    zd val = zmake(0.0, 0.0), a1 = zmake(0.0, 0.0), a2 = zmake(0.0, 0.0);
    if (BASE) val = zmul(cz<ADJ>(zadd(e.d1[i + (int64_t)n * c], e.d2[j + (int64_t)n * c])), pc[at]);
    const int pi0 = e.ptr[i], pi1 = e.ptr[i + 1];
    for (int k = pi0; k < pi1; ++k) {
        const zd x = pc[e.col[k] + (int64_t)n * j];
        if (BASE) val = zfma(cz<ADJ>(e.tv[k]), x, val);
        if (DIPOLE) a1 = zfma(cz<ADJ>(e.tb[k]), x, a1);
    }
    const int pj0 = e.ptr[j], pj1 = e.ptr[j + 1];
    for (int k = pj0; k < pj1; ++k) {
        const zd x = pc[i + (int64_t)n * e.col[k]];
        if (BASE) val = zfma(cz<ADJ>(e.tv[k]), x, val);
        if (DIPOLE) a2 = zfma(cz<ADJ>(e.tb[k]), x, a2);
    }
    out[bt.out_stride * member + nn * c + at] = val;
    if (DIPOLE) {
        y1[ystride * member + nn * c + at] = a1;
        y2[ystride * member + nn * c + at] = a2;
    }
}

// Staged radial walks.  A block covers 64 rows i x 16 rows j (4 per wave) of
// one channel slab: the slab's columns [iblo, ibhi) that the 64 i rows touch
// are staged in LDS for the i walk (lanes of one element read the same LDS
// words), and each wave walks the union of its 4 j rows' bands once, every
// coalesced 1 KiB row load feeding up to 4 outputs with wave-uniform band
// coefficients.
template <bool ADJ, bool BASE, bool DIPOLE>
__global__ __launch_bounds__(kApplyBX * 4) void k_radial_staged(EngineView e, BatchArgs bt, int64_t ystride,
                                                                const zd *__restrict__ psi, zd *__restrict__ out,
                                                                zd *__restrict__ y1, zd *__restrict__ y2) {
    extern __shared__ zd Ls[];   // [4 kStageJ][stage_width]
    const int sw = e.stage_width;
    const int nc = e.nc;
    const int n = e.n;
    // logical block: j block fastest (consecutive blocks share j-walk rows),
    // then i block, then (member, channel); XCD-contiguous numbering
    const int nxb = (n + kApplyBX - 1) / kApplyBX, nyb = (n + 4 * kStageJ - 1) / (4 * kStageJ);
    int64_t L = xcd_contiguous(blockIdx.x, (int64_t)gridDim.x);
    const int by = (int)(L % nyb);
    L /= nyb;
    const int bx = (int)(L % nxb);
    const int mc = (int)(L / nxb);
    const int member = mc / nc, c = mc - member * nc;
    if (bt.active && !bt.active[member]) return;
    // keep every lane of a wave on the same path (no divergent exits), so the
    // compiler sees the band walk and its coefficients as wave-uniform
    // (scalar loads); lanes past n compute on row n-1 and store nothing
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int i = bx * kApplyBX + lane;
    const bool live = i < n;
    const int ic = live ? i : n - 1;
    const int j0 = by * (4 * kStageJ);
    const int64_t nn = (int64_t)n * n;
    const zd *pc = psi + bt.in_stride * member + nn * c;
    const int ilo = ro(e.iblo)[bx], width = ro(e.ibhi)[bx] - ilo;
    const int rows = min(4 * kStageJ, n - j0);
    // eight loads in flight per thread before their LDS writes
    for (int idx0 = lane + kApplyBX * w; idx0 < rows * width; idx0 += 8 * kApplyBX * 4) {
        zd v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = min(idx0 + u * kApplyBX * 4, rows * width - 1);
            const int jj = idx / width, ii = idx - jj * width;
            v[u] = pc[(ilo + ii) + (int64_t)n * (j0 + jj)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = idx0 + u * kApplyBX * 4;
            if (idx < rows * width) {
                const int jj = idx / width, ii = idx - jj * width;
                Ls[jj * sw + ii] = v[u];
            }
        }
    }
    __syncthreads();
    const int jw = j0 + kStageJ * w;   // first j of this wave (wave-uniform)
    if (jw >= n) return;
    const int nq = min(kStageJ, n - jw);
    zd val[kStageJ], a1[kStageJ], a2[kStageJ];
#pragma unroll
    for (int q = 0; q < kStageJ; ++q) {
        val[q] = zmake(0.0, 0.0);
        a1[q] = zmake(0.0, 0.0);
        a2[q] = zmake(0.0, 0.0);
    }
    // i walk from LDS over the block's widest band, whole groups of four
    // offsets (the band tables are zero-padded to a multiple of four); entries
    // past a row's own band are zero and read a clamped LDS column.  Every
    // wave computes all kStageJ rows (rows past n read unused LDS) and stores
    // only the rows below n, so the walks carry no per-offset or per-row tests.
    // Blocks whose rows lie on the real grid (T and T R - R T real) take the
    // real-coefficient path: two FMAs per term instead of four.
    const zd *Lw = Ls + kStageJ * w * sw;
    {
        const int lo_i = ro(e.jlo)[ic] - ilo;
        const int bw = ro(e.ibw)[bx];
        if (ro(e.ireal)[bx]) {
            const double *tT = e.tbandTr + ic;
            const double *bT = e.bbandTr + ic;
            for (int off0 = 0; off0 < bw; off0 += 4) {
                double t[4], bq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    t[u] = BASE ? tT[(off0 + u) * n] : 0.0;
                    bq[u] = DIPOLE ? bT[(off0 + u) * n] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int col = min(lo_i + off0 + u, width - 1);
#pragma unroll
                    for (int q = 0; q < kStageJ; ++q) {
                        const zd x = Lw[q * sw + col];
                        if (BASE) val[q] = zmake(fma(t[u], x.x, val[q].x), fma(t[u], x.y, val[q].y));
                        if (DIPOLE) a1[q] = zmake(fma(bq[u], x.x, a1[q].x), fma(bq[u], x.y, a1[q].y));
                    }
                }
            }
        } else {
            const zd *tT = e.tbandT + ic;
            const zd *bT = e.bbandT + ic;
            for (int off0 = 0; off0 < bw; off0 += 4) {
                // coefficients of four band offsets in flight before any is used
                zd t[4], bq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    t[u] = BASE ? cz<ADJ>(tT[(off0 + u) * n]) : zmake(0.0, 0.0);
                    bq[u] = DIPOLE ? cz<ADJ>(bT[(off0 + u) * n]) : zmake(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int col = min(lo_i + off0 + u, width - 1);
#pragma unroll
                    for (int q = 0; q < kStageJ; ++q) {
                        const zd x = Lw[q * sw + col];
                        if (BASE) val[q] = zfma(t[u], x, val[q]);
                        if (DIPOLE) a1[q] = zfma(bq[u], x, a1[q]);
                    }
                }
            }
        }
    }
    // j walk over the union of the wave's four bands: per row group a dense,
    // zero-padded coefficient block [kk][q] (wave-uniform scalar loads), whole
    // groups of four rows (rows past n are clamped and carry zero coefficients);
    // four rows are loaded before any is used
    {
        const int g = jw / kStageJ;
        const int klo = ro(e.jg_lo)[g], kw = ro(e.jg_w)[g];
        const int64_t gbase = (int64_t)g * e.jg_max;
        const zd *pcol = pc + ic;
        if (ro(e.jg_real)[g]) {
            const auto *ct = ro(e.jg_tr) + gbase * kStageJ;
            const auto *cb = ro(e.jg_br) + gbase * kStageJ;
            for (int k0 = 0; k0 < kw; k0 += 4) {
                zd xs[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) xs[u] = pcol[(int64_t)n * min(klo + k0 + u, n - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int at = (k0 + u) * kStageJ;
#pragma unroll
                    for (int q = 0; q < kStageJ; ++q) {
                        if (BASE) {
                            const double t = ct[at + q];
                            val[q] = zmake(fma(t, xs[u].x, val[q].x), fma(t, xs[u].y, val[q].y));
                        }
                        if (DIPOLE) {
                            const double t = cb[at + q];
                            a2[q] = zmake(fma(t, xs[u].x, a2[q].x), fma(t, xs[u].y, a2[q].y));
                        }
                    }
                }
            }
        } else {
            const auto *ct = ro(e.jg_t) + gbase * kStageJ;
            const auto *cb = ro(e.jg_b) + gbase * kStageJ;
            for (int k0 = 0; k0 < kw; k0 += 4) {
                zd xs[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) xs[u] = pcol[(int64_t)n * min(klo + k0 + u, n - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int at = (k0 + u) * kStageJ;
#pragma unroll
                    for (int q = 0; q < kStageJ; ++q) {
                        if (BASE) val[q] = zfma(cz<ADJ>(ldz(ct + at + q)), xs[u], val[q]);
                        if (DIPOLE) a2[q] = zfma(cz<ADJ>(ldz(cb + at + q)), xs[u], a2[q]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kStageJ; ++q) {
        if (q < nq) {
            const int j = jw + q;
            const int64_t at = ic + (int64_t)n * j;
            zd v = val[q];
            if (BASE) {
                const zd dg = cz<ADJ>(zadd(e.d1[ic + (int64_t)n * c], e.d2[j + (int64_t)n * c]));
                v = zadd(v, zmul(dg, Lw[q * sw + (ic - ilo)]));
            }
            if (live) {
                st_stream(out + bt.out_stride * member + nn * c + at, v);
                if (DIPOLE) {
                    st_stream(y1 + ystride * member + nn * c + at, a1[q]);
                    st_stream(y2 + ystride * member + nn * c + at, a2[q]);
                }
            }
        }
    }
}

// Radial walks on the matrix cores.  One wave owns a 16 x 16 tile of points
// (i block bi, j block bj) of one channel slab; both walks are FP64 GEMMs that
// accumulate into the same 16 x 16 result D[m = j - j0][n = i - i0]:
//   i walk: D += Psi^T(j rows, band of bi) * T^T(band of bi, i rows)
//   j walk: D += T(j rows, band of bj) * Psi(band of bj, i rows)
// so the coefficient tile of block b serves as the B operand of one walk and
// the A operand of the other (the layouts coincide).  Complex values enter as
// separate real and imaginary products; real blocks need two per step and
// complex (irECS) blocks four.  The T R - R T walks go to Y1 (i) and Y2 (j),
// or with the mixed sources into the result itself.  A workgroup is 4 waves
// on consecutive i blocks of one j block.
typedef double hex_v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ hex_v4d mfma4(double a, double b, hex_v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One walk of a 16 x 16 tile: steps of four band columns, four steps per
// round with all their operand loads issued before the first product (the
// step counts are padded to a multiple of four with zero tiles).  IWALK: the
// state slice is the A operand and the coefficient tile the B operand (i
// walk); otherwise the roles swap (j walk).  `line` points at the lane's row
// (i walk) or column (j walk) of the slab, `stride` is the distance between
// consecutive band columns along it.  The T R - R T products go to (yr, yi)
// -- or, with ZMIX, act on the mixed sources Z = sum_q zc[q] psi_(zoff[q])
// (the field-coupling sources of this electron, zoff their slab offsets from
// `line`, q < nz) and add into (tr, ti).
template <bool ADJ, bool BASE, bool DIPOLE, bool IWALK, bool ZMIX>
__device__ __forceinline__ void mfma_walk(const EngineView &e, const zd *__restrict__ line, int64_t stride, int blk,
                                          int lane, hex_v4d &tr, hex_v4d &ti, hex_v4d &yr, hex_v4d &yi,
                                          const int64_t *zoff, const zd *zc, int nz) {
    const int n = e.n, kq = lane >> 4;
    const int ks = ro(e.kst)[blk], ns = ro(e.kns)[blk];
    const bool re = ro(e.kreal)[blk];
    const double sg = ADJ ? -1.0 : 1.0;   // conjugate coefficients for the adjoint
    const double *mt = e.mt + (int64_t)ro(e.koff)[blk] * 64 + lane;
    const double *mti = e.mti + (int64_t)ro(e.koff)[blk] * 64 + lane;
    const double *mb = e.mb + (int64_t)ro(e.koff)[blk] * 64 + lane;
    const double *mbi = e.mbi + (int64_t)ro(e.koff)[blk] * 64 + lane;
    for (int s0 = 0; s0 < ns; s0 += 4) {
        zd p[4], zq[4];
        double t[4], tm[4], b[4], bm[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int s = s0 + u;
            const int64_t at = stride * min(ks + 4 * s + kq, n - 1);
            p[u] = line[at];
            t[u] = BASE ? mt[64 * s] : 0.0;
            b[u] = DIPOLE ? mb[64 * s] : 0.0;
            tm[u] = (BASE && !re) ? sg * mti[64 * s] : 0.0;
            bm[u] = (DIPOLE && !re) ? sg * mbi[64 * s] : 0.0;
            zq[u] = zmake(0.0, 0.0);
        }
        if (ZMIX && DIPOLE) {
            zd x[kZSrc][4];
#pragma unroll
            for (int q = 0; q < kZSrc; ++q)
                if (q < nz) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[q][u] = line[zoff[q] + stride * min(ks + 4 * (s0 + u) + kq, n - 1)];
                }
#pragma unroll
            for (int q = 0; q < kZSrc; ++q)
                if (q < nz) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) zq[u] = zfma(zc[q], x[q][u], zq[u]);
                }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // operands of the T R - R T products: the slab itself or the mixed sources
            const zd d = ZMIX ? zq[u] : p[u];
            hex_v4d &ar = ZMIX ? tr : yr;
            hex_v4d &ai = ZMIX ? ti : yi;
            if (IWALK) {
                if (BASE) {
                    tr = mfma4(p[u].x, t[u], tr);
                    ti = mfma4(p[u].y, t[u], ti);
                }
                if (DIPOLE) {
                    ar = mfma4(d.x, b[u], ar);
                    ai = mfma4(d.y, b[u], ai);
                }
                if (!re) {
                    if (BASE) {
                        tr = mfma4(-p[u].y, tm[u], tr);
                        ti = mfma4(p[u].x, tm[u], ti);
                    }
                    if (DIPOLE) {
                        ar = mfma4(-d.y, bm[u], ar);
                        ai = mfma4(d.x, bm[u], ai);
                    }
                }
            } else {
                if (BASE) {
                    tr = mfma4(t[u], p[u].x, tr);
                    ti = mfma4(t[u], p[u].y, ti);
                }
                if (DIPOLE) {
                    ar = mfma4(b[u], d.x, ar);
                    ai = mfma4(b[u], d.y, ai);
                }
                if (!re) {
                    if (BASE) {
                        tr = mfma4(-tm[u], p[u].y, tr);
                        ti = mfma4(tm[u], p[u].x, ti);
                    }
                    if (DIPOLE) {
                        ar = mfma4(-bm[u], d.y, ar);
                        ai = mfma4(bm[u], d.x, ai);
                    }
                }
            }
        }
    }
}

// The field-coupling sources of output channel c for electron el (0, 1):
// slab offsets from channel c and the coefficients i coef (velocity gauge).
__device__ __forceinline__ int zsources(const EngineView &e, const zd *dcz, int c, int el, int64_t nn,
                                        int64_t *zoff, zd *zc) {
    int nz = 0;
#pragma unroll
    for (int q = 0; q < kZSrc; ++q) {
        const int k = ro(e.zsrc)[(2 * c + el) * kZSrc + q];
        zoff[q] = 0;
        zc[q] = zmake(0.0, 0.0);
        if (k >= 0) {
            const zd coef = ldz(ro(dcz) + k);
            zoff[q] = nn * (ro(e.dc)[k] - c);
            zc[q] = zmake(-coef.y, coef.x);
            nz = q + 1;
        }
    }
    return nz;
}

template <bool ADJ, bool BASE, bool DIPOLE, bool ZMIX>
__global__ __launch_bounds__(256) void k_radial_mfma(EngineView e, BatchArgs bt, int64_t ystride,
                                                     const zd *__restrict__ psi, zd *__restrict__ out,
                                                     zd *__restrict__ y1, zd *__restrict__ y2,
                                                     const zd *__restrict__ dcz) {
    const int n = e.n, nc = e.nc, nb = e.nb16;
    const int ngrp = (nb + 3) / 4;
    // logical block: without mixing, j block fastest (neighbouring j walks
    // share rows in L2), then the group of four i blocks, then (member,
    // channel); with mixing, channel fastest, so the field-coupling sources
    // of a tile are read while their own blocks have them in L2
    int64_t L = xcd_contiguous(blockIdx.x, (int64_t)gridDim.x);
    int bj, grp, member, c;
    if (ZMIX) {
        c = (int)(L % nc);
        L /= nc;
        bj = (int)(L % nb);
        L /= nb;
        grp = (int)(L % ngrp);
        member = (int)(L / ngrp);
    } else {
        bj = (int)(L % nb);
        L /= nb;
        grp = (int)(L % ngrp);
        const int mc = (int)(L / ngrp);
        member = mc / nc;
        c = mc - member * nc;
    }
    if (bt.active && !bt.active[member]) return;
    const int bi = 4 * grp + __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (bi >= nb) return;   // whole wave; the kernel has no barrier
    const int lane = threadIdx.x, col = lane & 15, kq = lane >> 4;
    const int i0 = 16 * bi, j0 = 16 * bj;
    const int64_t nn = (int64_t)n * n;
    const zd *pc = psi + bt.in_stride * member + nn * c;
    const hex_v4d zero = {0.0, 0.0, 0.0, 0.0};
    hex_v4d tr = zero, ti = zero, y1r = zero, y1i = zero, y2r = zero, y2i = zero;
    int64_t zoff[kZSrc];
    zd zc[kZSrc];
    int nz = 0;
    const zd *dm = ZMIX ? dcz + (int64_t)member * e.nd : nullptr;
    // i walk: A = Psi^T (lane: row j0 + col, column kst + 4 s + kq), B = tile of bi
    if (ZMIX && DIPOLE) nz = zsources(e, dm, c, 0, nn, zoff, zc);
    mfma_walk<ADJ, BASE, DIPOLE, true, ZMIX>(e, pc + (int64_t)n * min(j0 + col, n - 1), 1, bi, lane, tr, ti, y1r,
                                             y1i, zoff, zc, nz);
    // j walk: A = tile of bj, B = Psi (lane: row i0 + col, column kst + 4 s + kq)
    if (ZMIX && DIPOLE) nz = zsources(e, dm, c, 1, nn, zoff, zc);
    mfma_walk<ADJ, BASE, DIPOLE, false, ZMIX>(e, pc + min(i0 + col, n - 1), n, bj, lane, tr, ti, y2r, y2i, zoff, zc,
                                              nz);
    // register r of lane (kq, col) holds D[4 r + kq][col]: the point
    // i = i0 + col, j = j0 + 4 r + kq
    const int i = i0 + col;
    if (i >= n) return;
    const int64_t ob = bt.out_stride * member + nn * c, yb = ystride * member + nn * c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = j0 + 4 * r + kq;
        if (j < n) {
            const int64_t at = i + (int64_t)n * j;
            zd v = zmake(tr[r], ti[r]);
            if (BASE) v = zfma(cz<ADJ>(zadd(e.d1[i + (int64_t)n * c], e.d2[j + (int64_t)n * c])), pc[at], v);
            st_stream(out + ob + at, v);
            if (DIPOLE && !ZMIX) {
                st_stream(y1 + yb + at, zmake(y1r[r], y1i[r]));
                st_stream(y2 + yb + at, zmake(y2r[r], y2i[r]));
            }
        }
    }
}

template <bool ADJ, bool BASE, bool DIPOLE, bool ZMIX = false>
static void radial_mfma(int members, hipStream_t st, const EngineView &v, BatchArgs bt, int64_t ys, const zd *psi,
                        zd *out, zd *y1, zd *y2, const zd *dcz = nullptr) {
    const int64_t blocks = (int64_t)v.nb16 * ((v.nb16 + 3) / 4) * members * v.nc;
    hipLaunchKernelGGL((k_radial_mfma<ADJ, BASE, DIPOLE, ZMIX>), dim3((unsigned)blocks), dim3(64, 4), 0, st, v, bt,
                       ys, psi, out, y1, y2, dcz);
}

// Field-coupling coefficients of this apply, once per (member, edge):
// coef = sum_a field_a dcoef(a, k) (src/kernels.f90:145), set to exactly zero
// where the reference skips the edge (|coef| < 1e-30).
__global__ void k_dipole_coefs(EngineView e, double f0, double f1, double f2, BatchArgs bt, int members,
                               zd *__restrict__ dcz) {
    const int64_t total = (int64_t)members * e.nd;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int member = (int)(x / e.nd), k = (int)(x - (int64_t)member * e.nd);
        double g0 = f0, g1 = f1, g2 = f2;
        if (bt.fields) {
            g0 = bt.fields[3 * member];
            g1 = bt.fields[3 * member + 1];
            g2 = bt.fields[3 * member + 2];
        }
        const zd *dk = e.dcoef + 3 * k;
        zd coef = zmake(g0 * dk[0].x + g1 * dk[1].x + g2 * dk[2].x, g0 * dk[0].y + g1 * dk[1].y + g2 * dk[2].y);
        if (sqrt(zabs2(coef)) < 1e-30) coef = zmake(0.0, 0.0);
        dcz[x] = coef;
    }
}

// Mixed-gauge coefficients of one apply from the 12 field coefficients
// c = [A, E, A_a A_b]: per dipole edge sum_a A_a dcoef(a, k) (dcz) and
// sum_a E_a dcoef(a, k) (dczE); per angular-square edge
// 1/2 sum_m c[6 + m] scoef(m, k) (dczQ).  No edge is skipped
// (python/mixed_gauge.py:96-104 combines them as they are).
struct Coef12 {
    double c[12];
};

__global__ void k_mixed_coefs(EngineView e, Coef12 f, zd *__restrict__ dcz) {
    const int64_t total = (int64_t)e.nd + e.nsq;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        if (x < e.nd) {
            const zd *dk = e.dcoef + 3 * x;
            dcz[x] = zmake(f.c[0] * dk[0].x + f.c[1] * dk[1].x + f.c[2] * dk[2].x,
                           f.c[0] * dk[0].y + f.c[1] * dk[1].y + f.c[2] * dk[2].y);
            e.dczE[x] = zmake(f.c[3] * dk[0].x + f.c[4] * dk[1].x + f.c[5] * dk[2].x,
                              f.c[3] * dk[0].y + f.c[4] * dk[1].y + f.c[5] * dk[2].y);
        } else {
            const int64_t k = x - e.nd;
            const zd *sk = e.scoef + 6 * k;
            double re = 0.0, im = 0.0;
            for (int m = 0; m < 6; ++m) {
                re += f.c[6 + m] * sk[m].x;
                im += f.c[6 + m] * sk[m].y;
            }
            e.dczQ[k] = zmake(0.5 * re, 0.5 * im);
        }
    }
}

// MIXED: the mixed gauge (python/mixed_gauge.py:96-117) -- the dipole edges take
// q/(2 r^2) for 1/(2 r) (their walks Y were built from q), add the E coupling
// coefE F_e psi_s, the angular-square edges add dczQ c2_e psi_s and the
// diagonal 1/2 |A|^2 (c0_i + c0_j) replaces the velocity gauge's |A|^2.
// With the mixed sources (no Y reads) the pass is held to 64 VGPRs, 8 waves
// per SIMD, which measures faster; the other forms keep the compiler's choice.
template <bool VELOCITY, bool ADJ, bool MIXED = false, bool ZMIX = false>
__global__ __launch_bounds__(kApplyBX *kAngWaves, ZMIX ? 8 : 1) void k_angular(EngineView e, double f0, double f1, double f2, zd ca,
                                                                 zd cb, int affine, int terms, BatchArgs bt,
                                                                 int64_t ystride, const zd *__restrict__ psi,
                                                                 zd *__restrict__ out, const zd *__restrict__ y1,
                                                                 const zd *__restrict__ y2, const zd *dcz) {
    extern __shared__ zd lds[];
    const int ng = e.ngroups;
    const int n = e.n;
    // logical block: group fastest (the groups of one tile read each other's
    // channels through the field coupling), then i tile, then j, then member;
    // XCD-contiguous numbering
    const int nxb = (n + kApplyBX - 1) / kApplyBX;
    int64_t L = xcd_contiguous(blockIdx.x, (int64_t)gridDim.x);
    const int grp = (int)(L % ng);
    L /= ng;
    const int bx = (int)(L % nxb);
    L /= nxb;
    const int j = (int)(L % n);
    const int member = (int)(L / n);
    if (bt.active && !bt.active[member]) return;
    if (bt.fields) {
        f0 = bt.fields[3 * member];
        f1 = bt.fields[3 * member + 1];
        f2 = bt.fields[3 * member + 2];
    }
    // no divergent exits: lanes past n work on row n-1 and store nothing, so
    // the edge loops stay wave-uniform (scalar loads of the edge records)
    const int lane = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int i = bx * kApplyBX + lane;
    const bool live = i < n;
    const int ic = live ? i : n - 1;
    const int64_t nn = (int64_t)n * n;
    const int64_t at = ic + (int64_t)n * j;
    psi += bt.in_stride * member;
    out += bt.out_stride * member;
    const zd *Y1 = y1 + ystride * member;
    const zd *Y2 = y2 + ystride * member;
    const int g0 = ro(e.gptr)[grp], g = ro(e.gptr)[grp + 1] - g0;
    zd *ps = lds;                          // [gmax][64]
    zd *rad = lds + e.gmax * kApplyBX;     // [lam_count][64]

    const bool base = terms & kTermBase;
    const zd ri = cz<ADJ>(e.r[ic]);
    const zd rj = cz<ADJ>(e.r[j]);
    const double cc = e.cut[ic] * e.cut[j];
    // every wave of the block covers the same 64 points: the multipoles of
    // the tile are real when both radii of every point are real
    const bool realrad = e.vreal && __syncthreads_and(ri.y == 0.0 && rj.y == 0.0);
    const bool coulomb = base && __syncthreads_or(live && cc != 0.0);
    double *radr = reinterpret_cast<double *>(rad);
    if (coulomb && wave == 0) {
        // r_<^lam / r_>^(lam+1) cut_i cut_j by the recurrence rad_lam = rad_(lam-1) t,
        // t = r_< / r_>: two divisions per point instead of one per order
        if (realrad) {
            const double lo = i < j ? ri.x : rj.x;
            const double hi = i < j ? rj.x : ri.x;
            const double t = lo / hi;
            double v = cc / hi;
            for (int lam = 0; lam < e.lam_count; ++lam) {
                radr[lam * kApplyBX + lane] = v;
                v *= t;
            }
        } else {
            const zd lo = i < j ? ri : rj;
            const zd hi = i < j ? rj : ri;
            const zd t = zdiv(lo, hi);
            zd v = zscale(zdiv(zmake(1.0, 0.0), hi), cc);
            for (int lam = 0; lam < e.lam_count; ++lam) {
                rad[lam * kApplyBX + lane] = v;
                v = zmul(v, t);
            }
        }
    }
    // the group's values, eight loads in flight per lane before their LDS writes
    for (int t0 = wave; t0 < g; t0 += 8 * kAngWaves) {
        zd v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = psi[nn * ro(e.gch)[g0 + min(t0 + u * kAngWaves, g - 1)] + at];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (t0 + u * kAngWaves < g) ps[(t0 + u * kAngWaves) * kApplyBX + lane] = v[u];
    }
    __syncthreads();
    const double a2 = f0 * f0 + f1 * f1 + f2 * f2;
    const bool dipole = MIXED ? e.nd > 0 : (terms & kTermDipole) && (f0 != 0.0 || f1 != 0.0 || f2 != 0.0);
    const zd half_ri = MIXED ? e.hq[ic] : zdiv(zmake(1.0, 0.0), zscale(ri, 2.0));
    const zd half_rj = MIXED ? e.hq[j] : zdiv(zmake(1.0, 0.0), zscale(rj, 2.0));
    zd Fi = zmake(0.0, 0.0), Fj = Fi, c2i = Fi, c2j = Fi;
    double c0s = 0.0;
    if (MIXED) {
        Fi = e.mF[ic];
        Fj = e.mF[j];
        c2i = e.c2[ic];
        c2j = e.c2[j];
        c0s = 0.5 * a2 * (e.c0[ic] + e.c0[j]);
    }
    // the radial pass's partial sum of the next channel is fetched one
    // channel ahead, so its latency hides behind the current channel's edges
    zd pending = wave < g ? out[nn * ro(e.gch)[g0 + wave] + at] : zmake(0.0, 0.0);
    for (int t = wave; t < g; t += kAngWaves) {
        const int c = ro(e.gch)[g0 + t];
        const zd own = ps[t * kApplyBX + lane];
        zd val = pending;
        if (t + kAngWaves < g) pending = out[nn * ro(e.gch)[g0 + t + kAngWaves] + at];
        if (coulomb) {
            // edge by edge in the reference's order (src/kernels.f90:139-143),
            // out += (vcoef rad_lam) psi_b; branch-free, so the edge records
            // (scalar loads) and LDS reads of several edges are in flight together
            const int k0 = ro(e.vptr)[c], k1 = ro(e.vptr)[c + 1];
            const auto *ed = ro(e.aedge);
            if (realrad) {
#pragma unroll 4
                for (int k = k0; k < k1; ++k) {
                    const double w = ed[k].cre * radr[ed[k].lam * kApplyBX + lane];
                    val = zadd(val, zscale(ps[ed[k].src * kApplyBX + lane], w));
                }
            } else {
#pragma unroll 4
                for (int k = k0; k < k1; ++k) {
                    const zd w = zmul(zmake(ed[k].cre, ed[k].cim), rad[ed[k].lam * kApplyBX + lane]);
                    val = zadd(val, zmul(w, ps[ed[k].src * kApplyBX + lane]));
                }
            }
        }
        if (dipole) {
            // field-coupling edges four at a time: the source values of all
            // four are loaded before any is used (src/kernels.f90:144-168 order)
            const int k0 = ro(e.dptr)[c], k1 = ro(e.dptr)[c + 1];
            for (int kb = k0; kb < k1; kb += 4) {
                zd xs[4], ys[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    xs[u] = ys[u] = zmake(0.0, 0.0);
                    if (kb + u < k1) {
                        const int s = ro(e.dc)[kb + u];
                        xs[u] = psi[nn * s + at];
                        if (VELOCITY && !ZMIX) ys[u] = ro(e.de)[kb + u] == 1 ? Y1[nn * s + at] : Y2[nn * s + at];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = kb + u;
                    if (k >= k1) continue;
                    const zd coef = ldz(ro(dcz) + (int64_t)member * e.nd + k);
                    if (MIXED) {
                        const bool first = ro(e.de)[k] == 1;
                        zd part = ys[u];
                        const double ld = ro(e.ldiff)[k];
                        if (ld != 0.0) part = zfma(zscale(first ? half_ri : half_rj, ld), xs[u], part);
                        val = zfma(zmake(-coef.y, coef.x), part, val);
                        val = zfma(zmul(ldz(ro(e.dczE) + k), first ? Fi : Fj), xs[u], val);
                        continue;
                    }
                    if (coef.x == 0.0 && coef.y == 0.0) continue;   // skipped edge
                    const bool first = ro(e.de)[k] == 1;
                    if (!VELOCITY) {
                        val = zfma(zmul(coef, first ? ri : rj), xs[u], val);
                    } else {
                        zd part = ys[u];
                        const double ld = ro(e.ldiff)[k];
                        if (ld != 0.0) part = zfma(zscale(first ? half_ri : half_rj, ld), xs[u], part);
                        val = zfma(zmake(-coef.y, coef.x), part, val);
                    }
                }
            }
        }
        if (MIXED) {
            const int s0 = ro(e.sptr)[c], s1 = ro(e.sptr)[c + 1];
            for (int k = s0; k < s1; ++k) {
                const zd w = zmul(ldz(ro(e.dczQ) + k), ro(e.sele)[k] == 1 ? c2i : c2j);
                val = zfma(w, psi[nn * ro(e.ssrc)[k] + at], val);
            }
            val = zadd(val, zscale(own, c0s));
        } else if (VELOCITY && (terms & kTermA2)) {
            val = zadd(val, zscale(own, a2));
        }
        if (affine) val = zadd(zmul(ca, val), zmul(cb, own));
        if (live) st_stream(out + nn * c + at, val);
    }
}

template <bool ADJ, bool BASE, bool DIPOLE>
static void radial_by(int by, dim3 grid, hipStream_t st, const EngineView &v, BatchArgs bt, int64_t ys,
                      const zd *psi, zd *out, zd *y1, zd *y2) {
    switch (by) {
    case 1:
        hipLaunchKernelGGL((k_radial<ADJ, BASE, DIPOLE, 1>), grid, dim3(kApplyBX, 1), 0, st, v, bt, ys, psi, out, y1,
                           y2);
        break;
    case 2:
        hipLaunchKernelGGL((k_radial<ADJ, BASE, DIPOLE, 2>), grid, dim3(kApplyBX, 2), 0, st, v, bt, ys, psi, out, y1,
                           y2);
        break;
    case 8:
        hipLaunchKernelGGL((k_radial<ADJ, BASE, DIPOLE, 8>), grid, dim3(kApplyBX, 8), 0, st, v, bt, ys, psi, out, y1,
                           y2);
        break;
    default:
        hipLaunchKernelGGL((k_radial<ADJ, BASE, DIPOLE, 4>), grid, dim3(kApplyBX, 4), 0, st, v, bt, ys, psi, out, y1,
                           y2);
    }
}

template <bool ADJ, bool BASE, bool DIPOLE>
static void radial_staged(int members, hipStream_t st, const EngineView &v, BatchArgs bt, int64_t ys, const zd *psi,
                          zd *out, zd *y1, zd *y2) {
    const int64_t blocks = (int64_t)((v.n + kApplyBX - 1) / kApplyBX) * ((v.n + 4 * kStageJ - 1) / (4 * kStageJ)) *
                           members * v.nc;
    dim3 grid((unsigned)blocks);
    const size_t shmem = sizeof(zd) * 4 * kStageJ * v.stage_width;
    hipLaunchKernelGGL((k_radial_staged<ADJ, BASE, DIPOLE>), grid, dim3(kApplyBX, 4), shmem, st, v, bt, ys, psi, out,
                       y1, y2);
}

template <bool ADJ>
static void radial_sel(bool staged, bool base, bool dip, int by, dim3 grid, hipStream_t st, const EngineView &v,
                       BatchArgs bt, int64_t ys, const zd *psi, zd *out, zd *y1, zd *y2) {
    if (by == kLaunchMfma) {
        const int members = grid.z / v.nc;
        if (base && dip)
            radial_mfma<ADJ, true, true>(members, st, v, bt, ys, psi, out, y1, y2);
        else if (base)
            radial_mfma<ADJ, true, false>(members, st, v, bt, ys, psi, out, y1, y2);
        else if (dip)
            radial_mfma<ADJ, false, true>(members, st, v, bt, ys, psi, out, y1, y2);
        else
            radial_mfma<ADJ, false, false>(members, st, v, bt, ys, psi, out, y1, y2);
        return;
    }
    if (staged) {
        const int members = grid.z / v.nc;
        if (base && dip)
            radial_staged<ADJ, true, true>(members, st, v, bt, ys, psi, out, y1, y2);
        else if (base)
            radial_staged<ADJ, true, false>(members, st, v, bt, ys, psi, out, y1, y2);
        else if (dip)
            radial_staged<ADJ, false, true>(members, st, v, bt, ys, psi, out, y1, y2);
        else
            radial_staged<ADJ, false, false>(members, st, v, bt, ys, psi, out, y1, y2);
        return;
    }
    if (base && dip)
        radial_by<ADJ, true, true>(by, grid, st, v, bt, ys, psi, out, y1, y2);
    else if (base)
        radial_by<ADJ, true, false>(by, grid, st, v, bt, ys, psi, out, y1, y2);
    else if (dip)
        radial_by<ADJ, false, true>(by, grid, st, v, bt, ys, psi, out, y1, y2);
    else
        radial_by<ADJ, false, false>(by, grid, st, v, bt, ys, psi, out, y1, y2);
}

// Two-pass apply for `members` states; y1/y2 scratch hold members x N each.
int hex_launch_apply2(const EngineView &v, int by, const double *field, int velocity, int adjoint, zd ca, zd cb,
                      BatchArgs bt, int members, const zd *psi, zd *out, zd *y1, zd *y2, zd *dcz, hipStream_t st,
                      int terms) {
    // launch shape: 0 = staged radial walks (64 i x 16 j tiles through LDS),
    // 1, 2, 4, 8 = direct walks with that many j rows per workgroup
    const bool staged = by == 0 && v.stage_width > 0 && v.stage_width <= kStageMax && v.tband;
    if (by != 1 && by != 2 && by != 4 && by != 8 && !(by == kLaunchMfma && v.nb16 > 0)) by = kApplyBY;
    const int64_t ys = (int64_t)v.n * v.n * v.nc;
    const bool base = terms & kTermBase;
    // the B walks are needed only for a velocity-gauge coupling with a field on
    // some member; with per-member fields that is decided per member in k_angular
    const bool anyfield = bt.fields != nullptr ||
                          (field && (field[0] != 0.0 || field[1] != 0.0 || field[2] != 0.0));
    const bool dip = velocity && (terms & kTermDipole) && anyfield;
    // matrix-core walks with a velocity-gauge field: the T R - R T walks act
    // on each channel's mixed field-coupling sources and add straight into
    // the output, so no Y1/Y2 scratch is written or read
    const bool zmix = by == kLaunchMfma && dip && v.zsrc_ok;
    const double f0 = field ? field[0] : 0.0, f1 = field ? field[1] : 0.0, f2 = field ? field[2] : 0.0;
    if ((terms & kTermDipole) && anyfield && v.nd > 0) {
        const int64_t total = (int64_t)members * v.nd;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1024);
        hipLaunchKernelGGL(k_dipole_coefs, dim3(blocks), dim3(256), 0, st, v, f0, f1, f2, bt, members, dcz);
        HEX_HIP(hipGetLastError());
    }
    dim3 g1((v.n + kApplyBX - 1) / kApplyBX, (v.n + by - 1) / by, members * v.nc);
    if (zmix) {
        if (adjoint)
            base ? radial_mfma<true, true, true, true>(members, st, v, bt, ys, psi, out, y1, y2, dcz)
                 : radial_mfma<true, false, true, true>(members, st, v, bt, ys, psi, out, y1, y2, dcz);
        else
            base ? radial_mfma<false, true, true, true>(members, st, v, bt, ys, psi, out, y1, y2, dcz)
                 : radial_mfma<false, false, true, true>(members, st, v, bt, ys, psi, out, y1, y2, dcz);
    } else if (adjoint) {
        radial_sel<true>(staged, base, dip, by, g1, st, v, bt, ys, psi, out, y1, y2);
    } else {
        radial_sel<false>(staged, base, dip, by, g1, st, v, bt, ys, psi, out, y1, y2);
    }
    HEX_HIP(hipGetLastError());
    const int affine = !(ca.x == 1.0 && ca.y == 0.0 && cb.x == 0.0 && cb.y == 0.0);
    dim3 g2((unsigned)((int64_t)((v.n + kApplyBX - 1) / kApplyBX) * v.n * members * v.ngroups));
    const size_t shmem = sizeof(zd) * kApplyBX * (v.gmax + v.lam_count);
    if (velocity && zmix) {
        if (adjoint)
            hipLaunchKernelGGL((k_angular<true, true, false, true>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, f0,
                               f1, f2, ca, cb, affine, terms, bt, ys, psi, out, y1, y2, dcz);
        else
            hipLaunchKernelGGL((k_angular<true, false, false, true>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, f0,
                               f1, f2, ca, cb, affine, terms, bt, ys, psi, out, y1, y2, dcz);
    } else if (velocity) {
        if (adjoint)
            hipLaunchKernelGGL((k_angular<true, true>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, f0, f1, f2, ca, cb, affine,
                               terms, bt, ys, psi, out, y1, y2, dcz);
        else
            hipLaunchKernelGGL((k_angular<true, false>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, f0, f1, f2, ca, cb, affine,
                               terms, bt, ys, psi, out, y1, y2, dcz);
    } else {
        if (adjoint)
            hipLaunchKernelGGL((k_angular<false, true>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, f0, f1, f2, ca, cb, affine,
                               terms, bt, ys, psi, out, y1, y2, dcz);
        else
            hipLaunchKernelGGL((k_angular<false, false>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, f0, f1, f2, ca, cb, affine,
                               terms, bt, ys, psi, out, y1, y2, dcz);
    }
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// Mixed-gauge apply: out = ca H(c) psi + cb psi for the 12 coefficients c.
// The radial pass always forms the field-coupling walks (built from q), so the
// coupling pass reads defined values whatever the coefficients.
int hex_launch_apply_mixed(const EngineView &v, int by, const double *coeff, zd ca, zd cb, const zd *psi, zd *out,
                           hipStream_t st) {
    const bool staged = by == 0 && v.stage_width > 0 && v.stage_width <= kStageMax && v.tband;
    if (by != 1 && by != 2 && by != 4 && by != 8 && !(by == kLaunchMfma && v.nb16 > 0)) by = kApplyBY;
    const int64_t N = (int64_t)v.n * v.n * v.nc;
    BatchArgs bt{nullptr, nullptr, N, N};
    dim3 g1((v.n + kApplyBX - 1) / kApplyBX, (v.n + by - 1) / by, v.nc);
    radial_sel<false>(staged, true, true, by, g1, st, v, bt, N, psi, out, v.y1, v.y2);
    HEX_HIP(hipGetLastError());
    Coef12 f;
    for (int m = 0; m < 12; ++m) f.c[m] = coeff[m];
    const int64_t total = (int64_t)v.nd + v.nsq;
    if (total > 0) {
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1024);
        hipLaunchKernelGGL(k_mixed_coefs, dim3(blocks), dim3(256), 0, st, v, f, v.dcz);
        HEX_HIP(hipGetLastError());
    }
    const int affine = !(ca.x == 1.0 && ca.y == 0.0 && cb.x == 0.0 && cb.y == 0.0);
    dim3 g2((unsigned)((int64_t)((v.n + kApplyBX - 1) / kApplyBX) * v.n * v.ngroups));
    const size_t shmem = sizeof(zd) * kApplyBX * (v.gmax + v.lam_count);
    hipLaunchKernelGGL((k_angular<true, false, true>), g2, dim3(kApplyBX, kAngWaves), shmem, st, v, coeff[0], coeff[1],
                       coeff[2], ca, cb, affine, kTermAll, bt, N, psi, out, v.y1, v.y2, v.dcz);
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}
