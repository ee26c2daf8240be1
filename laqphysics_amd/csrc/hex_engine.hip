// laqphysics_hip engine: operator tables on the device and the H.psi apply.
//
// The apply computes, for every output point (i, j, c) of the F-ordered state
// psi(i, j, c) (index i + n j + n^2 c), the same sum as the reference kernel
// (src/kernels.f90:130-175):
//   diag psi + T_i psi + T_j psi + sum_edges vcoef rad_lam psi_b
//   + dipole edges (length: coef r psi_s; velocity: i coef [ldiff/2r psi_s +
//     sum_b T_ib (r_b - r_i) psi_s(b)]) + |A|^2 psi (velocity gauge)
// with the diagonal regenerated from two n-vectors per channel and the
// multipoles regenerated per radial pair from r and cut.
#include "hex_common.h"
#include "../../include/laqphysics_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hex {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
}  // namespace hex

extern "C" const char *hex_last_error(void) { return hex::g_error.c_str(); }
extern "C" int hex_version(void) { return 1; }

#include "hex_internal.h"

template <typename T>
static int upload(hex_engine *e, const T *host, size_t count, const T **dst) {
    if (count == 0) {
        *dst = nullptr;
        return HEX_OK;
    }
    void *p = nullptr;
    HEX_HIP(hipMalloc(&p, count * sizeof(T)));
    e->owned.push_back(p);
    e->bytes += (int64_t)(count * sizeof(T));
    HEX_HIP(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<const T *>(p);
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// Apply kernel.  One thread per radial pair (i, j); a wave covers 64
// consecutive i at one j, so every psi access of the j-direction walk and of
// the channel couplings is a coalesced 64 x 16 B row.  Each thread loops over
// all output channels; the multipoles of its pair live in LDS.
//
// Batching: blockIdx.z is the member; members are in/out_stride entries apart,
// carry their own field A (fields[3 b]) and are skipped when active[b] == 0.
// ADJ applies H^dagger: on every grid, H's channel-coupling matrices are
// Hermitian and its radial pieces are analytic in (r, T), so H^dagger is H with
// conjugated r, T, T R - R T, diagonal and multipoles (the irECS contour
// reflected), and unchanged edge coefficients.
// ---------------------------------------------------------------------------

template <bool ADJ>
__device__ __forceinline__ zd rz(zd a) { return ADJ ? zconj(a) : a; }

template <bool VELOCITY, bool ADJ, int BY>
__global__ __launch_bounds__(kApplyBX *BY) void apply_direct(EngineView e, double f0, double f1, double f2, zd ca,
                                                             zd cb, int affine, int terms, BatchArgs bt,
                                                             const zd *__restrict__ psi, zd *__restrict__ out) {
    extern __shared__ zd rad_s[];   // [lam_count][kApplyBX * BY]
    const int member = blockIdx.z;
    if (bt.active && !bt.active[member]) return;
    if (bt.fields) {
        f0 = bt.fields[3 * member];
        f1 = bt.fields[3 * member + 1];
        f2 = bt.fields[3 * member + 2];
    }
    psi += bt.in_stride * member;
    out += bt.out_stride * member;
    const int tid = threadIdx.x + kApplyBX * threadIdx.y;
    const int i = blockIdx.x * kApplyBX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int n = e.n;
    if (i >= n || j >= n) return;
    const int64_t nn = (int64_t)n * n;
    const int64_t at = i + (int64_t)n * j;

    const zd ri = rz<ADJ>(e.r[i]), rj = rz<ADJ>(e.r[j]);
    const double cc = e.cut[i] * e.cut[j];
    const bool base = terms & kTermBase;
    const bool streamed = e.rad_full != nullptr;
    const bool coulomb = base && (streamed || cc != 0.0);
    if (!coulomb) {
    } else if (streamed) {
        for (int lam = 0; lam < e.lam_count; ++lam) rad_s[lam * (kApplyBX * BY) + tid] = rz<ADJ>(e.rad_full[at + nn * lam]);
    } else if (coulomb) {
        const zd lo = i < j ? ri : rj;
        const zd hi = i < j ? rj : ri;
        // rad_lam = lo^lam / hi^(lam+1) * cut_i cut_j
        zd num = zmake(1.0, 0.0);
        zd den = hi;
        for (int lam = 0; lam < e.lam_count; ++lam) {
            rad_s[lam * (kApplyBX * BY) + tid] = zscale(zdiv(num, den), cc);
            num = zmul(num, lo);
            den = zmul(den, hi);
        }
    }
    const double a2 = f0 * f0 + f1 * f1 + f2 * f2;
    const int pi0 = e.ptr[i], pi1 = e.ptr[i + 1];
    const int pj0 = e.ptr[j], pj1 = e.ptr[j + 1];
    const zd half_ri = zdiv(zmake(1.0, 0.0), zscale(ri, 2.0));
    const zd half_rj = zdiv(zmake(1.0, 0.0), zscale(rj, 2.0));

    for (int c = 0; c < e.nc; ++c) {
        const zd *pc = psi + nn * c;
        const zd own = pc[at];
        zd val = zmake(0.0, 0.0);
        if (base) {
            const zd dg = rz<ADJ>(e.diag_full ? e.diag_full[at + nn * c]
                                              : zadd(e.d1[i + (int64_t)n * c], e.d2[j + (int64_t)n * c]));
            val = zmul(dg, own);
            for (int k = pi0; k < pi1; ++k) val = zfma(rz<ADJ>(e.tv[k]), pc[e.col[k] + (int64_t)n * j], val);
            for (int k = pj0; k < pj1; ++k) val = zfma(rz<ADJ>(e.tv[k]), pc[i + (int64_t)n * e.col[k]], val);
        }
        if (coulomb) {
            for (int k = e.vptr[c]; k < e.vptr[c + 1]; ++k)
                val = zfma(zmul(e.vcoef[k], rad_s[e.vl[k] * (kApplyBX * BY) + tid]), psi[at + nn * e.vc[k]], val);
        }
        for (int k = (terms & kTermDipole) ? e.dptr[c] : 0; k < ((terms & kTermDipole) ? e.dptr[c + 1] : 0); ++k) {
            const zd *dk = e.dcoef + 3 * k;
            const zd coef = zmake(f0 * dk[0].x + f1 * dk[1].x + f2 * dk[2].x,
                                  f0 * dk[0].y + f1 * dk[1].y + f2 * dk[2].y);
            if (sqrt(zabs2(coef)) < 1e-30) continue;
            const zd *ps = psi + nn * e.dc[k];
            const bool first = e.de[k] == 1;
            if (!VELOCITY) {
                val = zfma(zmul(coef, first ? ri : rj), ps[at], val);
            } else {
                zd part = zmake(0.0, 0.0);
                const double ld = e.ldiff[k];
                if (first) {
                    if (ld != 0.0) part = zmul(zscale(half_ri, ld), ps[at]);
                    for (int q = pi0; q < pi1; ++q)
                        part = zfma(rz<ADJ>(e.tb[q]), ps[e.col[q] + (int64_t)n * j], part);
                } else {
                    if (ld != 0.0) part = zmul(zscale(half_rj, ld), ps[at]);
                    for (int q = pj0; q < pj1; ++q)
                        part = zfma(rz<ADJ>(e.tb[q]), ps[i + (int64_t)n * e.col[q]], part);
                }
                // i * coef * part
                const zd ic = zmake(-coef.y, coef.x);
                val = zfma(ic, part, val);
            }
        }
        if (VELOCITY && (terms & kTermA2)) val = zadd(val, zscale(own, a2));
        // out = ca * (H psi) + cb * psi
        if (affine) val = zadd(zmul(ca, val), zmul(cb, own));
        out[at + nn * c] = val;
    }
}

// ---------------------------------------------------------------------------
// Engine lifetime
// ---------------------------------------------------------------------------
static bool check_tables(const hex_tables *t) {
    if (!t || t->n < 1 || t->nc < 1 || t->lam_count < 1 || t->lam_count > kLamMax) {
        hex::set_error("tables: need n>=1, nc>=1, 1<=lam_count<=32");
        return false;
    }
    if (!t->ptr || !t->r || !t->cut || !t->d1 || !t->d2 || !t->vptr || !t->dptr) {
        hex::set_error("tables: missing required array");
        return false;
    }
    if (t->ptr[0] != 0 || t->ptr[t->n] != t->nt || t->vptr[t->nc] != t->nv || t->dptr[t->nc] != t->nd) {
        hex::set_error("tables: CSR/edge pointer arrays disagree with their lengths");
        return false;
    }
    for (int k = 0; k < t->nt; ++k)
        if (t->col[k] < 0 || t->col[k] >= t->n) {
            hex::set_error("tables: kinetic column index out of range");
            return false;
        }
    for (int k = 0; k < t->nv; ++k)
        if (t->vc[k] < 0 || t->vc[k] >= t->nc || t->vl[k] < 0 || t->vl[k] >= t->lam_count) {
            hex::set_error("tables: Coulomb edge index out of range");
            return false;
        }
    for (int k = 0; k < t->nd; ++k)
        if (t->dc[k] < 0 || t->dc[k] >= t->nc || (t->de[k] != 1 && t->de[k] != 2)) {
            hex::set_error("tables: dipole edge index out of range");
            return false;
        }
    return true;
}

int hex_engine_alloc_scratch(hex_engine *e);

// bq: the radial function of the field-coupling walks B = T Q - Q T (r when
// null; q = r (1 - f) for the mixed gauge).
static int engine_create(const hex_tables *t, const zd *bq, int device, hex_engine **out) {
    if (!out) {
        hex::set_error("engine: null output pointer");
        return HEX_EINVAL;
    }
    *out = nullptr;
    if (!check_tables(t)) return HEX_EINVAL;
    HEX_HIP(hipSetDevice(device));
    auto *e = new hex_engine();
    e->device = device;
    e->nt = t->nt;
    e->nv = t->nv;
    e->nd = t->nd;
    EngineView &v = e->v;
    v.n = t->n;
    v.nc = t->nc;
    v.lam_count = t->lam_count;
    const zd *tv = reinterpret_cast<const zd *>(t->tv);
    const zd *r = reinterpret_cast<const zd *>(t->r);
    // B = T R - R T on the CSR pattern: B_ik = T_ik (r_k - r_i) (q for r when given)
    const zd *rb = bq ? bq : r;
    std::vector<zd> tb(t->nt);
    for (int i = 0; i < t->n; ++i)
        for (int k = t->ptr[i]; k < t->ptr[i + 1]; ++k) {
            const zd d = zsub(rb[t->col[k]], rb[i]);
            tb[k] = zmul(tv[k], d);
        }
    int rc = HEX_OK;
    auto up = [&](auto *host, size_t count, auto **dst) {
        if (rc == HEX_OK) rc = upload(e, host, count, dst);
    };
    up(t->ptr, (size_t)t->n + 1, &v.ptr);
    up(t->col, (size_t)t->nt, &v.col);
    up(tv, (size_t)t->nt, &v.tv);
    up(tb.data(), (size_t)t->nt, &v.tb);
    up(r, (size_t)t->n, &v.r);
    up(t->cut, (size_t)t->n, &v.cut);
    up(reinterpret_cast<const zd *>(t->d1), (size_t)t->n * t->nc, &v.d1);
    up(reinterpret_cast<const zd *>(t->d2), (size_t)t->n * t->nc, &v.d2);
    up(t->vptr, (size_t)t->nc + 1, &v.vptr);
    up(t->vc, (size_t)t->nv, &v.vc);
    up(t->vl, (size_t)t->nv, &v.vl);
    up(reinterpret_cast<const zd *>(t->vcoef), (size_t)t->nv, &v.vcoef);
    up(t->dptr, (size_t)t->nc + 1, &v.dptr);
    up(t->dc, (size_t)t->nd, &v.dc);
    up(t->de, (size_t)t->nd, &v.de);
    up(reinterpret_cast<const zd *>(t->dcoef), (size_t)3 * t->nd, &v.dcoef);
    up(t->ldiff, (size_t)t->nd, &v.ldiff);
    // field-coupling sources per (output channel, electron), for the walks of
    // the mixed sources Z_e = sum_k i coef_k psi_src(k)
    {
        std::vector<int32_t> zsrc((size_t)2 * t->nc * kZSrc, -1);
        v.zsrc_ok = t->nd > 0;
        for (int c = 0; c < t->nc; ++c) {
            int cnt[2] = {0, 0};
            for (int k = t->dptr[c]; k < t->dptr[c + 1]; ++k) {
                const int el = t->de[k] == 1 ? 0 : 1;
                if (cnt[el] < kZSrc)
                    zsrc[((size_t)2 * c + el) * kZSrc + cnt[el]] = k;
                else
                    v.zsrc_ok = 0;
                ++cnt[el];
            }
        }
        up(zsrc.data(), zsrc.size(), &v.zsrc);
    }
    // Coulomb groups (union-find over the multipole edges)
    {
        std::vector<int> parent(t->nc);
        for (int c = 0; c < t->nc; ++c) parent[c] = c;
        auto find = [&](int a) {
            while (parent[a] != a) a = parent[a] = parent[parent[a]];
            return a;
        };
        for (int c = 0; c < t->nc; ++c)
            for (int k = t->vptr[c]; k < t->vptr[c + 1]; ++k) {
                const int a = find(c), b = find(t->vc[k]);
                if (a != b) parent[a] = b;
            }
        std::vector<int> root_to_group(t->nc, -1);
        std::vector<std::vector<int>> groups;
        for (int c = 0; c < t->nc; ++c) {
            const int r = find(c);
            if (root_to_group[r] < 0) {
                root_to_group[r] = (int)groups.size();
                groups.emplace_back();
            }
            groups[root_to_group[r]].push_back(c);
        }
        std::vector<int32_t> gptr{0}, gch, pos(t->nc);
        int gmax = 0;
        for (auto &g : groups) {
            for (size_t q = 0; q < g.size(); ++q) {
                pos[g[q]] = (int)q;
                gch.push_back(g[q]);
            }
            gptr.push_back((int32_t)gch.size());
            gmax = std::max(gmax, (int)g.size());
        }
        std::vector<int32_t> vcl(t->nv);
        for (int k = 0; k < t->nv; ++k) vcl[k] = pos[t->vc[k]];
        v.ngroups = (int)groups.size();
        v.vreal = 1;
        for (int k = 0; k < t->nv; ++k)
            if (reinterpret_cast<const zd *>(t->vcoef)[k].y != 0.0) v.vreal = 0;
        v.gmax = gmax;
        up(gptr.data(), gptr.size(), &v.gptr);
        up(gch.data(), gch.size(), &v.gch);
        up(vcl.data(), vcl.size(), &v.vcl);
        std::vector<AngEdge> aedge(t->nv);
        for (int k = 0; k < t->nv; ++k) {
            const zd w = reinterpret_cast<const zd *>(t->vcoef)[k];
            aedge[k] = AngEdge{w.x, w.y, t->vl[k], vcl[k], {0, 0}};
        }
        up(aedge.data(), aedge.size(), &v.aedge);
        if (rc == HEX_OK && gmax + t->lam_count > 64) {
            hex::set_error("tables: Coulomb-coupled channel group plus multipole orders exceed 64 (LDS budget)");
            rc = HEX_EINVAL;
        }
    }
    // staged-walk tables
    {
        const int n = t->n;
        std::vector<int32_t> lo(n), hi(n);
        int band = 1;
        for (int i = 0; i < n; ++i) {
            lo[i] = i;
            hi[i] = i;
            for (int k = t->ptr[i]; k < t->ptr[i + 1]; ++k) {
                lo[i] = std::min(lo[i], t->col[k]);
                hi[i] = std::max(hi[i], t->col[k]);
            }
            band = std::max(band, hi[i] - lo[i] + 1);
        }
        const int nb = (n + kApplyBX - 1) / kApplyBX;
        std::vector<int32_t> iblo(nb), ibhi(nb);
        for (int b = 0; b < nb; ++b) {
            int a = n, z = 0;
            for (int i = b * kApplyBX; i < std::min(n, (b + 1) * kApplyBX); ++i) {
                a = std::min(a, lo[i]);
                z = std::max(z, hi[i] + 1);
            }
            iblo[b] = a;
            ibhi[b] = z;
        }
        std::vector<zd> tband((size_t)n * band, zmake(0.0, 0.0)), bband((size_t)n * band, zmake(0.0, 0.0));
        for (int i = 0; i < n; ++i)
            for (int k = t->ptr[i]; k < t->ptr[i + 1]; ++k) {
                tband[(size_t)i * band + (t->col[k] - lo[i])] = tv[k];
                bband[(size_t)i * band + (t->col[k] - lo[i])] = tb[k];
            }
        // the same bands transposed ([band][n]): lane i reads entry (off, i),
        // so a wave's coefficient loads are one contiguous 1 KiB row
        // padded to a multiple of 4 offsets with zero coefficients, so the
        // walks run whole groups of four offsets without per-offset tests
        const int band4 = (band + 3) / 4 * 4;
        std::vector<zd> tbandT((size_t)n * band4, zmake(0.0, 0.0)), bbandT((size_t)n * band4, zmake(0.0, 0.0));
        for (int i = 0; i < n; ++i)
            for (int o = 0; o < band; ++o) {
                tbandT[(size_t)o * n + i] = tband[(size_t)i * band + o];
                bbandT[(size_t)o * n + i] = bband[(size_t)i * band + o];
            }
        std::vector<int32_t> ibw(nb);
        for (int b = 0; b < nb; ++b) {
            ibw[b] = 1;
            for (int i = b * kApplyBX; i < std::min(n, (b + 1) * kApplyBX); ++i)
                ibw[b] = std::max(ibw[b], hi[i] - lo[i] + 1);
            ibw[b] = (ibw[b] + 3) / 4 * 4;
        }
        std::vector<double> tbandTr((size_t)n * band4), bbandTr((size_t)n * band4);
        for (size_t x = 0; x < tbandT.size(); ++x) {
            tbandTr[x] = tbandT[x].x;
            bbandTr[x] = bbandT[x].x;
        }
        std::vector<int32_t> rowreal(n, 1), ireal(nb, 1);
        for (int i = 0; i < n; ++i)
            for (int o = 0; o < band; ++o)
                if (tband[(size_t)i * band + o].y != 0.0 || bband[(size_t)i * band + o].y != 0.0) rowreal[i] = 0;
        for (int b = 0; b < nb; ++b)
            for (int i = b * kApplyBX; i < std::min(n, (b + 1) * kApplyBX); ++i)
                if (!rowreal[i]) ireal[b] = 0;
        up(tbandT.data(), tbandT.size(), &v.tbandT);
        up(bbandT.data(), bbandT.size(), &v.bbandT);
        up(tbandTr.data(), tbandTr.size(), &v.tbandTr);
        up(bbandTr.data(), bbandTr.size(), &v.bbandTr);
        up(ibw.data(), ibw.size(), &v.ibw);
        up(rowreal.data(), rowreal.size(), &v.rowreal);
        up(ireal.data(), ireal.size(), &v.ireal);
        // j-walk row groups
        const int ng = (n + kStageJ - 1) / kStageJ;
        std::vector<int32_t> glo(ng), gw(ng), greal(ng, 1);
        int gmaxw = 1;
        for (int g = 0; g < ng; ++g) {
            int a = n, z = 0;
            for (int j = g * kStageJ; j < std::min(n, (g + 1) * kStageJ); ++j) {
                a = std::min(a, lo[j]);
                z = std::max(z, hi[j] + 1);
                if (!rowreal[j]) greal[g] = 0;
            }
            glo[g] = a;
            gw[g] = (z - a + 3) / 4 * 4;   // whole groups of four rows; rows past n have zero coefficients
            gmaxw = std::max(gmaxw, gw[g]);
        }
        std::vector<zd> gt((size_t)ng * gmaxw * kStageJ, zmake(0.0, 0.0)), gb(gt.size(), zmake(0.0, 0.0));
        for (int g = 0; g < ng; ++g)
            for (int j = g * kStageJ; j < std::min(n, (g + 1) * kStageJ); ++j) {
                const int q = j - g * kStageJ;
                for (int o = 0; o <= hi[j] - lo[j]; ++o) {
                    const size_t at = ((size_t)g * gmaxw + (lo[j] + o - glo[g])) * kStageJ + q;
                    gt[at] = tband[(size_t)j * band + o];
                    gb[at] = bband[(size_t)j * band + o];
                }
            }
        std::vector<double> gtr(gt.size()), gbr(gb.size());
        for (size_t x = 0; x < gt.size(); ++x) {
            gtr[x] = gt[x].x;
            gbr[x] = gb[x].x;
        }
        v.jg_max = gmaxw;
        up(glo.data(), glo.size(), &v.jg_lo);
        up(gw.data(), gw.size(), &v.jg_w);
        up(greal.data(), greal.size(), &v.jg_real);
        up(gt.data(), gt.size(), &v.jg_t);
        up(gb.data(), gb.size(), &v.jg_b);
        up(gtr.data(), gtr.size(), &v.jg_tr);
        up(gbr.data(), gbr.size(), &v.jg_br);
        v.band = band;
        v.stage_width = 0;
        for (int b = 0; b < nb; ++b) v.stage_width = std::max(v.stage_width, ibhi[b] - iblo[b]);
        up(iblo.data(), iblo.size(), &v.iblo);
        up(ibhi.data(), ibhi.size(), &v.ibhi);
        up(lo.data(), lo.size(), &v.jlo);
        std::vector<int32_t> hi1(n);
        for (int i = 0; i < n; ++i) hi1[i] = hi[i] + 1;
        up(hi1.data(), hi1.size(), &v.jhi);
        up(tband.data(), tband.size(), &v.tband);
        up(bband.data(), bband.size(), &v.bband);
        // matrix-core walk tiles: per 16-row block the union of its rows'
        // bands in steps of four columns, one 64-entry operand tile per step
        const int nb16 = (n + 15) / 16;
        std::vector<int32_t> kst(nb16), kns(nb16), koff(nb16 + 1, 0), kreal(nb16, 1);
        for (int b = 0; b < nb16; ++b) {
            int a = n, z = 0;
            for (int i = 16 * b; i < std::min(n, 16 * b + 16); ++i) {
                a = std::min(a, lo[i]);
                z = std::max(z, hi[i] + 1);
                if (!rowreal[i]) kreal[b] = 0;
            }
            kst[b] = a;
            kns[b] = (z - a + 15) / 16 * 4;   // whole rounds of four steps; the extra tiles are zero
            koff[b + 1] = koff[b] + kns[b];
        }
        const size_t tiles = (size_t)koff[nb16];
        std::vector<double> mt(tiles * 64, 0.0), mti(tiles * 64, 0.0), mb(tiles * 64, 0.0), mbi(tiles * 64, 0.0);
        for (int b = 0; b < nb16; ++b)
            for (int s = 0; s < kns[b]; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int i = 16 * b + l % 16, k = kst[b] + 4 * s + l / 16;
                    if (i >= n || k < lo[i] || k > hi[i]) continue;
                    const size_t at = ((size_t)koff[b] + s) * 64 + l;
                    const zd tt = tband[(size_t)i * band + (k - lo[i])];
                    const zd bb = bband[(size_t)i * band + (k - lo[i])];
                    mt[at] = tt.x;
                    mti[at] = tt.y;
                    mb[at] = bb.x;
                    mbi[at] = bb.y;
                }
        v.nb16 = nb16;
        up(kst.data(), kst.size(), &v.kst);
        up(kns.data(), kns.size(), &v.kns);
        up(koff.data(), koff.size(), &v.koff);
        up(kreal.data(), kreal.size(), &v.kreal);
        up(mt.data(), mt.size(), &v.mt);
        up(mti.data(), mti.size(), &v.mti);
        up(mb.data(), mb.size(), &v.mb);
        up(mbi.data(), mbi.size(), &v.mbi);
    }
    if (rc == HEX_OK) {
        const size_t N = (size_t)t->n * t->n * t->nc;
        const size_t nd = (size_t)std::max(t->nd, 1);
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(zd) * (2 * N + nd)) != hipSuccess) {
            hex::set_error("engine: scratch allocation failed");
            rc = HEX_EDEVICE;
        } else {
            e->owned.push_back(p);
            e->bytes += (int64_t)(sizeof(zd) * (2 * N + nd));
            v.y1 = static_cast<zd *>(p);
            v.y2 = v.y1 + N;
            v.dcz = v.y2 + N;
            v.nd = t->nd;
        }
    }
    if (rc == HEX_OK) rc = hex_engine_alloc_scratch(e);
    if (rc != HEX_OK) {
        hex_engine_destroy(e);
        return rc;
    }
    *out = e;
    return HEX_OK;
}

extern "C" int hex_engine_create(const hex_tables *t, int device, hex_engine **out) {
    return engine_create(t, nullptr, device, out);
}

extern "C" int hex_engine_create_mixed(const hex_tables *t, const double *q, int device, hex_engine **out) {
    if (!q) {
        hex::set_error("engine: mixed gauge needs q = r (1 - f)");
        return HEX_EINVAL;
    }
    return engine_create(t, reinterpret_cast<const zd *>(q), device, out);
}

extern "C" int hex_engine_set_mixed(hex_engine *e, const double *hq, const double *F, const double *c0,
                                    const double *c2, int nsq, const int32_t *sptr, const int32_t *ssrc,
                                    const int32_t *sele, const double *scoef) {
    if (!e || !hq || !F || !c0 || !c2 || !sptr || nsq < 0 || (nsq > 0 && (!ssrc || !sele || !scoef))) {
        hex::set_error("mixed: bad argument");
        return HEX_EINVAL;
    }
    EngineView &v = e->v;
    if (v.mixed) {
        hex::set_error("mixed: tables already set");
        return HEX_EINVAL;
    }
    if (sptr[0] != 0 || sptr[v.nc] != nsq) {
        hex::set_error("mixed: square-edge pointer does not span the edge list");
        return HEX_EINVAL;
    }
    for (int k = 0; k < nsq; ++k)
        if (ssrc[k] < 0 || ssrc[k] >= v.nc || (sele[k] != 1 && sele[k] != 2)) {
            hex::set_error("mixed: square edge with a bad source channel or electron");
            return HEX_EINVAL;
        }
    HEX_HIP(hipSetDevice(e->device));
    int rc = HEX_OK;
    auto up = [&](auto *host, size_t count, auto **dst) {
        if (rc == HEX_OK) rc = upload(e, host, count, dst);
    };
    const int n = v.n;
    up(reinterpret_cast<const zd *>(hq), (size_t)n, &v.hq);
    up(reinterpret_cast<const zd *>(F), (size_t)n, &v.mF);
    up(c0, (size_t)n, &v.c0);
    up(reinterpret_cast<const zd *>(c2), (size_t)n, &v.c2);
    up(sptr, (size_t)v.nc + 1, &v.sptr);
    up(ssrc, (size_t)nsq, &v.ssrc);
    up(sele, (size_t)nsq, &v.sele);
    up(reinterpret_cast<const zd *>(scoef), (size_t)6 * nsq, &v.scoef);
    if (rc != HEX_OK) return rc;
    void *p = nullptr;
    const size_t count = (size_t)std::max(e->nd, 1) + (size_t)std::max(nsq, 1);
    HEX_HIP(hipMalloc(&p, sizeof(zd) * count));
    e->owned.push_back(p);
    e->bytes += (int64_t)(sizeof(zd) * count);
    v.dczE = static_cast<zd *>(p);
    v.dczQ = v.dczE + std::max(e->nd, 1);
    v.nsq = nsq;
    v.mixed = 1;
    return HEX_OK;
}

extern "C" int hex_workspace_create(int device, hex_engine **out) {
    if (!out) {
        hex::set_error("workspace: null output pointer");
        return HEX_EINVAL;
    }
    *out = nullptr;
    HEX_HIP(hipSetDevice(device));
    auto *e = new hex_engine();
    e->device = device;
    int rc = hex_engine_alloc_scratch(e);
    if (rc != HEX_OK) {
        hex_engine_destroy(e);
        return rc;
    }
    *out = e;
    return HEX_OK;
}

extern "C" int hex_engine_destroy(hex_engine *e) {
    if (!e) return HEX_OK;
    (void)hipSetDevice(e->device);
    for (void *p : e->owned) (void)hipFree(p);
    if (e->hbuf_host) (void)hipHostFree(e->hbuf_host);
    delete e;
    return HEX_OK;
}

extern "C" int64_t hex_engine_bytes(const hex_engine *e) { return e ? e->bytes : 0; }

// Reduction scratch lives with the engine (allocated once at create time).
static constexpr int kRedBlocks = 1024;
static constexpr int kRedCols = 64;

struct hex_engine_scratch {
    zd *partials;
    zd *zbuf;
    zd *hbuf_host;
};
hex_engine_scratch hex_scratch_of(hex_engine *e) { return {e->partials, e->zbuf, e->hbuf_host}; }

int hex_engine_alloc_scratch(hex_engine *e) {
    void *p = nullptr;
    HEX_HIP(hipMalloc(&p, sizeof(zd) * kRedBlocks * kRedCols));
    e->owned.push_back(p);
    e->partials = static_cast<zd *>(p);
    HEX_HIP(hipMalloc(&p, sizeof(zd) * kRedCols * 2));
    e->owned.push_back(p);
    e->zbuf = static_cast<zd *>(p);
    e->bytes += (int64_t)sizeof(zd) * (kRedBlocks * kRedCols + 2 * kRedCols);
    HEX_HIP(hipHostMalloc((void **)&e->hbuf_host, sizeof(zd) * kRedCols * 2, hipHostMallocDefault));
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// tSURFF source (python/helium3d.py:69-77, python/torch_backend.py:133-138):
//   out[c][s][i] = sum_k J[s,k] psi(i,k,c)
//                + sum_{electron-2 dipole edges c<-b} coef * sum_k JP[s,k] psi(i,k,b)
// One block per (s, c), threads along i (coalesced rows of psi).
// ---------------------------------------------------------------------------
__global__ void surface_flux_kernel(EngineView e, int ns, int k0, int kc, const zd *__restrict__ J,
                                    const zd *__restrict__ JP, double f0, double f1, double f2, int use_base,
                                    int use_field, const zd *__restrict__ psi, zd *__restrict__ out) {
    const int s = blockIdx.x, c = blockIdx.y;
    const int n = e.n;
    const int64_t nn = (int64_t)n * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const zd *pc = psi + nn * c + i;
        zd acc = zmake(0.0, 0.0);
        if (use_base)
            for (int k = 0; k < kc; ++k) acc = zfma(J[s * kc + k], pc[(int64_t)n * (k0 + k)], acc);
        if (use_field) {
            for (int q = e.dptr[c]; q < e.dptr[c + 1]; ++q) {
                if (e.de[q] != 2) continue;
                const zd *dk = e.dcoef + 3 * q;
                const zd coef = zmake(f0 * dk[0].x + f1 * dk[1].x + f2 * dk[2].x,
                                      f0 * dk[0].y + f1 * dk[1].y + f2 * dk[2].y);
                if (sqrt(zabs2(coef)) <= 1e-30) continue;
                const zd *pb = psi + nn * e.dc[q] + i;
                zd part = zmake(0.0, 0.0);
                for (int k = 0; k < kc; ++k) part = zfma(JP[s * kc + k], pb[(int64_t)n * (k0 + k)], part);
                acc = zfma(coef, part, acc);
            }
        }
        out[((int64_t)c * ns + s) * n + i] = acc;
    }
}

extern "C" int hex_engine_set_surface(hex_engine *e, int ns, int k0, int kc, const double *J, const double *JP) {
    if (!e || ns < 1 || kc < 1 || k0 < 0 || k0 + kc > e->v.n || !J || !JP) {
        hex::set_error("set_surface: bad surface window");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(e->device));
    const zd *dj = nullptr, *dp = nullptr;
    int rc = upload(e, reinterpret_cast<const zd *>(J), (size_t)ns * kc, &dj);
    if (rc == HEX_OK) rc = upload(e, reinterpret_cast<const zd *>(JP), (size_t)ns * kc, &dp);
    if (rc != HEX_OK) return rc;
    e->ns = ns;
    e->k0 = k0;
    e->kc = kc;
    e->surfJ = const_cast<zd *>(dj);
    e->surfP = const_cast<zd *>(dp);
    return HEX_OK;
}

extern "C" int hex_flux(hex_engine *e, const double field[3], const double *psi, double *out, void *stream) {
    if (!e || !field || !psi || !out) {
        hex::set_error("flux: null argument");
        return HEX_EINVAL;
    }
    if (!e->surfJ) {
        hex::set_error("flux: no surface prepared (call prepare_surface first)");
        return HEX_EINVAL;
    }
    return hex_flux_terms(e, field, 3, psi, out, stream);
}

extern "C" int hex_flux_terms(hex_engine *e, const double field[3], int terms, const double *psi, double *out,
                              void *stream) {
    if (!e || !field || !psi || !out || terms < 1 || terms > 3) {
        hex::set_error("flux: null argument or terms outside 1..3");
        return HEX_EINVAL;
    }
    if (!e->surfJ) {
        hex::set_error("flux: no surface prepared (call prepare_surface first)");
        return HEX_EINVAL;
    }
    const int use_field = (terms & 2) && (fabs(field[0]) > 1e-30 || fabs(field[1]) > 1e-30 ||
                                          fabs(field[2]) > 1e-30) ? 1 : 0;
    hipLaunchKernelGGL(surface_flux_kernel, dim3(e->ns, e->v.nc), dim3(256), 0, static_cast<hipStream_t>(stream),
                       e->v, e->ns, e->k0, e->kc, e->surfJ, e->surfP, field[0], field[1], field[2], terms & 1,
                       use_field, reinterpret_cast<const zd *>(psi), reinterpret_cast<zd *>(out));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// Adjoint of the surface map F0 = J psi, Fd = (field part for a unit z field)
// psi (surface_flux_kernel with field (0, 0, 1)):
//   u(i, k, b) = sum_s conj(J[s,k]) g0[b][s][i]
//              + sum_{electron-2 edges a <- b} conj(coef) sum_s conj(JP[s,k]) gd[a][s][i],
// coef = dcoef_z of the edge, skipped where the forward map skips it.  One
// block per (k, b), threads along i; rows k outside the surface window are zero.
__global__ void surface_flux_adjoint_kernel(EngineView e, int ns, int k0, int kc, const zd *__restrict__ J,
                                            const zd *__restrict__ JP, const zd *__restrict__ g0,
                                            const zd *__restrict__ gd, zd *__restrict__ u) {
    const int k = blockIdx.x, b = blockIdx.y;
    const int n = e.n, nc = e.nc;
    const int64_t nn = (int64_t)n * n;
    const int kk = k - k0;
    const bool inside = kk >= 0 && kk < kc;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        zd acc = zmake(0.0, 0.0);
        if (inside) {
            for (int s = 0; s < ns; ++s) acc = zcfma(J[s * kc + kk], g0[((int64_t)b * ns + s) * n + i], acc);
            for (int a = 0; a < nc; ++a) {
                for (int q = e.dptr[a]; q < e.dptr[a + 1]; ++q) {
                    if (e.de[q] != 2 || e.dc[q] != b) continue;
                    const zd coef = e.dcoef[3 * q + 2];
                    if (sqrt(zabs2(coef)) <= 1e-30) continue;
                    zd part = zmake(0.0, 0.0);
                    for (int s = 0; s < ns; ++s) part = zcfma(JP[s * kc + kk], gd[((int64_t)a * ns + s) * n + i], part);
                    acc = zcfma(coef, part, acc);
                }
            }
        }
        u[nn * b + (int64_t)n * k + i] = acc;
    }
}

extern "C" int hex_flux_adjoint(hex_engine *e, const double *g0, const double *gd, double *u, void *stream) {
    if (!e || !g0 || !gd || !u) {
        hex::set_error("flux_adjoint: null argument");
        return HEX_EINVAL;
    }
    if (!e->surfJ) {
        hex::set_error("flux_adjoint: no surface prepared (call prepare_surface first)");
        return HEX_EINVAL;
    }
    hipLaunchKernelGGL(surface_flux_adjoint_kernel, dim3(e->v.n, e->v.nc), dim3(128), 0,
                       static_cast<hipStream_t>(stream), e->v, e->ns, e->k0, e->kc, e->surfJ, e->surfP,
                       reinterpret_cast<const zd *>(g0), reinterpret_cast<const zd *>(gd), reinterpret_cast<zd *>(u));
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// Apply entry points
// ---------------------------------------------------------------------------
// dynamic LDS of apply_direct: the multipoles of every point of the block
static size_t rad_bytes(const EngineView &v, int by) { return sizeof(zd) * (size_t)v.lam_count * kApplyBX * by; }

template <bool V, bool ADJ>
static void launch_by(int by, dim3 grid, hipStream_t stream, const EngineView &v, const double *field, zd ca, zd cb,
                      int affine, int terms, BatchArgs bt, const zd *psi, zd *out) {
    const double f0 = field ? field[0] : 0.0, f1 = field ? field[1] : 0.0, f2 = field ? field[2] : 0.0;
    switch (by) {
    case 1:
        hipLaunchKernelGGL((apply_direct<V, ADJ, 1>), grid, dim3(kApplyBX, 1), rad_bytes(v, 1), stream, v, f0, f1,
                           f2, ca, cb,
                           affine, terms, bt, psi, out);
        break;
    case 2:
        hipLaunchKernelGGL((apply_direct<V, ADJ, 2>), grid, dim3(kApplyBX, 2), rad_bytes(v, 2), stream, v, f0, f1,
                           f2, ca, cb,
                           affine, terms, bt, psi, out);
        break;
    case 8:
        hipLaunchKernelGGL((apply_direct<V, ADJ, 8>), grid, dim3(kApplyBX, 8), rad_bytes(v, 8), stream, v, f0, f1,
                           f2, ca, cb,
                           affine, terms, bt, psi, out);
        break;
    default:
        hipLaunchKernelGGL((apply_direct<V, ADJ, 4>), grid, dim3(kApplyBX, 4), rad_bytes(v, 4), stream, v, f0, f1,
                           f2, ca, cb,
                           affine, terms, bt, psi, out);
    }
}

// General launcher: `members` states `stride` apart, optional per-member
// fields/mask, j rows per block `by` in {1, 2, 4, 8}.
int hex_launch_apply(const EngineView &v, int by, const double *field, int velocity, int adjoint, zd ca, zd cb,
                     BatchArgs bt, int members, const zd *psi, zd *out, hipStream_t stream, int terms) {
    if (!v.diag_full && !v.rad_full && v.y1 && members == 1)
        return hex_launch_apply2(v, by, field, velocity, adjoint, ca, cb, bt, members, psi, out, v.y1, v.y2, v.dcz,
                                 stream, terms);
    if (by != 1 && by != 2 && by != 4 && by != 8) by = kApplyBY;
    while (by > 1 && rad_bytes(v, by) > 65536) by /= 2;   // high multipole orders: fewer rows per block
    dim3 grid((v.n + kApplyBX - 1) / kApplyBX, (v.n + by - 1) / by, members);
    const int affine = !(ca.x == 1.0 && ca.y == 0.0 && cb.x == 0.0 && cb.y == 0.0);
    if (velocity) {
        if (adjoint)
            launch_by<true, true>(by, grid, stream, v, field, ca, cb, affine, terms, bt, psi, out);
        else
            launch_by<true, false>(by, grid, stream, v, field, ca, cb, affine, terms, bt, psi, out);
    } else {
        if (adjoint)
            launch_by<false, true>(by, grid, stream, v, field, ca, cb, affine, terms, bt, psi, out);
        else
            launch_by<false, false>(by, grid, stream, v, field, ca, cb, affine, terms, bt, psi, out);
    }
    HEX_HIP(hipGetLastError());
    return HEX_OK;
}

static int launch_apply(const hex_engine *e, const double *field, int velocity, zd ca, zd cb, const zd *psi, zd *out,
                        hipStream_t stream, int adjoint = 0, int terms = kTermAll) {
    BatchArgs bt{nullptr, nullptr, 0, 0};
    return hex_launch_apply(e->v, e->apply_by, field, velocity, adjoint, ca, cb, bt, 1, psi, out, stream, terms);
}

extern "C" int hex_apply_terms(hex_engine *e, const double field[3], int velocity, int adjoint, int terms,
                               const double a[2], const double b[2], const double *psi, double *out, void *stream) {
    if (!e || !field || !a || !b || !psi || !out || terms < 0 || terms > kTermAll) {
        hex::set_error("apply_terms: bad argument");
        return HEX_EINVAL;
    }
    if (psi == out) {
        hex::set_error("apply_terms: psi and out must not alias");
        return HEX_EINVAL;
    }
    return launch_apply(e, field, velocity, zmake(a[0], a[1]), zmake(b[0], b[1]), reinterpret_cast<const zd *>(psi),
                        reinterpret_cast<zd *>(out), static_cast<hipStream_t>(stream), adjoint, terms);
}

extern "C" int hex_apply(hex_engine *e, const double field[3], int velocity, double scale, const double *psi,
                         double *out, void *stream) {
    if (!e || !field || !psi || !out) {
        hex::set_error("apply: null argument");
        return HEX_EINVAL;
    }
    if (psi == out) {
        hex::set_error("apply: psi and out must not alias");
        return HEX_EINVAL;
    }
    return launch_apply(e, field, velocity, zmake(scale, 0.0), zmake(0.0, 0.0), reinterpret_cast<const zd *>(psi),
                        reinterpret_cast<zd *>(out), static_cast<hipStream_t>(stream));
}

extern "C" int hex_apply_affine(hex_engine *e, const double field[3], int velocity, const double a[2],
                                const double b[2], const double *psi, double *out, void *stream) {
    if (!e || !field || !a || !b || !psi || !out) {
        hex::set_error("apply_affine: null argument");
        return HEX_EINVAL;
    }
    if (psi == out) {
        hex::set_error("apply_affine: psi and out must not alias");
        return HEX_EINVAL;
    }
    return launch_apply(e, field, velocity, zmake(a[0], a[1]), zmake(b[0], b[1]),
                        reinterpret_cast<const zd *>(psi), reinterpret_cast<zd *>(out),
                        static_cast<hipStream_t>(stream));
}

extern "C" int hex_apply_mixed(hex_engine *e, const double coeff[12], const double a[2], const double b[2],
                               const double *psi, double *out, void *stream) {
    if (!e || !coeff || !a || !b || !psi || !out) {
        hex::set_error("apply_mixed: null argument");
        return HEX_EINVAL;
    }
    if (psi == out) {
        hex::set_error("apply_mixed: psi and out must not alias");
        return HEX_EINVAL;
    }
    if (!e->v.mixed || !e->v.y1) {
        hex::set_error("apply_mixed: the engine has no mixed-gauge tables (hex_engine_create_mixed + hex_engine_set_mixed)");
        return HEX_EINVAL;
    }
    HEX_HIP(hipSetDevice(e->device));
    return hex_launch_apply_mixed(e->v, e->apply_by, coeff, zmake(a[0], a[1]), zmake(b[0], b[1]),
                                  reinterpret_cast<const zd *>(psi), reinterpret_cast<zd *>(out),
                                  static_cast<hipStream_t>(stream));
}

extern "C" int hex_apply_batched(hex_engine *e, int members, const double *fields, int velocity, const double *psi,
                                 double *out, void *stream) {
    if (!e || members < 1 || !fields || !psi || !out) {
        hex::set_error("apply_batched: bad argument");
        return HEX_EINVAL;
    }
    if (psi == out) {
        hex::set_error("apply_batched: psi and out must not alias");
        return HEX_EINVAL;
    }
    if (e->v.diag_full || e->v.rad_full) {
        hex::set_error("apply_batched: needs an engine built from separable tables");
        return HEX_EINVAL;
    }
    const int64_t N = (int64_t)e->v.n * e->v.n * e->v.nc;
    const int64_t nd = std::max(e->v.nd, 1);
    if (members > e->bscratch_members) {
        HEX_HIP(hipSetDevice(e->device));
        // release the old buffer only after the work queued on it has finished
        HEX_HIP(hipDeviceSynchronize());
        if (e->bscratch) {
            for (auto it = e->owned.begin(); it != e->owned.end(); ++it)
                if (*it == e->bscratch) {
                    e->owned.erase(it);
                    break;
                }
            HEX_HIP(hipFree(e->bscratch));
            e->bytes -= (int64_t)sizeof(zd) * e->bscratch_members * (2 * N + nd);
            e->bscratch = nullptr;
            e->bscratch_members = 0;
        }
        void *p = nullptr;
        HEX_HIP(hipMalloc(&p, sizeof(zd) * (size_t)members * (2 * N + nd)));
        e->owned.push_back(p);
        e->bytes += (int64_t)sizeof(zd) * members * (2 * N + nd);
        e->bscratch = static_cast<zd *>(p);
        e->bscratch_members = members;
    }
    zd *y1 = e->bscratch, *y2 = y1 + (int64_t)members * N, *dcz = y2 + (int64_t)members * N;
    BatchArgs bt{fields, nullptr, N, N};
    return hex_launch_apply2(e->v, e->apply_by, nullptr, velocity, 0, zmake(1.0, 0.0), zmake(0.0, 0.0), bt, members,
                             reinterpret_cast<const zd *>(psi), reinterpret_cast<zd *>(out), y1, y2, dcz,
                             static_cast<hipStream_t>(stream), kTermAll);
}

extern "C" int hex_apply_adjoint(hex_engine *e, const double field[3], int velocity, const double *psi, double *out,
                                 void *stream) {
    if (!e || !field || !psi || !out) {
        hex::set_error("apply_adjoint: null argument");
        return HEX_EINVAL;
    }
    if (psi == out) {
        hex::set_error("apply_adjoint: psi and out must not alias");
        return HEX_EINVAL;
    }
    return launch_apply(e, field, velocity, zmake(1.0, 0.0), zmake(0.0, 0.0), reinterpret_cast<const zd *>(psi),
                        reinterpret_cast<zd *>(out), static_cast<hipStream_t>(stream), 1);
}

extern "C" int hex_engine_set_launch(hex_engine *e, int rows_per_block) {
    if (!e || (rows_per_block != 0 && rows_per_block != 1 && rows_per_block != 2 && rows_per_block != 4 &&
               rows_per_block != 8 && rows_per_block != kLaunchMfma)) {
        hex::set_error("set_launch: launch shape must be 0 (staged), 1, 2, 4, 8 rows per block or 16 (matrix cores)");
        return HEX_EINVAL;
    }
    e->apply_by = rows_per_block;
    return HEX_OK;
}

// ---------------------------------------------------------------------------
// Drop-in tensor_apply (host buffers, stateless like the reference kernel).
// The caller hands full diag(n,n,nc) and rad(n,n,nlam) tables, so the
// streamed-table variant of the kernel reads them instead of regenerating.
// ---------------------------------------------------------------------------
extern "C" int tensor_apply(int n, int nc, int nt, const int32_t *ptr, const int32_t *col, const double *tv,
                            const double *diag, int nv, const int32_t *vptr, const int32_t *vc,
                            const int32_t *vl, const double *vcoef, const double *rad, int nd,
                            const int32_t *dptr, const int32_t *dc, const int32_t *de, const double *dcoef,
                            const double *ldiff, const double *r, const double *field, int velocity,
                            const double *psi, double *out) {
    if (n < 1 || nc < 1 || nt < 0 || nv < 0 || nd < 0 || !ptr || !diag || !r || !field || !psi || !out ||
        !vptr || !dptr || (nv && !rad)) {
        hex::set_error("tensor_apply: bad argument");
        return HEX_EINVAL;
    }
    int lam_count = 1;
    for (int k = 0; k < nv; ++k) lam_count = vl[k] + 1 > lam_count ? vl[k] + 1 : lam_count;
    std::vector<double> zeros_n(2 * (size_t)n * nc, 0.0), ones(n, 1.0);
    hex_tables t{};
    t.n = n;
    t.nc = nc;
    t.lam_count = lam_count;
    t.nt = nt;
    t.ptr = ptr;
    t.col = col;
    t.tv = tv;
    t.r = r;
    t.cut = ones.data();
    t.d1 = zeros_n.data();
    t.d2 = zeros_n.data();
    t.nv = nv;
    t.vptr = vptr;
    t.vc = vc;
    t.vl = vl;
    t.vcoef = vcoef;
    t.nd = nd;
    t.dptr = dptr;
    t.dc = dc;
    t.de = de;
    t.dcoef = dcoef;
    t.ldiff = ldiff;
    int dev = 0;
    HEX_HIP(hipGetDevice(&dev));
    hex_engine *e = nullptr;
    int rc = hex_engine_create(&t, dev, &e);
    if (rc != HEX_OK) return rc;
    const int64_t N = (int64_t)n * n * nc;
    const int64_t nrad = (int64_t)n * n * lam_count;
    rc = upload(e, reinterpret_cast<const zd *>(diag), (size_t)N, &e->v.diag_full);
    if (rc == HEX_OK && nv) rc = upload(e, reinterpret_cast<const zd *>(rad), (size_t)nrad, &e->v.rad_full);
    const zd *dpsi = nullptr;
    if (rc == HEX_OK) rc = upload(e, reinterpret_cast<const zd *>(psi), (size_t)N, &dpsi);
    zd *dout = nullptr;
    if (rc == HEX_OK) {
        if (hipMalloc(&dout, sizeof(zd) * N) != hipSuccess) {
            hex::set_error("tensor_apply: device allocation failed");
            rc = HEX_EDEVICE;
        } else {
            e->owned.push_back(dout);
        }
    }
    if (rc == HEX_OK) rc = launch_apply(e, field, velocity, zmake(1.0, 0.0), zmake(0.0, 0.0), dpsi, dout, nullptr);
    if (rc == HEX_OK && hipMemcpy(out, dout, sizeof(zd) * N, hipMemcpyDeviceToHost) != hipSuccess) {
        hex::set_error("tensor_apply: copy back failed");
        rc = HEX_EDEVICE;
    }
    hex_engine_destroy(e);
    return rc;
}
