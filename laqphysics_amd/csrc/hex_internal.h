// Engine internals shared by the plain and the batched entry points.
#pragma once
#include "hex_common.h"

#include <vector>

static constexpr int kLamMax = 32;       // lam_count <= 32 (lmax <= 15)
static constexpr int kApplyBX = 64;      // i per block (one wave)
static constexpr int kApplyBY = 4;       // j per block (one wave each)
static constexpr int kLaunchMfma = 16;   // launch shape id: radial walks on the matrix cores
static constexpr int kStageMax = 192;    // widest staged i window (columns) of the radial walks
static constexpr int kStageJ = 4;        // j rows per wave in the staged radial walks
static constexpr int kAngWaves = 4;      // waves sharing one staged channel group in the coupling pass
static constexpr int kZSrc = 4;          // field-coupling sources per channel and electron mixed in the walks

// One multipole edge of the channel coupling, packed for one scalar load:
// coefficient, multipole order and the source channel's slot in its group.
struct AngEdge {
    double cre, cim;
    int32_t lam, src;
    int32_t pad[2];
};

// Read-only tables are read through the constant address space: loads with a
// wave-uniform address then become scalar (SMEM) loads, which the compiler
// does not do for global pointers it cannot prove unclobbered.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *ro(const T *p) {
    return (const __attribute__((address_space(4))) T *)p;
}
__device__ __forceinline__ zd ldz(const __attribute__((address_space(4))) zd *p) { return zmake(p->x, p->y); }

// Workgroups are handed to the 8 XCDs round-robin (block b runs on XCD b % 8),
// each with its own L2.  Renumbering b so that every XCD receives one
// contiguous range of logical ids keeps neighbouring logical blocks -- the ones
// that re-read each other's rows -- in the same L2.  A bijection on [0, G).
__device__ __forceinline__ int64_t xcd_contiguous(int64_t b, int64_t G) {
    constexpr int64_t X = 8;
    const int64_t q = G / X, r = G % X, x = b % X;
    return x * q + (x < r ? x : r) + b / X;
}

struct EngineView {
    int n, nc, lam_count;
    const int32_t *ptr, *col;
    const zd *tv, *tb;        // T and (T R - R T) values on the CSR pattern
    const zd *r;
    const double *cut;
    const zd *d1, *d2;
    const int32_t *vptr, *vc, *vl;
    const zd *vcoef;
    const int32_t *dptr, *dc, *de;
    const zd *dcoef;
    const double *ldiff;
    // Optional full tables (drop-in path): diag(n,n,nc) and rad(n,n,lam_count).
    const zd *diag_full;
    const zd *rad_full;
    // Coulomb groups: channels linked by any multipole edge, listed group by
    // group in gch[gptr[g] .. gptr[g+1]); vcl[k] is the position of edge k's
    // source channel inside its group; gmax = largest group.
    int ngroups, gmax, vreal;   // vreal: every multipole coefficient is real
    const int32_t *gptr, *gch, *vcl;
    const AngEdge *aedge;       // [nv], same order as vc/vl/vcoef
    // radial-walk scratch of the two-pass apply (members x N each)
    zd *y1, *y2;
    // field-coupling coefficients sum_a field_a dcoef(a, k) of the current
    // apply, one per dipole edge (members x nd; zero where the reference
    // skips the edge, |coef| < 1e-30)
    int nd;
    zd *dcz;
    // field-coupling sources by output channel and electron: the dipole edge
    // indices zsrc[(2 c + e - 1) kZSrc + q] (-1 past the channel's count).
    // zsrc_ok: no channel has more than kZSrc sources per electron (the
    // matrix-core pass then walks the mixed sources instead of writing Y1/Y2)
    int zsrc_ok;
    const int32_t *zsrc;
    // staged radial walks: per 64-row i block the column window [iblo, ibhi)
    // its rows touch; per row j the dense band lo[j] .. lo[j] + band - 1 of T
    // and of T R - R T (zero-padded to `band` entries)
    int band, stage_width;
    const int32_t *iblo, *ibhi, *jlo, *jhi;
    const zd *tband, *bband;     // [n][band]
    const zd *tbandT, *bbandT;   // [band][n]
    const double *tbandTr, *bbandTr;   // real parts of the same, [band][n]
    const int32_t *ibw;          // per 64-row i block: widest band of its rows
    const int32_t *ireal;        // per 64-row i block: every band entry of its rows is real
    const int32_t *rowreal;      // per row: every band entry (T and T R - R T) is real
    // j walk per group of kStageJ consecutive rows g: columns [jg_lo[g], jg_lo[g] + jg_w[g])
    // cover the group's bands; coefficients [g][kk < jg_max][q] zero-padded (T, T R - R T),
    // complex and real-part copies; jg_real[g]: the group's coefficients are real
    int jg_max;
    const int32_t *jg_lo, *jg_w, *jg_real;
    const zd *jg_t, *jg_b;
    const double *jg_tr, *jg_br;
    // matrix-core radial walks: rows in blocks of 16 (nb16 blocks).  Block b
    // covers columns kst[b] .. kst[b] + 4 kns[b] - 1 in kns[b] steps of four
    // (kns a multiple of four, the steps past the band holding zero tiles);
    // its coefficient tiles start at tile koff[b], one tile per step, each
    // 64 doubles with lane l holding the entry (16 b + l % 16, kst[b] + 4 s +
    // l / 16) -- the A (and, read as its transpose, B) operand layout of
    // v_mfma_f64_16x16x4f64 -- zero outside the band and past n.  Real and
    // imaginary parts of T (mt, mti) and of T R - R T (mb, mbi); kreal[b]: the
    // block's coefficients are all real.
    int nb16;
    const int32_t *kst, *kns, *koff, *kreal;
    const double *mt, *mti, *mb, *mbi;
    // Mixed gauge (set by hex_engine_set_mixed on an engine whose field-coupling
    // walks were built from q = r (1 - f)): per radial point hq = q / (2 r^2),
    // F = r f, c0 = (1 - f)^2, c2 = q'^2 - (1 - f)^2; the angular-square edges
    // (per output channel sptr[c] .. sptr[c+1]: source ssrc, electron sele,
    // six coefficients scoef[6 k ..]); per-apply coefficients dczE (E part of
    // the dipole edges, nd) and dczQ (square edges, nsq).
    int mixed, nsq;
    const zd *hq, *mF, *c2;
    const double *c0;
    const int32_t *sptr, *ssrc, *sele;
    const zd *scoef;
    zd *dczE, *dczQ;
};

struct hex_engine {
    int device = 0;
    EngineView v{};
    int nt = 0, nv = 0, nd = 0;
    std::vector<void *> owned;
    int64_t bytes = 0;
    // reduction scratch
    zd *partials = nullptr;   // [kRedBlocks][kRedCols]
    zd *zbuf = nullptr;       // [kRedCols] device coefficients
    zd *hbuf_host = nullptr;  // pinned host results
    // tSURFF surface rows: J = [T, Theta] and JP = i [T, Theta] o (r_k - r_s),
    // restricted to the column window [k0, k0 + kc) that holds every nonzero.
    int ns = 0, k0 = 0, kc = 0;
    zd *surfJ = nullptr, *surfP = nullptr;   // [ns][kc] row-major
    int apply_by = 0;                        // radial launch: 0 staged, 1/2/4/8 direct j rows per block (tunable)
    // scratch of the B-member apply (grown on demand): y1, y2 (B x N each), coefficients (B x nd)
    zd *bscratch = nullptr;
    int64_t bscratch_members = 0;
};


struct BatchArgs {
    const double *fields;    // [B][3] per-member A, or null (use f0..f2)
    const int32_t *active;   // [B] member mask, or null
    int64_t in_stride;       // entries between members of the input
    int64_t out_stride;      // entries between members of the output
};

// Term selection of the fused apply (bit mask).
static constexpr int kTermBase = 1;     // diagonal, kinetic, Coulomb
static constexpr int kTermDipole = 2;   // field-coupling edges
static constexpr int kTermA2 = 4;       // |A|^2 (velocity gauge)
static constexpr int kTermAll = 7;

// Two-pass apply (radial walks, then per-group channel coupling); dcz holds
// members x nd field-coupling coefficients (scratch).
int hex_launch_apply2(const EngineView &v, int by, const double *field, int velocity, int adjoint, zd ca, zd cb,
                      BatchArgs bt, int members, const zd *psi, zd *out, zd *y1, zd *y2, zd *dcz, hipStream_t st,
                      int terms);

// Mixed-gauge apply (engine built by hex_engine_create_mixed and given its
// tables by hex_engine_set_mixed); coeff = [A, E, A_a A_b] (12).
int hex_launch_apply_mixed(const EngineView &v, int by, const double *coeff, zd ca, zd cb, const zd *psi, zd *out,
                           hipStream_t st);

// Launch the fused apply for `members` states in/out_stride entries apart.
int hex_launch_apply(const EngineView &v, int by, const double *field, int velocity, int adjoint, zd ca, zd cb,
                     BatchArgs bt, int members, const zd *psi, zd *out, hipStream_t stream, int terms = kTermAll);
