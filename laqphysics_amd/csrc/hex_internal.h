// Engine internals shared by the plain and the batched entry points.
#pragma once
#include "hex_common.h"

#include <vector>

static constexpr int kLamMax = 16;       // lam_count <= 16 (lmax <= 7)
static constexpr int kApplyBX = 64;      // i per block (one wave)
static constexpr int kApplyBY = 4;       // j per block (one wave each)
static constexpr int kStageMax = 192;    // widest staged i window (columns) of the radial walks
static constexpr int kStageJ = 4;        // j rows per wave in the staged radial walks

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
    // radial-walk scratch of the two-pass apply (members x N each)
    zd *y1, *y2;
    // staged radial walks: per 64-row i block the column window [iblo, ibhi)
    // its rows touch; per row j the dense band lo[j] .. lo[j] + band - 1 of T
    // and of T R - R T (zero-padded to `band` entries)
    int band, stage_width;
    const int32_t *iblo, *ibhi, *jlo, *jhi;
    const zd *tband, *bband;   // [n][band]
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
    int apply_by = kApplyBY;                 // j rows per apply block (tunable)
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

// Two-pass apply (radial walks, then per-group channel coupling).
int hex_launch_apply2(const EngineView &v, int by, const double *field, int velocity, int adjoint, zd ca, zd cb,
                      BatchArgs bt, int members, const zd *psi, zd *out, zd *y1, zd *y2, hipStream_t st, int terms);

// Launch the fused apply for `members` states in/out_stride entries apart.
int hex_launch_apply(const EngineView &v, int by, const double *field, int velocity, int adjoint, zd ca, zd cb,
                     BatchArgs bt, int members, const zd *psi, zd *out, hipStream_t stream, int terms = kTermAll);
