// Layout of one batch of (ξ, m) round-trip-matrix blocks in HBM.
//
// A batch is a set of blocks (of one or several caps objects).  The host
// fills one BlockDesc per block; every kernel takes the descriptor array and
// the arena base pointer and finds its block through blockIdx (the batch
// dimension).
// All per-block tables live in one device arena, at the offsets recorded in
// the descriptor (in bytes, 256-byte aligned).
#pragma once
#include <stdint.h>

namespace capsd {

struct BlockDesc {
    // physics of the block
    double xi_;          // ξ(L+R)/c
    double alpha;        // 2 ξ(L+R)/c
    double epsm1_plate;  // ε(iξ)-1 of the plate (+inf: perfect reflector)
    double epsm1_sphere; // ε(iξ)-1 of the sphere
    double LbyR;
    int m, lmin, ldim, N;   // N = 2 ldim, window l in [lmin, lmin+ldim)
    int is_pr;              // plate is a perfect reflector: I_TE = -|I_TM|
    // K_ν table: ν in [mu, nu_max]
    int mu, nu_max, nK;
    // I table: l in [lo, lo+nl)
    int lo, nl;
    // quadrature grid of the K integrals: v_k = v0 + k h, k < n_nodes,
    // x(v) = 1 + A² softplus(v)²
    double v0, h, A;
    int n_nodes, n_waves;
    // matrix
    int lda;
    int pad_;
    // byte offsets into the arena
    int64_t off_mie;    // 4 * ldim doubles: log|a_l|, log|b_l|, then the mantissas
                        // of e^{ea(l)}, e^{eb(l)}; then 2 * ldim int exponents;
                        // then nl + (2 nl - 1) doubles: parts of log a0 of I
    int64_t off_kpart;  // n_waves * nK * KPart
    int64_t off_k;      // 2 pols * nK * KEnt (K = m * 2^e), then 2 pols * nK
                        // doubles log|K|
    int64_t off_i;      // I = mant * 2^exp: 2 pols * nl*nl signed double mantissas,
                        // then 2 pols * nl*nl int exponents
    int64_t off_mat;    // lda * N doubles (column-major)
    int64_t off_diag;   // N doubles: diagonal of M
    int64_t off_piv;    // N ints: LU pivots
    int64_t off_cpart;  // n_rtiles x N doubles: per row tile, the column sums
                        // of |M_ij| over i != j (written with the matrix)
    int n_rtiles;       // row tiles of off_cpart (0: none)
    int pad2_;
};

// Blocks whose I tables share the coefficients a_q: same m, ℓ-window and
// polarizations (the blocks of one m at several frequencies).
constexpr int IG_MAX = 8;
struct IGroup {
    int count;
    int blk[IG_MAX];
};

// Partial K sums of one wave of quadrature nodes for one ν.
struct KPart {
    double s_te, s_tm;
    int e;
    int pad_;
};

// One K_ν value as signed mantissa times 2^e, read by the I recurrence with
// a single 16-byte load.
struct KEnt {
    double m;
    int e;
    int pad_;
};
// zeroed KEnt entries in front of every block's K table (BlockDesc::off_k
// is at least this far into the arena; 256 bytes are reserved)
constexpr int K_GUARD = 16;

// Per-block scalar results.
struct BlockOut {
    double trace;       // tr M (Kahan)
    double logdet;      // log|det(1-M)| from the LU (NAN if skipped)
    double frob2;       // ||M||_F^2 (LU mode only)
    double tr2;         // tr M^2 (LU mode only)
    int skip;           // LU skipped (trace shortcut)
    int status;         // 0 ok; nonzero: numerical failure flag
    int delta;          // 1: the matrix holds -M (1 - M without its unit
                        // diagonal; 1 - M is column diagonally dominant, so
                        // partial pivoting exchanges no rows) and the LU
                        // runs without pivot search on u = 1 + d, summing
                        // log1p(d); 0: the matrix holds 1 - M (or a user
                        // matrix), LU with partial pivoting
    int pad_;
};

}  // namespace capsd
