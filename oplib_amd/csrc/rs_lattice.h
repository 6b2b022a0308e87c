// rs_lattice.h -- FP64 regime-switching local-vol lattice, element by
// element (LV is nr = 1).  Shared by the assembly kernels (lattice.hip)
// and the W form's fused assembly (wsq.hip).
//
// States y = x + nx*r (S2DGrid.cs:89).  Row y: rates to x-1 and x+1 in the
// same regime from the closed-form inverse of the 2x2 stencil of
// S1DGrid.cs:78-147 applied to (drift, vol^2) as in CLVModel.cs:99-114;
// rate q_off to every other regime at the same x; diagonal = -sum.  Rows at
// x = 0 and x = nx-1 have no price moves (absorbing in x).
#pragma once
#include "ophip_internal.h"

namespace ophip {

struct RSLattice {
    int nx, nr;
    const double *xval;   // nx nodes
    const double *drift;  // per state (nx*nr), in price units
    const double *vol;    // per state (nx*nr), absolute vol in price units
    double q_off;
};

__device__ inline void rs_rates(const RSLattice &L, int y, double &am, double &ap) {
    int x = y % L.nx;
    am = 0.0;
    ap = 0.0;
    if (x <= 0 || x >= L.nx - 1) return;
    double m00 = L.xval[x - 1] - L.xval[x];
    double m01 = L.xval[x + 1] - L.xval[x];
    double m10 = m00 * m00, m11 = m01 * m01;
    double det = m00 * m11 - m01 * m10;
    double i00 = m11 / det, i10 = -m10 / det, i01 = -m01 / det, i11 = m00 / det;
    double mu = L.drift[y], v = L.vol[y];
    double s2 = v * v;
    am = i00 * mu + i01 * s2;
    ap = i10 * mu + i11 * s2;
}

// MODE 0: generator (FP64 out); MODE 1: I + dt*L; MODE 2: dt*L (U0 - I).
// NM: node-major states y = nr*x + r (the same operator in another basis:
// the 8 regimes of a node are neighbours, so the matrix is one band and
// the W form's zero-cell skipping applies); drift/vol stay indexed by the
// canonical state x + nx*r.
template <int MODE, bool NM>
__device__ inline double rs_elem(const RSLattice &L, double dt, int i, int j) {
    const int nx = L.nx, nr = L.nr;
    const int xi = NM ? i / nr : i % nx, ri = NM ? i % nr : i / nx;
    const int xj = NM ? j / nr : j % nx, rj = NM ? j % nr : j / nx;
    double g = 0.0;
    if (ri == rj) {
        if (xj == xi - 1 || xj == xi + 1 || xj == xi) {
            double am, ap;
            rs_rates(L, xi + nx * ri, am, ap);
            if (xj == xi - 1) g = am;
            else if (xj == xi + 1) g = ap;
            else g = -(am + ap + (double)(L.nr - 1) * L.q_off);
        }
    } else if (xi == xj) {
        g = L.q_off;
    }
    if (MODE == 0) return g;
    if (MODE == 1) return (i == j ? 1.0 : 0.0) + dt * g;
    return dt * g;
}

// false only if every generator entry (i, j) with i in [i0, i0+n) and j in
// [j0, j0+n) is structurally zero.  An entry can be nonzero only where
// |x_i - x_j| <= 1: node-major, x = y / nr, so the node ranges of the rows
// and columns must overlap after widening by one; canonical, it needs
// j - i in {-1, 0, 1} (same regime) or j - i a multiple of nx (same node),
// i.e. some multiple of nx within [lo - 1, hi + 1] of the differences j - i.
template <bool NM>
__host__ __device__ inline bool rs_block_may_be_nonzero(int nx, int nr, int i0, int j0, int n) {
    if (NM) {
        const int xr0 = i0 / nr, xr1 = (i0 + n - 1) / nr;
        const int xc0 = j0 / nr, xc1 = (j0 + n - 1) / nr;
        return xc0 <= xr1 + 1 && xr0 <= xc1 + 1;
    }
    const int lo = j0 - i0 - (n - 1) - 1, hi = j0 - i0 + (n - 1) + 1;
    // smallest multiple of nx >= lo (floor division for negative lo)
    const int q = lo >= 0 ? (lo + nx - 1) / nx : -((-lo) / nx);
    return (long)q * nx <= hi;
}

// lattice arguments of the assembly entry points (handles resolved)
inline bool rs_args(int nx, int nr, unsigned xval, unsigned drift, unsigned vol, RSLattice &L) {
    if (nx < 3 || nr < 1 || (long)nx * nr > 65535) return false;
    L.nx = nx;
    L.nr = nr;
    L.xval = P<const double>(xval);
    L.drift = P<const double>(drift);
    L.vol = P<const double>(vol);
    return L.xval && L.drift && L.vol;
}

}  // namespace ophip
