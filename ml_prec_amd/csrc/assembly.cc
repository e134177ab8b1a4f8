// Seven-point matrix assembly on the host, rows filled in parallel.
#include "assembly.h"

#include "mlp_error.h"
#include "par.h"

namespace mlp {

uint64_t spe_nnz(uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint64_t X = nx, Y = ny, Z = nz;
    const uint64_t faces = (X ? X - 1 : 0) * Y * Z + X * (Y ? Y - 1 : 0) * Z + X * Y * (Z ? Z - 1 : 0);
    return X * Y * Z + 2 * faces;
}

// spe.cc:69-131 (construct_matrix, XYZ cell order).  Row of cell (i, j, k):
// the diagonal c first, then the neighbours z-, y-, x-, x+, y+, z+ that exist
// (ascending columns), each with -v, and v added to the diagonal in that
// order; v = 2/(1/k_this + 1/k_neighbour) / h^2 on the face's axis (the ADD
// macro, spe.cc:80-86).  TPFA scales the harmonic mean by area/spacing
// instead and starts the diagonal at c + the Dirichlet half-cell terms
// k * (hy*hz) / (hx/2) of the cells on x = 0 and x = nx-1.
void spe_assemble(SpeKind kind, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx, double kz_ratio, double c,
                  double hx, double hy, double hz, uint32_t* ia, uint32_t* ja, double* a) {
    if (!nx || !ny || !nz) throw Error("grid dimensions must be positive", 1);
    const uint64_t n = uint64_t(nx) * ny * nz;
    if (n >= (uint64_t(1) << 32) || spe_nnz(nx, ny, nz) >= (uint64_t(1) << 32))
        throw Error("grid too large for 32-bit indices", 1);
    const uint32_t sy = nx, sz = nx * ny;
    // row starts: 1 + the neighbours present
    ia[0] = 0;
    for (uint32_t k = 0, r = 0; k < nz; k++)
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t i = 0; i < nx; i++, r++) {
                const uint32_t cnt = 1 + (k > 0) + (j > 0) + (i > 0) + (i + 1 < nx) + (j + 1 < ny) + (k + 1 < nz);
                ia[r + 1] = ia[r] + cnt;
            }
    const bool spe = kind == SpeKind::SPE;
    const double sxx = hx * hx, syy = hy * hy, szz = hz * hz;                       // spe: divide by h^2
    const double fx = hy * hz / hx, fy = hx * hz / hy, fz = hx * hy / hz;           // tpfa: multiply
    auto face = [&](double ka, double kb, int axis) {
        const double m = 2 / (1 / ka + 1 / kb);
        if (spe) return axis == 0 ? m / sxx : axis == 1 ? m / syy : m / szz;
        return axis == 0 ? m * fx : axis == 1 ? m * fy : m * fz;
    };
    parallel_for(nz, 1, [&](uint64_t k0, uint64_t k1) {
        for (uint32_t k = uint32_t(k0); k < k1; k++)
            for (uint32_t j = 0; j < ny; j++)
                for (uint32_t i = 0; i < nx; i++) {
                    const uint32_t r = i + sy * j + sz * k;
                    uint32_t p = ia[r];
                    const uint32_t d = p;
                    double diag = c;
                    if (!spe) {
                        double tb = 0.0;
                        if (i == 0) tb += kx[r] * (hy * hz) / (hx / 2);
                        if (i + 1 == nx) {
                            if (nx > 1) tb += kx[r] * (hy * hz) / (hx / 2);
                            else tb *= 2.0;
                        }
                        diag = diag + tb;
                    }
                    ja[p] = r;
                    p++;
                    auto add = [&](uint32_t q, double v) {
                        ja[p] = q;
                        a[p] = -v;
                        p++;
                        diag = diag + v;
                    };
                    const double kzr = kz_ratio * kx[r];
                    if (k > 0) add(r - sz, face(kzr, kz_ratio * kx[r - sz], 2));
                    if (j > 0) add(r - sy, face(kx[r], kx[r - sy], 1));
                    if (i > 0) add(r - 1, face(kx[r], kx[r - 1], 0));
                    if (i + 1 < nx) add(r + 1, face(kx[r], kx[r + 1], 0));
                    if (j + 1 < ny) add(r + sy, face(kx[r], kx[r + sy], 1));
                    if (k + 1 < nz) add(r + sz, face(kzr, kz_ratio * kx[r + sz], 2));
                    a[d] = diag;
                }
    });
}

}  // namespace mlp
