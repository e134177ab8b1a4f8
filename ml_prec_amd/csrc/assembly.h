// Host assembly of the seven-point two-point-flux matrices of the SPE10 grids.
#pragma once

#include <cstdint>

namespace mlp {

enum class SpeKind : int {
    SPE = 0,    // modules/mesh/spe.cc:69-131: face value 2/(1/k_a + 1/k_b) / h^2, diagonal c + sum of faces
    TPFA = 1,   // face value 2/(1/k_a + 1/k_b) * (area / spacing), plus Dirichlet half-cells on x = 0 and x = max
};

// Entries of the matrix of an nx*ny*nz grid: one diagonal per cell and two per interior face.
uint64_t spe_nnz(uint32_t nx, uint32_t ny, uint32_t nz);

// Skyline CSR (diagonal first, then columns ascending) of the grid with cell
// index i + nx*(j + ny*k), permeabilities kx = ky (kx[cell]) and kz =
// kz_ratio * kx, cell sizes hx, hy, hz.  ia has n+1 entries, ja and a spe_nnz.
void spe_assemble(SpeKind kind, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx, double kz_ratio, double c,
                  double hx, double hy, double hz, uint32_t* ia, uint32_t* ja, double* a);

}  // namespace mlp
