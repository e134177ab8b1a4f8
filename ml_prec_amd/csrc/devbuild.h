// uh_cheb construction on a device (devbuild.hip): every level built by the
// data-parallel steps of cheb_steps.h on the given HIP stream, with the same
// arrays as the host builder.
#pragma once

#include <cstdint>
#include <memory>

#include "hierarchy.h"

namespace mlp {

// A matrix already in device memory (Skyline CSR).
struct DevCSRView {
    const uint32_t* ia = nullptr;
    const uint32_t* ja = nullptr;
    const double* a = nullptr;
    uint32_t nnz = 0;
};

// stream: the hipStream_t the construction's kernels and copies run on.  With
// level0's pointers set, level 0 is read there instead of being copied from the
// host matrix (whose nrow must still be the size).
std::unique_ptr<ChebLevelBuilder> device_cheb_builder(void* stream, const DevCSRView& level0 = DevCSRView());

// The seven-point layout's arrays, written by spe_assemble_device when given.
struct SpeStencilOut {
    uint8_t* pres;
    double *d, *lx, *ly, *lz;
};

// spe_assemble (assembly.h) on the device: kx and the outputs ia (n+1), ja, a
// (spe_nnz) in device memory; spe = 1 for the reference's assembly, 0 for TPFA.
// With zfirst/zlast (device arrays of nz entries) the planes form a stack of
// independent systems: plane k's system spans planes zfirst[k] .. zlast[k],
// and the matrix has no z-link between systems (block diagonal, each block the
// matrix of that system alone, bit for bit).
void spe_assemble_device(void* stream, int spe, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx,
                         double kz_ratio, double c, double hx, double hy, double hz, uint32_t* ia, uint32_t* ja,
                         double* a, const SpeStencilOut* st = nullptr, const uint32_t* zfirst = nullptr,
                         const uint32_t* zlast = nullptr);

}  // namespace mlp
