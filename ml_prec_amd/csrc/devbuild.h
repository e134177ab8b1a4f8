// uh_cheb construction on a device (devbuild.hip): every level built by the
// data-parallel steps of cheb_steps.h on the given HIP stream, with the same
// arrays as the host builder.
#pragma once

#include <memory>

#include "hierarchy.h"

namespace mlp {

// stream: the hipStream_t the construction's kernels and copies run on
std::unique_ptr<ChebLevelBuilder> device_cheb_builder(void* stream);

}  // namespace mlp
