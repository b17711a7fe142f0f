// pdlp_noalloc.h — the per-device scratch region of pdlp_solve_noalloc
// (pdlp_noalloc.hip).
#pragma once
#include <stddef.h>

#include <string>

namespace pdlp {

// Size of the region (128 MiB per device): room for an FP64 LP of about
// 1.5 million nonzeros.
constexpr size_t kScratchBytes = (size_t)128 << 20;

// Takes the region's device address on `device`, which must be current;
// once per device (later calls return at once).  0, or an error code with
// *err set.
int scratch_setup(int device, std::string *err);

// The region's device address on `device`, or null before scratch_setup.
void *scratch_region(int device);

}  // namespace pdlp
