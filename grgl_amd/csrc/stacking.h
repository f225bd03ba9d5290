// Column layout of requests stacked into one traversal (grgl_hip_matmul_batch, server).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace grgl_amd {

// First stacked row of each request; `total` receives the stacked row count.  Requests of
// at least 16 bytes start on 16-byte boundaries and own whole 16-byte slots; narrower ones
// are packed at the end.
std::vector<uint32_t> stackOffsets(const std::vector<uint32_t>& rows, size_t es, uint64_t& total);

} // namespace grgl_amd
