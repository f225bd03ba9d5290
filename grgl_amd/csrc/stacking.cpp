// Column layout of requests stacked into one traversal.
#include "stacking.h"

#include <algorithm>
#include <stdexcept>

namespace grgl_amd {

std::vector<uint32_t> stackOffsets(const std::vector<uint32_t>& rows, size_t es, uint64_t& total) {
    // Largest requests first, each on a 16-byte boundary with its slot rounded up to whole
    // 16-byte vectors; requests narrower than 16 bytes go last, packed, and use scalar
    // accesses.  The K ladder 1..256 (f32) stacks into exactly 511 rows this way.
    const uint64_t v = std::max<size_t>(1, 16 / es);
    std::vector<uint32_t> idx(rows.size());
    for (uint32_t i = 0; i < idx.size(); i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return rows[a] > rows[b]; });
    std::vector<uint32_t> off(rows.size(), 0);
    uint64_t k = 0;
    for (uint32_t i : idx) {
        const bool wide = (uint64_t)rows[i] * es >= 16;
        if (wide) k = (k + v - 1) / v * v;
        if (k > 0xffffffffULL) throw std::runtime_error("too many stacked rows");
        off[i] = (uint32_t)k;
        k += wide ? (rows[i] + v - 1) / v * v : rows[i];
    }
    // The last wide slot may have been rounded up past its rows; trim to the used rows.
    uint64_t used = 0;
    for (uint32_t i = 0; i < rows.size(); i++) used = std::max<uint64_t>(used, (uint64_t)off[i] + rows[i]);
    total = used;
    return off;
}

} // namespace grgl_amd
