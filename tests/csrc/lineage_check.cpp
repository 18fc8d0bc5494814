// Checks the lineage partition of a segment's paths (pcr_kernels.h): for every
// count tried, the paths lineage_path(j, r), r < lineage_count(count, j), of
// the kLineages lineages cover [0, count) exactly once, and lineage_base is
// the running sum of the counts.  Prints "ok" or the first failure.
#include <cstdio>
#include <vector>

#include "pcr_kernels.h"

int main() {
    using namespace pcr;
    std::vector<unsigned char> seen;
    for (uint32_t count = 0; count < 5000; count += (count < 1100 ? 1 : 37)) {
        seen.assign(count, 0);
        uint32_t total = 0;
        for (int j = 0; j < kLineages; j++) {
            if (lineage_base(count, j) != total) {
                std::printf("count %u lineage %d: base %u, expected %u\n", count, j, lineage_base(count, j), total);
                return 1;
            }
            const uint32_t n = lineage_count(count, j);
            for (uint32_t r = 0; r < n; r++) {
                const uint32_t p = lineage_path(j, r);
                if (p >= count || seen[p]) {
                    std::printf("count %u lineage %d item %u: path %u %s\n", count, j, r, p,
                                p >= count ? "out of range" : "repeated");
                    return 1;
                }
                seen[p] = 1;
            }
            total += n;
        }
        if (total != count) {
            std::printf("count %u: lineages hold %u paths\n", count, total);
            return 1;
        }
    }
    // the largest segments: counts only (a path index fits 32 bits)
    for (uint32_t count : {1u << 24, (1u << 26) + 12345u, 4000000000u}) {
        uint64_t total = 0;
        for (int j = 0; j < kLineages; j++) total += lineage_count(count, j);
        if (total != count || lineage_base(count, kLineages - 1) + lineage_count(count, kLineages - 1) != count) {
            std::printf("count %u: lineages hold %llu paths\n", count, (unsigned long long)total);
            return 1;
        }
    }
    std::printf("ok\n");
    return 0;
}
