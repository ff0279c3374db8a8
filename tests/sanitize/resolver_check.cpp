// resolver_check.cpp -- the host-side geometry and boundary resolver of the
// library (imexlb-1.0_amd/csrc/geometry.cpp) for one lattice, for builds
// with AddressSanitizer / UndefinedBehaviorSanitizer.  Prints the resolved
// table's size, its order-sensitive read count, whether a needed value lay
// past the reference's arrays, and a checksum of the entries, so a caller
// can compare with the library's own (uninstrumented) resolution.
//
//   resolver_check dim nx ny nz pos radius aspect angle
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "geometry.hpp"

int main(int argc, char **argv)
{
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s dim nx ny nz pos radius aspect angle\n", argv[0]);
        return 2;
    }
    imexlb::RefGeometry g;
    g.init(std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    g.build(std::atof(argv[5]), std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8]));
    const imexlb::BcTable t = imexlb::resolve_boundary(g);
    uint64_t h = 1469598103934665603ull;  // FNV-1a over (slot, source) pairs
    for (const imexlb::FillEntry &e : t.entries)
        for (int64_t v : {e.ref_slot, e.ref_src}) {
            h ^= uint64_t(v);
            h *= 1099511628211ull;
        }
    std::printf("entries %zu order_sensitive %lld undefined %d fluid %zu user %zu hash %llu error '%s'\n",
                t.entries.size(), (long long)t.order_sensitive, int(t.undefined), g.fluid_id.size(),
                g.b_user.size(), (unsigned long long)h, t.error.c_str());
    return 0;
}
