// geometry.hpp -- host-side set-up of the lattice: the reference-layout mask
// and boundary lists, and the resolved boundary table that lets the device
// kernels replace PassF + the five copy phases with one gather.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace imexlb {

// Lattice velocities in the reference's numbering
// (initialization.f90:112-123; initialization_3d.f90:108-122).
struct LatticeTable {
    int dim = 2, nq = 9;
    int e[27][3] = {};
    void build(int d);
};

// Mask and lists exactly as the reference's SetGeometry leaves them, in its
// own flat layout (3-wide ghost shell, ids 0-based).
struct RefGeometry {
    int dim = 2, nq = 9;
    int n[3] = {0, 0, 0};  // largest node index per axis
    int L[3] = {1, 1, 1};  // nodes per axis
    int64_t A = 0, dy = 0, dz = 0;
    LatticeTable lat;
    std::vector<int32_t> geo;                      // A cells
    std::vector<int32_t> fluid_id, bl, br, bu, bd, bf, bb, b_user;
    int64_t scan_oob = 0;  // body-scan reads past the array that mattered

    void init(int d, int nx, int ny, int nz);
    // body: the reference's circle/sphere of `radius` at (pos, ny/2[, nz/2]);
    // in 2D optionally an ellipse of the same area, axis ratio `aspect`
    // (semi-axes radius*sqrt(aspect) along, radius/sqrt(aspect) across its
    // axis), the axis turned by `angle_deg` from x (aspect 1, angle 0 is the
    // reference's test itself)
    void build(double pos, double radius, double aspect = 1.0, double angle_deg = 0.0);

    int64_t id(int x, int y, int z) const {
        return int64_t(x + 3) + dy * int64_t(y + 3) + (dim == 3 ? dz * int64_t(z + 3) : 0);
    }
    void coords(int64_t c, int &x, int &y, int &z) const {
        x = int(c % dy) - 3;
        y = int((c / dy) % int64_t(L[1] + 6)) - 3;
        z = dim == 3 ? int(c / dz) - 3 : 0;
    }
    bool interior(int x, int y, int z) const {
        return x >= 0 && x < L[0] && y >= 0 && y < L[1] && (dim == 2 || (z >= 0 && z < L[2]));
    }
    int64_t eoff(int q, int mult) const {
        int64_t o = int64_t(mult) * lat.e[q][0] + int64_t(mult) * dy * lat.e[q][1];
        if (dim == 3) o += int64_t(mult) * dz * lat.e[q][2];
        return o;
    }
};

// One resolved slot: the reference slot the boundary fill writes and where
// its value comes from.  Encoding of ref_src:
//   >= 0  cell*nq + q of an interior fluid slot, post-collision value of
//         this step;
//   -1    zero (a solid cell the reference never writes);
//   <= -2 the value slot (-ref_src - 2) of an interior body cell held at the
//         end of the previous step's BoundaryCondition: Collision never
//         touches solid cells, so what the User phase wrote there persists
//         into the next step, where a wall phase can read it through a
//         periodic image before the User phase rewrites it.
struct FillEntry {
    int64_t ref_slot;  // cell*nq + q in the reference layout
    int64_t ref_src;
};

inline int64_t encode_prev(int64_t slot) { return -slot - 2; }
inline int64_t decode_prev(int64_t src) { return -src - 2; }

struct BcTable {
    std::vector<FillEntry> entries;
    int64_t order_sensitive = 0;  // reads inside one phase of a cell the same
                                  // phase also writes (order-dependent there)
    bool undefined = false;       // some needed value came from past the array
    std::string error;
};

// Resolve, through PassF and the Up, Down, User, Inlet, Outlet phases of
// BoundaryCondition (lbm.f90:96-155, each phase walked in list order):
//  (a) every slot a fluid node pulls from a ghost cell, and
//  (b) every slot of every interior body (b_user) cell -- the ones fluid
//      nodes pull now and the ones the next step may read as its previous
//      value.
BcTable resolve_boundary(const RefGeometry &g);

}  // namespace imexlb
