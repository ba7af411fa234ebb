// Octree (host): bit-exact restatement of libtree's scaling, Morton universal
// index, level selection and interaction domains.
#pragma once
#include <stdint.h>
#include <vector>

namespace ifs {
namespace tree {

constexpr int kThresholdLevel = 21;  // 8-byte universal index, 3 dimensions (lib_tree.f90:184-189)

struct Scaling {
    double x_min[3];
    double D;  // common edge (lib_tree.f90:320-374)
};

// Scales points to [0,1)^3 exactly as lib_tree_get_scaled_element_list
// (libtree/src/lib/lib_tree.f90:320-374).
Scaling scale_points(const double *pts, int n, std::vector<double> &scaled);
// floor(f * 2^64) from the IEEE-754 fields of f in [0,1)
// (libtree/src/lib/lib_tree_helper_functions.f90:532-614).
uint64_t coordinate_binary(double f);
// Morton index of the box at level l that contains p (x is the most significant
// bit of every triple) (lib_tree_helper_functions.f90:202-364).
int64_t universal_index(const double *p, int l);
inline int64_t parent(int64_t n, int step = 1) { return n >> (3 * step); }
// children in ascending order (lib_tree_helper_functions.f90:741-755)
void children(int64_t n, int64_t out[8]);
// box centre in scaled coordinates (lib_tree_helper_functions.f90:774-799)
void centre(int64_t n, int l, double out[3]);
// 26 neighbours at distance k in the reference's order, -1 where outside
// (lib_tree_helper_functions.f90:837-997)
void neighbours(int64_t k, int64_t n, int l, int64_t out[26]);
// I2 = own box + existing neighbours (lib_tree.f90:898-948)
std::vector<int64_t> domain_i2(int64_t n, int l, int64_t k = 1);
// I4 = children(I2(parent)) \ I2(self), ascending (lib_tree.f90:994-1143)
std::vector<int64_t> domain_i4(int64_t n, int l, int64_t k = 1);
// l_min = 1 + floor(log2(k+1)) (lib_tree.f90:1162-1166)
int level_min(int64_t k);
// smallest level at which no box holds more than s of the sorted indices
// (lib_tree.f90:1188-1242); sorted_uindex at kThresholdLevel
int level_max(const std::vector<int64_t> &sorted_uindex, int64_t s);

}  // namespace tree
}  // namespace ifs
