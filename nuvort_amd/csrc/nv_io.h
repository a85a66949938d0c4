// Host-side formats of the reference's output files (nv_io.cpp).
#pragma once
#include <string>

#include "../../include/nuvort.h"

namespace nvio {
// print_info's ts.log header and one line (src/output.f90:23-37)
std::string ts_header();
std::string ts_line(int shots, const nv_diag &d, double delta);
// one tree_mesh line of tree_print (src/tree.f90:96-102)
std::string tree_leaf_line(const nv_tree_node &nd);
}  // namespace nvio
