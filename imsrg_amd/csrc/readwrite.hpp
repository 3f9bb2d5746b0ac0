// readwrite.hpp -- interaction files into the two-body layout of an operator.
#pragma once
#include <string>
#include <vector>

#include "basis.hpp"

namespace imsrg {

// Reads a Darmstadt me2j two-body file (plain text, .gz or .bin) into `two` (the
// two-body layout of hb, tbsize doubles, zero on entry) as the reference's
// ReadBareTBME_Darmstadt does for a Hermitian operator; emax < 0, E2max <= 0 and
// lmax <= 0 take the defaults (the basis's emax, 2 emax, emax).  Returns "" or an
// error message; *nrecords: the (a b c d J) records read.
std::string read_me2j(const HostBasis &hb, const std::string &path, int emax, int E2max, int lmax,
                      std::vector<double> &two, long *nrecords);

}  // namespace imsrg
