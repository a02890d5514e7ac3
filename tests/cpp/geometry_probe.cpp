// geometry_probe.cpp -- prints the block sizes and padded sizes the drop-in
// header's allocGEMMContext chooses (contextGeometry, the same steps
// allocGEMMContext_base runs) for each "M K N" triple on stdin, one line
// "M K N lhsBlock rhsBlock lhs.nrows_a lhs.ncols_a rhs.nrows_a rhs.ncols_a"
// per triple.  Host only: no device call, so it runs without a GPU.  The
// pytest wrapper compares the lines with the reference's own allocator
// (gemmbitserial.hpp:415-452 + arch-generic.hpp:78-100, compiled by
// oracle/Makefile).
#include <cstdio>

#include "gemmbitserial.hpp"

int main() {
  unsigned long long M, K, N;
  while (std::scanf("%llu %llu %llu", &M, &K, &N) == 3) {
    const gemmbitserial::ContextGeometry g = gemmbitserial::contextGeometry(M, K, N);
    std::printf("%llu %llu %llu %llu %llu %llu %llu %llu %llu\n", M, K, N,
                (unsigned long long)g.lhsBlock, (unsigned long long)g.rhsBlock,
                (unsigned long long)g.lhs_nrows_a, (unsigned long long)g.lhs_ncols_a,
                (unsigned long long)g.rhs_nrows_a, (unsigned long long)g.rhs_ncols_a);
  }
  return 0;
}
